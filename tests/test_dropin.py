"""The reference's import lines resolve to the drop-in after chimera_amd.dropin.install()
(no device work: imports only)."""
import subprocess
import sys


def test_reference_imports_resolve():
    code = (
        "import chimera_amd.dropin as d; r = d.install();"
        "from exafs_gpu.forward import ForwardModel, thermal_sigma2;"
        "from exafs_gpu.fourier import FourierTransform, r_grid, hanning_window;"
        "from chimera.multift import MultiEdgeFT;"
        "from methods.method_a_rmc import RMCSampler;"
        "import chimera_amd.forward as f, chimera_amd.method_a_rmc as m, chimera_amd.multift as mm;"
        "assert ForwardModel is f.ForwardModel and RMCSampler is m.RMCSampler and MultiEdgeFT is mm.MultiEdgeFT;"
        "import exafs_gpu, methods; assert exafs_gpu.forward is f and methods.method_a_rmc is m;"
        "assert sorted(r['installed']) == sorted(d.ALIASES) and not r['skipped'];"
        "print('ok')"
    )
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().endswith("ok")


def test_already_imported_name_is_reported():
    code = (
        "import sys, types; sys.modules['methods.method_a_rmc'] = types.ModuleType('other');"
        "import chimera_amd.dropin as d; r = d.install();"
        "assert 'methods.method_a_rmc' in r['skipped'] and 'exafs_gpu.forward' in r['installed'];"
        "print('ok')"
    )
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
