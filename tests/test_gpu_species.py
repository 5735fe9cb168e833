"""The forward-model kernels at every species count they are built for (2..6): FP64 and
FP32 path sums, dL/dpos, the relaxed-occupancy path sum and its adjoint, through the C ABI
against the FP64 oracle, on a small fcc cell (two shells) with per-pair sigma^2 and
cumulant tables; also a k grid without a k = 0 point (every lane live), a non-uniform
one, the workloads' default grid and a fine grid of 801 points.  dL/dpos is checked with C3 != 0 (per-term sine and cosine) and with C3 = 0, where a
uniform grid whose tables fit takes the k-rotation recurrence."""
import sys

import numpy as np
import pytest

from oracle import forward as OF

pytestmark = pytest.mark.gpu

ZS = (24.0, 27.0, 28.0, 26.0, 25.0, 13.0)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _model(nsp, k):
    from chimera_amd.forward import ForwardModel
    from chimera_amd.lattice import CrystalSupercell
    from chimera_amd.scattering import synthetic
    cell = CrystalSupercell("fcc", (3, 3, 3), 3.59)
    i, j, sh = cell.neighbour_list(3.6)
    shell, _ = cell.shells(i, j, sh)
    tab = synthetic(ZS[:nsp], k)
    rng = np.random.default_rng(nsp)
    sig2 = 0.004 + 0.003 * rng.random(i.size)
    fm = ForwardModel(cell, tab, i, j, sh, sigma2=sig2, shell=shell)
    return fm, cell, i, j, sh, shell, tab, sig2


@pytest.mark.parametrize("nsp,kind", [(2, "grid"), (4, "grid"), (5, "grid"), (6, "grid"),
                                      (2, "no_zero"), (4, "no_zero"), (3, "nonuniform"),
                                      (6, "default"), (3, "fine"), (6, "fine")])
def test_forward_kernels_all_species_counts(nsp, kind):
    """kind "default": the workloads' 0..16 grid at 0.05 (321 points; at six species the
    relaxed-occupancy tables need two k chunks); "fine": 0..20 at 0.025 (801 points: more
    k points than one workgroup has lanes, so every kernel works in k chunks)."""
    from chimera_amd.scattering import k_grid
    if kind == "grid":
        k = k_grid(16.0, 0.1)
    elif kind == "default":
        k = k_grid()
    elif kind == "fine":
        k = k_grid(20.0, 0.025)
    elif kind == "no_zero":
        k = np.linspace(0.5, 18.0, 320)
    else:
        k = np.linspace(0.5, 18.0, 320)
        k = k + 0.02 * np.sin(k)            # increasing, not uniform
    _say(f"[test] forward kernels nsp={nsp} k={kind} ({k.size} points)")
    fm, cell, i, j, sh, shell, tab, sig2 = _model(nsp, k)
    rng = np.random.default_rng(10 + nsp)
    ns = int(shell.max()) + 1
    dr = rng.normal(0, 0.01, (ns, nsp, nsp))
    c3 = rng.normal(0, 2e-4, ns)
    s2t = np.abs(rng.normal(0, 5e-4, (ns, nsp, nsp)))
    pos = cell.ideal[None] + rng.normal(0, 0.04, (2,) + cell.ideal.shape)
    sp = rng.integers(0, nsp, (2, cell.natoms)).astype(np.int32)
    oargs = (tab.k, tab.f, tab.delta, tab.lam, tab.s02)

    # hard occupancy with the cumulant tables
    fm.set_dr(dr).set_c3(c3).set_sig2_types(s2t)
    got = fm.chi(pos, sp, normalise=False)
    got32 = fm.chi(pos, sp, normalise=False, precision="f32")
    g = rng.normal(0, 1, (2, nsp, tab.k.size))
    gp = fm.grad_pos(pos, sp, g, normalise=False)
    for c in range(2):
        ref = OF.chi_pairs(*oargs, pos[c], sp[c], i, j, sh, sig2, shell=shell, dr=dr, c3=c3,
                           s2types=s2t, normalise=False)
        scale = np.abs(ref).max()
        assert np.abs(got[c] - ref).max() <= 1e-12 * scale
        assert np.abs(got32[c] - ref).max() <= 1e-5 * scale
        rg = OF.grad_pos(*oargs, pos[c], sp[c], i, j, sh, sig2, g[c], shell=shell, dr=dr, c3=c3,
                         s2types=s2t)
        assert np.abs(gp[c] - rg).max() <= 1e-11 * np.abs(rg).max()
    fm.set_c3(np.zeros_like(c3))
    gp0 = fm.grad_pos(pos, sp, g, normalise=False)
    for c in range(2):
        rg = OF.grad_pos(*oargs, pos[c], sp[c], i, j, sh, sig2, g[c], shell=shell, dr=dr,
                         s2types=s2t)
        assert np.abs(gp0[c] - rg).max() <= 1e-11 * np.abs(rg).max()
    fm.set_dr(np.zeros_like(dr)).set_sig2_types(np.zeros_like(s2t))

    # relaxed occupancy
    w = rng.dirichlet(np.ones(nsp), (2, cell.natoms))
    soft = fm.chi_soft(pos, w)
    gsp, gsw = fm.grad_soft(pos, w, g)
    for c in range(2):
        ref = OF.chi_soft(*oargs, pos[c], w[c], i, j, sh, sig2)
        assert np.abs(soft[c] - ref).max() <= 1e-12 * np.abs(ref).max()
        rp, rw = OF.grad_soft(*oargs, pos[c], w[c], i, j, sh, sig2, g[c])
        assert np.abs(gsp[c] - rp).max() <= 1e-11 * np.abs(rp).max()
        assert np.abs(gsw[c] - rw).max() <= 1e-11 * np.abs(rw).max()
