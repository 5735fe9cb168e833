"""A multiple-scattering path library read from FEFF-format files (library.json +
feffNNNN.dat per entry) drives the device table and the sampler; both match the oracle,
which reads the same files with its own parser."""
import os
import sys

import numpy as np
import pytest

from oracle import fourier as OT
from oracle import lattice as OL
from oracle import mspaths as OM
from oracle import rmc as OR
from oracle import tables as OTab

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_file_library_table_and_sampler(tmp_path):
    from chimera_amd.forward import ForwardModel
    from chimera_amd.fourier import FourierTransform
    from chimera_amd.lattice import CrystalSupercell
    from chimera_amd.method_a_rmc import RMCSampler
    from chimera_amd.mspaths import MSLibrary, PathLibrary
    from chimera_amd.multift import MultiEdgeFT
    from chimera_amd.scattering import k_grid, synthetic
    from tests.feffio import write_library
    from tests.test_oracle import _fdl

    a0_lib, a0 = 3.59, 3.61                      # the library at its own lattice, used scaled
    _say("[build] fcc 3x3x3 FeNi with a FEFF-format library")
    cell = CrystalSupercell("fcc", (3, 3, 3), a0)
    i, j, sh = cell.neighbour_list(5.2)
    shell, radii = cell.shells(i, j, sh)
    k = k_grid()
    tab = synthetic(("Fe", "Ni"), k)
    fm = ForwardModel(cell, tab, i, j, sh, sigma2=np.full(i.size, 0.006), shell=shell)
    enum = MSLibrary(CrystalSupercell("fcc", (3, 3, 3), a0_lib), 2, 5.2)
    root = write_library(str(tmp_path / "lib"), ("Fe", "Ni"), enum.classes, a0_lib / 2,
                         a0_lib / np.sqrt(2), _fdl([26, 28]))
    lib = PathLibrary(root).prepare(k)
    ms = MSLibrary(cell, 2, 5.2, sigma2_nn=0.006, r_nn=float(radii[0]), c_ms=1.3, amp=0.9,
                   library=lib)
    assert abs(ms.a_scale - a0 / a0_lib) < 1e-15

    # oracle: its own reading of every file, table at the scaled lattice
    F, D, L, reff = [], [], [], []
    for el in ("Fe", "Ni"):
        for n in range(lib.nblk):
            _, _, rf, cols = OTab.read_feff_columns(os.path.join(root, el, f"feff{n + 1:04d}.dat"))
            p, f, d, lam = OTab.feff_columns_to_path(cols, rf)
            F.append(np.interp(k, p, f))
            D.append(np.interp(k, p, d))
            L.append(np.interp(k, p, lam))
            reff.append(rf)
    T_or = OTab.tabulated_table(k, np.array(F), np.array(D), np.array(L), reff, 0.9, 0.006,
                                float(radii[0]), c_ms=1.3, amp=0.9, a_scale=a0 / a0_lib)
    _say("[test] device table from the files")
    T = ms.table(fm)
    assert np.abs(T - T_or).max() <= 1e-12 * np.abs(T_or).max()

    # the sampler on the file library, move by move against the oracle
    R = np.arange(1.4, 5.0001, 0.02)
    fts = [FourierTransform(k, 3.0, 14.0, w, dk_win=1.0, R=R) for e in range(2) for w in (1, 2, 3)]
    ce = np.repeat(np.arange(2), 3).astype(np.int32)
    mft = MultiEdgeFT(fts, chan_edge=ce)
    ocell = OL.Supercell("fcc", (3, 3, 3), a0)
    ocl = [OM.PathClass(c.rep) for c in lib.classes]
    sc = OM.SupercellPaths(ocl, 2, ocell.ideal, a0, (3, 3, 3))
    ktab, f_, d_, lam_, s02 = OTab.synthetic_ss(k, [26, 28], s02=0.9)
    trs = [OT.Transform(k, 3.0, 14.0, w, 1.0) for e in range(2) for w in (1, 2, 3)]
    prob0 = OR.Problem(k, f_, d_, lam_, s02, ocell.ideal, i, j, sh, shell, 0.006, trs, ce,
                       np.zeros((6, R.size), complex), np.ones((6, R.size)), R=R, ms=sc, T=T_or)
    sp_t = OL.random_config(108, [1, 1], 31)
    chi_t = prob0.chi_full(sp_t) / np.maximum(np.bincount(sp_t, minlength=2), 1)[:, None]
    tgt = OT.multi_to_R(trs, ce, chi_t, R=R)
    sig = np.full(tgt.shape, 0.02 * np.abs(tgt).max())
    smp = RMCSampler(fm, mft, tgt, np.ones(R.size, bool), sig, nrep=6, seed=4, swap_frac=1.0,
                     ms_library=ms)
    nsteps = 40
    smp.set_recording(log_steps=nsteps)
    sp0 = OL.random_config(108, [1, 1], 2)
    smp.init_state(sp0, cell.ideal)
    _say("[run] device sampler on the file library")
    smp.run(nsteps, exchange_every=10, log_every=10 ** 9)
    log = smp.move_log(nsteps)
    prob = OR.Problem(k, f_, d_, lam_, s02, ocell.ideal, i, j, sh, shell, 0.006, trs, ce, tgt,
                      sig, R=R, ms=sc, T=T_or)
    Ro = OR.PlainSampler(prob, nrep=6, seed=4)
    Ro.init_state(sp0)
    _say("[run] oracle sampler")
    Ro.run(nsteps, exchange_every=10)
    for t in range(nsteps):
        by_rung = {int(r[3]): r for r in log[t]}
        for rung in range(6):
            a, b, acc, c2, x, u = Ro.log[t][rung]
            g = by_rung[rung]
            assert (int(g[0]), int(g[1]), int(g[2])) == (a, b, acc), f"step {t} rung {rung}"
    assert np.array_equal(smp.species, Ro.species)
    np.testing.assert_allclose(smp.chi2, Ro.chi2, rtol=1e-9)
    assert smp.accepted_per_rung().sum() > 0
