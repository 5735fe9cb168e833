"""Adjoints of the path sum (grad_pos), the relaxed-occupancy path sum (chi_soft) and its
adjoint (grad_soft) on the device, through the C ABI, against the FP64 oracle (whose
gradients are pinned by finite differences in tests/test_oracle.py)."""
import sys

import numpy as np
import pytest

from oracle import forward as OF

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tern():
    from chimera_amd.forward import ForwardModel
    from chimera_amd.lattice import CrystalSupercell
    from chimera_amd.scattering import k_grid, synthetic
    _say("[build] fcc 3x3x3 CrCoNi, rmax 5.2")
    cell = CrystalSupercell("fcc", (3, 3, 3), 3.59)
    i, j, sh = cell.neighbour_list(5.2)
    shell, radii = cell.shells(i, j, sh)
    tab = synthetic(("Cr", "Co", "Ni"), k_grid())
    rng = np.random.default_rng(7)
    sig2 = 0.004 + 0.003 * rng.random(i.size)
    fm = ForwardModel(cell, tab, i, j, sh, sigma2=sig2, shell=shell)
    return fm, cell, i, j, sh, shell, tab, sig2


def _pos(cell, nconf, seed):
    rng = np.random.default_rng(seed)
    return cell.ideal[None] + rng.normal(0, 0.04, (nconf,) + cell.ideal.shape)


def test_grad_pos_matches_oracle(tern):
    fm, cell, i, j, sh, shell, tab, sig2 = tern
    _say("[test] grad_pos vs oracle (cumulant tables set)")
    rng = np.random.default_rng(1)
    ns = int(shell.max()) + 1
    dr = rng.normal(0, 0.01, (ns, 3, 3))
    c3 = rng.normal(0, 2e-4, ns)
    s2t = np.abs(rng.normal(0, 5e-4, (ns, 3, 3)))
    fm.set_dr(dr).set_c3(c3).set_sig2_types(s2t)
    try:
        pos = _pos(cell, 2, 3)
        sp = rng.integers(0, 3, (2, cell.natoms)).astype(np.int32)
        g = rng.normal(0, 1, (2, 3, tab.k.size))
        got = fm.grad_pos(pos, sp, g, normalise=False)
        for c in range(2):
            ref = OF.grad_pos(tab.k, tab.f, tab.delta, tab.lam, tab.s02, pos[c], sp[c], i, j, sh,
                              sig2, g[c], shell=shell, dr=dr, c3=c3, s2types=s2t)
            assert np.abs(got[c] - ref).max() <= 1e-11 * np.abs(ref).max()
        # normalise: the same as dividing dL/dchi by the species counts
        cnt = np.stack([(sp == s).sum(1) for s in range(3)], 1)
        np.testing.assert_allclose(fm.grad_pos(pos, sp, g),
                                   fm.grad_pos(pos, sp, g / cnt[:, :, None], normalise=False),
                                   rtol=1e-14, atol=0)
        # deterministic: the same bits twice
        assert np.array_equal(got, fm.grad_pos(pos, sp, g, normalise=False))
    finally:
        fm.set_dr(np.zeros((ns, 3, 3))).set_c3(np.zeros(ns)).set_sig2_types(np.zeros((ns, 3, 3)))


def test_chi_soft_matches_oracle_and_one_hot(tern):
    fm, cell, i, j, sh, shell, tab, sig2 = tern
    _say("[test] chi_soft vs oracle")
    rng = np.random.default_rng(2)
    pos = _pos(cell, 2, 4)
    w = rng.dirichlet([1.0, 1.0, 1.0], (2, cell.natoms))
    got = fm.chi_soft(pos, w)
    for c in range(2):
        ref = OF.chi_soft(tab.k, tab.f, tab.delta, tab.lam, tab.s02, pos[c], w[c], i, j, sh, sig2)
        assert np.abs(got[c] - ref).max() <= 1e-12 * np.abs(ref).max()
    sp = rng.integers(0, 3, (2, cell.natoms)).astype(np.int32)
    one = np.eye(3)[sp]
    np.testing.assert_allclose(fm.chi_soft(pos, one), fm.chi(pos, sp), rtol=0,
                               atol=1e-12 * np.abs(fm.chi(pos, sp)).max())


def test_grad_soft_matches_oracle(tern):
    fm, cell, i, j, sh, shell, tab, sig2 = tern
    _say("[test] grad_soft vs oracle")
    rng = np.random.default_rng(3)
    pos = _pos(cell, 2, 5)
    w = rng.dirichlet([1.0, 1.0, 1.0], (2, cell.natoms))
    g = rng.normal(0, 1, (2, 3, tab.k.size))
    gp, gw = fm.grad_soft(pos, w, g)
    for c in range(2):
        rp, rw = OF.grad_soft(tab.k, tab.f, tab.delta, tab.lam, tab.s02, pos[c], w[c], i, j, sh,
                              sig2, g[c])
        assert np.abs(gp[c] - rp).max() <= 1e-11 * np.abs(rp).max()
        assert np.abs(gw[c] - rw).max() <= 1e-11 * np.abs(rw).max()


def test_profile_entry_points(tern):
    fm, cell, i, j, sh, shell, tab, sig2 = tern
    _say("[test] forward-model profile hooks")
    pos = _pos(cell, 2, 6)
    sp = np.zeros((2, cell.natoms), np.int32)
    w = np.full((2, cell.natoms, 3), 1.0 / 3.0)
    g = np.ones((2, 3, tab.k.size))
    fm.chi(pos, sp)
    fm.grad_pos(pos, sp, g)
    fm.chi_soft(pos, w)
    fm.grad_soft(pos, w, g)
    for what in ("chi", "chi_f32", "adjoint", "soft", "soft_adjoint"):
        assert fm.profile(what, 2, iters=3) > 0.0
    with pytest.raises(RuntimeError):
        fm.profile("adjoint", 5, iters=1)       # never run with nconf = 5
