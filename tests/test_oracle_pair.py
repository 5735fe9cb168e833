"""Pins the pair-space sampler oracle (oracle/pair_rmc.py): every move's Delta chi against
the full recompute, the energy change against the energy difference, the displacement
step's clamp and distribution, the |chi(R)| chi^2 against its minimum and its zero, and
the tracked state of a long run against a from-scratch recount."""
import math

import numpy as np
import pytest

from oracle import fourier as OT
from oracle import lattice as OL
from oracle import pair_rmc as PR
from oracle import rmc as OR
from oracle import tables as OTab


def _problem(nsp=2, cumulants=False, seed=99):
    cell = OL.Supercell("fcc", (3, 3, 3), 3.59)
    i, j, sh = cell.neighbour_list(5.2)
    shell, radii = OL.shell_index(cell.ideal, i, j, sh)
    zs = [24, 26, 28][:nsp]
    k, f, d, lam, s02 = OTab.synthetic_ss(OTab.k_grid(), zs)
    R = np.arange(1.4, 5.0001, 0.02)
    trs = [OT.Transform(k, 3.0, 14.0, w, 1.0) for e in range(nsp) for w in (1, 2)]
    ce = np.repeat(np.arange(nsp), 2)
    sp_t = OL.random_config(108, [1] * nsp, seed)
    rng = np.random.default_rng(seed)
    sig2 = 0.004 + 0.002 * rng.random(i.size)
    nshell = int(shell.max()) + 1
    kw = {}
    if cumulants:
        kw = dict(dr=rng.normal(0, 0.01, (nshell, nsp, nsp)), c3=rng.normal(0, 1e-4, nshell),
                  s2types=rng.normal(0, 5e-4, (nshell, nsp, nsp)))
    P = OR.Problem(k, f, d, lam, s02, cell.ideal, i, j, sh, shell, sig2, trs, ce,
                   np.zeros((len(trs), R.size), complex), np.full((len(trs), R.size), 0.05),
                   R=R, **kw)
    pos_t = cell.ideal + rng.normal(0, 0.04, cell.ideal.shape)
    S = PR.PairSampler(P)
    S.cnt = np.bincount(sp_t, minlength=nsp).astype(float)
    z = P.to_R(S.chi_full(sp_t, pos_t) / S.cnt[:, None])
    P.tm = z + 0.02 * (rng.normal(size=z.shape) + 1j * rng.normal(size=z.shape))
    return P, cell


@pytest.mark.parametrize("cumulants", [False, True])
def test_delta_chi_equals_full_recompute(cumulants):
    """Swaps and displacements (off the lattice, with shell cumulants): the tracked chi
    after 40 moves equals chi of the final configuration recomputed pair by pair."""
    P, cell = _problem(cumulants=cumulants)
    S = PR.PairSampler(P, nrep=1, swap_frac=0.5, disp_sigma=0.05, max_disp=0.3, seed=4)
    sp0 = OL.random_config(108, [1, 1], 3)
    S.init_state(sp0, cell.ideal)
    nswap = ndisp = 0
    for t in range(1, 41):
        a, b, na, nb, sw, newp = PR.propose_move(S.species[0], S.pos[0], P.ideal, 4, t, 0, 0.5,
                                                 0.05, 0.3)
        dchi, aff, r_new, t_new = S.delta_chi(0, a, b, na, nb, newp)
        one = S.chi_full(S.species[0], S.pos[0])
        sp, pos = S.species[0].copy(), S.pos[0].copy()
        sp[a], sp[b] = na, nb
        if not sw:
            pos[a] = newp
        two = S.chi_full(sp, pos)
        np.testing.assert_allclose(dchi, two - one, rtol=0, atol=1e-12 * np.abs(one).max())
        S.r[0, aff], S.ptype[0, aff] = r_new, t_new
        S.species[0], S.pos[0] = sp, pos
        S.chi[0] += dchi
        nswap += sw
        ndisp += 1 - sw
    assert nswap > 5 and ndisp > 5
    np.testing.assert_allclose(S.r[0], PR.distances(S.pos[0], P.i, P.j, P.shift), rtol=1e-15)
    np.testing.assert_array_equal(S.ptype[0], S.species[0][P.i] * 2 + S.species[0][P.j])


def test_planted_doubled_pair_is_caught():
    """A Delta chi that counts the pair (a, b) twice (no dedup of b's block) differs from the
    recompute: the check above would catch it."""
    P, cell = _problem()
    S = PR.PairSampler(P, nrep=1)
    sp = OL.random_config(108, [1, 1], 3)
    S.init_state(sp, cell.ideal)
    a = 0
    b = int(next(j for j in P.j[P.i == a] if sp[j] != sp[a]))   # a neighbour of a
    good, aff, r_new, t_new = S.delta_chi(0, a, b, int(sp[b]), int(sp[a]), cell.ideal[a])
    dup = np.r_[aff, np.nonzero((P.i == b) & (P.j == a))[0]]
    spn = sp.copy()
    spn[a], spn[b] = sp[b], sp[a]
    bad = (S.terms(dup, S.r[0, dup], spn[P.i[dup]] * 2 + spn[P.j[dup]])
           - S.terms(dup, S.r[0, dup], S.ptype[0, dup]))
    want = S.chi_full(spn, cell.ideal) - S.chi_full(sp, cell.ideal)
    tol = 1e-12 * np.abs(want).max()
    assert np.abs(good - want).max() < tol
    assert np.abs(bad - want).max() > 1e3 * tol


def test_energy_delta_equals_energy_difference():
    P, cell = _problem(nsp=3)
    rng = np.random.default_rng(1)
    V = rng.normal(0, 0.05, (2, 3, 3))
    V = 0.5 * (V + V.transpose(0, 2, 1))
    S = PR.PairSampler(P, nrep=1, energy_prior=(V, 800.0))
    sp = OL.random_config(108, [1, 1, 1], 8)
    S.init_state(sp, cell.ideal)
    assert S.ebeta == pytest.approx(1.0 / (PR.KB_EV * 800.0))
    for a, b in [(0, 1), (5, 77), (10, 10)]:
        na, nb = (int(sp[b]), int(sp[a])) if a != b else (int(sp[a]),) * 2
        aff, r_new, t_new = S.move_pairs(0, a, b, na, nb, cell.ideal[a])
        spn = sp.copy()
        spn[a], spn[b] = na, nb
        np.testing.assert_allclose(S.energy_delta(0, aff, t_new), S.energy(spn) - S.energy(sp),
                                   rtol=0, atol=1e-13)
    # the table holds V/2 per directed pair for the shells V gives, 0 beyond
    tab = S.etab.reshape(-1, 9)
    np.testing.assert_array_equal(tab[:2], 0.5 * V.reshape(2, 9))
    assert np.all(tab[2:] == 0)


def test_displacement_clamp_and_distribution():
    """Displacements: Gaussian steps of disp_sigma around the current position, the offset
    from the ideal site clamped to max_disp (kernels.py:405-421); a configuration with
    one species never swaps."""
    cell = OL.Supercell("fcc", (2, 2, 2), 3.59)
    ideal = cell.ideal
    sp = np.zeros(32, np.int64)
    pos = ideal.copy()
    steps = []
    for t in range(1, 3001):
        a, b, na, nb, sw, newp = PR.propose_move(sp, pos, ideal, 7, t, 0, 0.75, 0.02, 10.0)
        assert sw == 0 and a == b and na == nb == 0
        steps.append(newp - pos[a])
    g = np.array(steps) / 0.02
    assert abs(g.mean()) < 0.06 and abs(g.std() - 1.0) < 0.04
    # clamped: the offset from the ideal site never exceeds max_disp, and a step past it
    # lands on the sphere along the unclamped direction
    pos2 = ideal.copy()
    pos2[:, 0] += 0.25
    for t in range(1, 200):
        a, _, _, _, _, newp = PR.propose_move(sp, pos2, ideal, 9, t, 0, 0.0, 0.1, 0.3)
        off = newp - ideal[a]
        assert np.linalg.norm(off) <= 0.3 * (1 + 1e-12)
    # swap fraction: with two species the swap share is swap_frac times the chance of
    # finding a partner in 8 draws
    sp2 = OL.random_config(32, [1, 1], 2)
    nsw = sum(PR.propose_move(sp2, ideal, ideal, 3, t, 0, 0.75, 0.02, 0.3)[4] for t in range(1, 4001))
    assert abs(nsw / 4000 - 0.75 * (1 - 0.5 ** 8)) < 0.025


def test_magnitude_chi2_minimum_and_zero():
    """|chi(R)| target: the profiled scale minimises chi^2 per edge, and a target that is
    |chi(R)| of the configuration itself gives chi^2 = 0."""
    P, cell = _problem()
    S = PR.PairSampler(P, nrep=1, magnitude=True)
    sp = OL.random_config(108, [1, 1], 3)
    S.init_state(sp, cell.ideal)
    m = S.observe(S.chi[0])
    assert np.isrealobj(m) and S.nfit == P.tm.shape[1] * P.nchan
    s = S.profiled_scale(m)

    def c2(scale):
        return (((m * scale[:, None] - S.tm) / S.sm) ** 2).sum() / S.nfit

    assert S.chi2_of_chi(S.chi[0]) == pytest.approx(c2(s), rel=1e-12)
    for e in range(2):
        for eps in (1e-3, -1e-3):
            s2 = s.copy()
            s2[2 * e:2 * e + 2] *= 1 + eps
            assert c2(s2) > c2(s)
    T = PR.PairSampler(P, nrep=1, magnitude=True, target_abs=0.7 * m)
    T.init_state(sp, cell.ideal)
    assert T.chi2[0] < 1e-24


@pytest.mark.parametrize("magnitude,prior", [(False, False), (True, True)])
def test_run_keeps_state_consistent(magnitude, prior):
    """After a run with exchanges every replica's tracked chi, r and types equal a recount
    from its species and positions, chi^2 equals chi^2 of that chi, the best retained
    state reproduces its chi^2, and both move kinds were accepted."""
    P, cell = _problem()
    V = np.array([[[0.0, -0.02], [-0.02, 0.0]]])
    S = PR.PairSampler(P, nrep=4, seed=2, swap_frac=0.6, disp_sigma=0.03, magnitude=magnitude,
                       energy_prior=(V, 600.0) if prior else None)
    S.init_state(OL.random_config(108, [1, 1], 5), cell.ideal)
    S.run(60, exchange_every=10)
    for c in range(4):
        np.testing.assert_allclose(S.chi[c], S.chi_full(S.species[c], S.pos[c]), rtol=0,
                                   atol=1e-11 * np.abs(S.chi[c]).max())
        np.testing.assert_allclose(S.r[c], PR.distances(S.pos[c], P.i, P.j, P.shift), rtol=1e-15)
        np.testing.assert_array_equal(S.ptype[c], S.species[c][P.i] * 2 + S.species[c][P.j])
        assert S.chi2[c] == pytest.approx(S.chi2_of_chi(S.chi[c]), rel=1e-10)
        bc = S.chi2_of_chi(S.chi_full(S.best_sp[c], S.best_pos[c]))
        assert S.best_chi2[c] == pytest.approx(bc, rel=1e-10)
    kinds = {(r[2], r[3]) for step in S.log for r in step}
    assert (1, 1) in kinds and (0, 1) in kinds
    assert np.abs(S.pos - cell.ideal[None]).max() > 0.0
