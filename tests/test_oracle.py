"""Pins the FP64 oracle: published values, closed forms, invariants, brute force."""
import numpy as np
import pytest

from oracle import forward as OF
from oracle import fourier as OT
from oracle import lattice as OL
from oracle import mspaths as OM
from oracle import rmc as OR
from oracle import tables as OTab


def test_splitmix64_published_sequence():
    # Vigna's splitmix64 seeded with 0: first three outputs (public reference values)
    x = 0
    outs = []
    for _ in range(3):
        outs.append(OR.splitmix64(x))
        x = (x + OR.GOLDEN) & OR.M64
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_u01_range_and_resolution():
    assert OR.u01(0) == 0.0
    assert OR.u01(OR.M64) < 1.0
    assert OR.u01(1 << 11) == 2.0 ** -53


@pytest.fixture(scope="module")
def fcc108():
    cell = OL.Supercell("fcc", (3, 3, 3), 3.59)
    i, j, sh = cell.neighbour_list(5.2)
    shell, radii = OL.shell_index(cell.ideal, i, j, sh)
    return cell, i, j, sh, shell, radii


def test_fcc_shells(fcc108):
    cell, i, j, sh, shell, radii = fcc108
    assert i.size == 54 * cell.natoms
    assert np.bincount(shell).tolist() == [12 * 108, 6 * 108, 24 * 108, 12 * 108]
    a0 = 3.59
    np.testing.assert_allclose(radii, [a0 / np.sqrt(2), a0, a0 * np.sqrt(1.5), a0 * np.sqrt(2)],
                               rtol=1e-12)


def test_reverse_index_involution(fcc108):
    cell, i, j, sh, shell, radii = fcc108
    rev = OL.reverse_index(i, j, cell.natoms)
    assert np.array_equal(rev[rev], np.arange(i.size))
    assert np.array_equal(i[rev], j) and np.array_equal(shell[rev], shell)


def test_neighbour_list_brute_force():
    # every ordered pair within rmax found once, against an explicit all-images search
    cell = OL.Supercell("bcc", (3, 3, 3), 2.87)
    i, j, sh = cell.neighbour_list(3.0)
    got = set(zip(i.tolist(), j.tolist()))
    want = set()
    for a in range(cell.natoms):
        for b in range(cell.natoms):
            for img in cell._images:
                d = cell.ideal[b] + img - cell.ideal[a]
                if 1e-6 < np.linalg.norm(d) <= 3.0:
                    want.add((a, b))
    assert got == want
    assert np.bincount(i).tolist() == [14] * cell.natoms   # bcc 8 + 6


def test_warren_cowley_special_cases(fcc108):
    cell, i, j, sh, shell, radii = fcc108
    # L1_0: alternate (001) planes -> first-shell alpha_AB = 1 - (8/12)/0.5 = -1/3
    z = np.rint(cell.ideal[:, 2] / (3.59 / 2)).astype(int)
    sp = (z % 2).astype(np.int32)
    a = OL.warren_cowley(sp, i, shell, j, 0, 2)
    assert abs(a[0, 1] + 1.0 / 3.0) < 1e-12 and abs(a[1, 0] + 1.0 / 3.0) < 1e-12
    # random alloys average alpha -> 0
    vals = [OL.warren_cowley(OL.random_config(108, [1, 1], s), i, shell, j, 0, 2)[0, 1]
            for s in range(40)]
    assert abs(np.mean(vals)) < 0.03


def test_integer_counts_rounding():
    assert OL.integer_counts([1, 1, 1], 4000).tolist() == [1333, 1333, 1334]
    assert OL.integer_counts([0.5, 0.5], 500).tolist() == [250, 250]


def test_ms_classes_reference_figures():
    # the reference's own numbers: 25 fcc classes to 5.3 A, 327 ternary entries per
    # absorber (chimera/mslibrary.py:11,16; README.md:267)
    cl = OM.enumerate_classes("fcc", 3.56, 5.3)
    assert len(cl) == 25
    assert len(OM.entry_list(cl, 3)) == 327
    assert len(OM.entry_list(cl, 2)) == 122
    assert sum(c.deg for c in cl) == 984
    assert sum(c.nleg == 3 for c in cl) + sum(c.nleg == 4 for c in cl) == 25
    # every image is a distinct directed path of the class's length
    for c in cl:
        L = c.reff_units()
        for img, so in c.images:
            pts = np.vstack([np.zeros((1, 3)), img, np.zeros((1, 3))])
            assert abs(0.5 * np.linalg.norm(np.diff(pts, axis=0), axis=1).sum() - L) < 1e-9


def test_ms_counts_and_delta_brute_force(fcc108):
    cell, i, j, sh, shell, radii = fcc108
    cl = OM.enumerate_classes("fcc", 3.59, 5.2)
    sc = OM.SupercellPaths(cl, 3, cell.ideal, 3.59, (3, 3, 3))
    assert sc.npath == 108 * sum(c.deg for c in cl)
    sp = OL.random_config(108, [1, 1, 1], 3)
    C = sc.counts(sp)
    assert C.sum() == sc.npath
    # species marginal: each absorber species appears deg times per atom
    nb = sc.nblk
    for a in range(3):
        assert C[a * nb:(a + 1) * nb].sum() == (sp == a).sum() * sum(c.deg for c in cl)
    rng = np.random.default_rng(0)
    for _ in range(20):
        a, b = rng.integers(0, 108, 2)
        if sp[a] == sp[b]:
            continue
        d = sc.delta_counts(sp, a, b, sp[b], sp[a])
        sp2 = sp.copy()
        sp2[a], sp2[b] = sp[b], sp[a]
        assert np.array_equal(sc.counts(sp2) - C, d)


def test_pair_space_equals_count_space(fcc108):
    cell, i, j, sh, shell, radii = fcc108
    k, f, d, lam, s02 = OTab.synthetic_ss(OTab.k_grid(), [24, 27, 28])
    sp = OL.random_config(108, [1, 1, 1], 1)
    chi = OF.chi_pairs(k, f, d, lam, s02, cell.ideal, sp, i, j, sh, 0.006, normalise=False)
    N = OF.pair_counts(sp, i, j, shell, 3, 4)
    B = OF.ss_basis(k, f, d, lam, s02, radii, np.full(4, 0.006))
    np.testing.assert_allclose(OF.chi_counts(N, B), chi, rtol=0, atol=1e-13 * np.abs(chi).max())
    assert np.all(chi[:, 0] == 0.0)       # the k = 0 term is 0


def test_pair_term_closed_form():
    k = np.array([0.0, 1.0, 4.0])
    f = np.array([0.0, 2.0, 3.0])
    d = np.array([0.5, 0.25, -1.0])
    lam = np.array([6.0, 6.3, 7.2])
    v = OTab.pair_term(k, 2.5, f, d, lam, 0.006, 0.9, c3=0.001)
    want = [0.0,
            0.9 * 2.0 / (1 * 6.25) * np.exp(-5 / 6.3 - 2 * 0.006) * np.sin(5 + 0.25 - 4 / 3 * 0.001),
            0.9 * 3.0 / (4 * 6.25) * np.exp(-5 / 7.2 - 32 * 0.006) * np.sin(20 - 1 - 4 / 3 * 64 * 0.001)]
    np.testing.assert_allclose(v, want, rtol=1e-15)


def test_synthetic_tables_spot_values():
    k, f, d, lam, s02 = OTab.synthetic_ss(OTab.k_grid(), [26, 28])
    assert k.size == 321 and k[80] == 4.0
    assert f[0, 80] == 1.0 and abs(f[1, 80] - 28 / 26) < 1e-15
    assert abs(d[0, 1, 0] - (1.2 + 0.05 * 28)) < 1e-15 and lam[0] == 6.0


def test_fft_equals_direct_dft_and_hanning():
    k = OTab.k_grid()
    t = OT.Transform(k, 3.0, 14.0, 2, 1.0)
    rng = np.random.default_rng(0)
    chi = rng.standard_normal(k.size)
    z_fft = t.to_R(chi)
    z_dft = t.dft(chi, t.R)
    np.testing.assert_allclose(z_dft, z_fft, rtol=0, atol=1e-12 * np.abs(z_fft).max())
    # Hanning: flat top 1, half height at kmin and kmax, 239 non-zero points
    w = OT.hanning_window(t.k_uniform, 3.0, 14.0, 1.0)
    assert w[60] == 0.5 and w[280] == 0.5 and w[100] == 1.0
    assert np.count_nonzero(t.weight) == 239


def test_transform_peak_position():
    k = OTab.k_grid()
    t = OT.Transform(k, 3.0, 14.0, 0, 1.0)
    R = np.arange(1.0, 5.0, 0.005)
    z = t.dft(np.sin(2 * k * 2.7), R)
    assert abs(R[np.argmax(np.abs(z))] - 2.7) < 0.02


def _toy_problem():
    cell = OL.Supercell("fcc", (3, 3, 3), 3.59)
    i, j, sh = cell.neighbour_list(5.2)
    shell, radii = OL.shell_index(cell.ideal, i, j, sh)
    k, f, d, lam, s02 = OTab.synthetic_ss(OTab.k_grid(), [26, 28])
    R = np.arange(1.4, 5.0001, 0.02)
    trs = [OT.Transform(k, 3.0, 14.0, w, 1.0) for e in range(2) for w in (1, 2, 3)]
    ce = np.repeat(np.arange(2), 3)
    sp_t = OL.random_config(108, [1, 1], 99)
    chi_t = OF.chi_pairs(k, f, d, lam, s02, cell.ideal, sp_t, i, j, sh, 0.006)
    tgt = OT.multi_to_R(trs, ce, chi_t, R=R)
    P = OR.Problem(k, f, d, lam, s02, cell.ideal, i, j, sh, shell, 0.006, trs, ce, tgt,
                   np.full(tgt.shape, 0.05), R=R)
    return P


def test_profiled_scale_minimises_chi2():
    P = _toy_problem()
    sp = OL.random_config(108, [1, 1], 5)
    chi = P.chi_full(sp) / np.bincount(sp)[:, None]
    m = P.to_R(chi)
    s = P.profiled_scale(m)
    def c2(scale):
        d = (m * scale[:, None] - P.tm) / P.sm
        return (np.abs(d) ** 2).sum()
    base = c2(s)
    for e in range(2):
        for eps in (1e-3, -1e-3):
            s2 = s.copy()
            s2[3 * e:3 * e + 3] *= 1 + eps
            assert c2(s2) > base


def test_affected_pairs_dedup_brute_force():
    P = _toy_problem()
    for a, b in [(0, 1), (0, 50), (7, 7), (3, 107)]:
        aff = OR.affected_pairs(a, b, P.nbr, P.rev, P.i, P.j)
        want = sorted(np.nonzero((P.i == a) | (P.j == a) | ((P.i == b) | (P.j == b)) * (a != b))[0])
        assert sorted(aff.tolist()) == want
        assert len(set(aff.tolist())) == len(aff)


def test_incremental_chi_matches_recompute():
    P = _toy_problem()
    sp = OL.random_config(108, [1, 1], 7)
    chi = P.chi_full(sp)
    rng = np.random.default_rng(3)
    for _ in range(10):
        a, b = rng.integers(0, 108, 2)
        if sp[a] == sp[b]:
            continue
        d, _ = P.delta_chi(sp, a, b, sp[b], sp[a])
        sp[a], sp[b] = sp[b], sp[a]
        chi = chi + d
    np.testing.assert_allclose(chi, P.chi_full(sp), rtol=0, atol=1e-12 * np.abs(chi).max())


def test_exchange_acceptance_detailed_balance():
    # two rungs, two states: P(swap) pi(x) == P(swap back) pi(x')
    b1, b2, nfit = 2.0, 0.5, 10.0
    for E1, E2 in [(1.0, 1.3), (1.3, 1.0), (2.0, 2.0)]:
        def acc(Ea, Eb):
            d = (b1 - b2) * (Ea - Eb) * nfit
            return 1.0 if d >= 0 else np.exp(max(d, -60))
        pi = np.exp(-nfit * (b1 * E1 + b2 * E2))
        pi_sw = np.exp(-nfit * (b1 * E2 + b2 * E1))
        assert abs(pi * acc(E1, E2) - pi_sw * acc(E2, E1)) <= 1e-12 * max(pi, pi_sw)


def test_sampler_runs_and_improves():
    P = _toy_problem()
    S = OR.PlainSampler(P, nrep=4, seed=1)
    S.init_state(OL.random_config(108, [1, 1], 5))
    c0 = S.chi2.copy()
    S.run(30, exchange_every=10)
    assert S.best_chi2.min() <= c0.min()
    # tracked chi equals a recompute for every replica
    for c in range(4):
        np.testing.assert_allclose(S.chi[c], P.chi_full(S.species[c]), rtol=0,
                                   atol=1e-11 * np.abs(S.chi[c]).max())


def test_adapt_ladder_closed_forms():
    """Ladder retuning (method_a_rmc.py:389-398) pinned by cases with known answers."""
    beta = np.exp(-np.arange(3.0))                      # log gaps -1, -1
    # rates 0.8 / 0.2: sqrt(0.8) : sqrt(0.2) = 2 : 1, span 2 kept -> gaps -4/3, -2/3
    nb = OR.adapt_ladder(beta, [0.8, 0.2])
    np.testing.assert_allclose(np.log(nb), [0.0, -4.0 / 3.0, -2.0], rtol=0, atol=1e-14)
    # equal rates leave a ladder unchanged, whatever its spacing
    b = 1.0 / np.geomspace(0.02, 5.0, 8)
    np.testing.assert_allclose(OR.adapt_ladder(b, np.full(7, 0.37)), b, rtol=1e-13)
    b2 = np.sort(np.random.default_rng(0).uniform(0.1, 40.0, 6))[::-1]
    np.testing.assert_allclose(OR.adapt_ladder(b2, np.full(5, 0.6)), b2, rtol=1e-13)
    # endpoints fixed, order kept, and rates beyond [0.02, 0.98] act as the clip values
    r = np.array([0.0, 0.5, 1.0, 0.3, 0.9, 0.05, 0.7])
    nb = OR.adapt_ladder(b, r)
    assert abs(nb[0] / b[0] - 1) < 1e-14 and abs(nb[-1] / b[-1] - 1) < 1e-13
    assert np.all(np.diff(nb) < 0)
    np.testing.assert_array_equal(nb, OR.adapt_ladder(b, np.clip(r, 0.02, 0.98)))
    # a pair that exchanges more freely than average moves apart (its log gap grows)
    g0, g1 = np.diff(np.log(b)), np.diff(np.log(nb))
    assert abs(g1[2]) > abs(g0[2]) and abs(g1[0]) < abs(g0[0])


def test_exchange_rates_filled_and_adaptive_run():
    """Per-pair sweep outcomes: tried pairs 0/1, untried ones the mean of the tried
    (method_a_rmc.py:366-372); an adaptive run retunes the ladder inside its span."""
    P = _toy_problem()
    S = OR.PlainSampler(P, nrep=5, seed=3)
    S.init_state(OL.random_config(108, [1, 1], 5))
    _, rates = S.replica_exchange(return_rates=True)
    start, _ = OR.exchange_uniforms(S.seed, S.iex, 4)
    tried = np.zeros(4, bool)
    tried[start::2] = True
    assert set(np.unique(rates[tried])) <= {0.0, 1.0}
    np.testing.assert_allclose(rates[~tried], rates[tried].mean())
    b0 = S.beta.copy()
    S.run(40, exchange_every=5, adapt_every=20)
    assert not np.allclose(S.beta, b0) and abs(S.beta[0] / b0[0] - 1) < 1e-14
    assert abs(S.beta[-1] / b0[-1] - 1) < 1e-13 and np.all(np.diff(S.beta) < 0)


def test_ms_expected_counts_random_alloy(fcc108):
    """Random-alloy entry occurrences (chimera/mslibrary.py:315-321): exact for a pure
    crystal (every path of the class, all-species-s entry), and the mean of the exact
    counts over i.i.d. random occupations for a mixture."""
    cell, i, j, sh, shell, radii = fcc108
    cl = OM.enumerate_classes("fcc", 3.59, 5.2)
    sc = OM.SupercellPaths(cl, 2, cell.ideal, 3.59, (3, 3, 3))
    nb = sc.nblk
    for s in range(2):
        comp = np.eye(2)[s]
        E = OM.expected_counts(cl, 2, comp)
        C = sc.counts(np.full(108, s))
        np.testing.assert_allclose(E[s * nb:(s + 1) * nb], C[s * nb:(s + 1) * nb] / 108.0, rtol=0,
                                   atol=1e-12)
    comp = np.array([0.3, 0.7])
    E = OM.expected_counts(cl, 2, comp)
    rng = np.random.default_rng(4)
    acc = np.zeros(2 * nb)
    nab = np.zeros(2)
    for _ in range(300):
        sp = (rng.random(108) < comp[1]).astype(np.int64)
        acc += sc.counts(sp)
        nab += np.bincount(sp, minlength=2)
    mean = np.concatenate([acc[a * nb:(a + 1) * nb] / nab[a] for a in range(2)])
    big = E > 0.5
    np.testing.assert_allclose(mean[big], E[big], rtol=0.08)
    np.testing.assert_allclose(mean.sum(), E.sum(), rtol=1e-12)   # every path counted once


def test_k_exp_axis_closed_form():
    """chimera/model.py:120-130: k_exp^2 = k^2 - ETOK dE0 above the 1e-4 floor."""
    k = np.linspace(0.0, 16.0, 321)
    np.testing.assert_allclose(OTab.k_exp_axis(k, 0.0)[k >= 0.01], k[k >= 0.01], rtol=1e-15)
    for dE0 in (-3.0, 2.5):
        ke = OTab.k_exp_axis(k, dE0)
        ok = k * k - OTab.ETOK * dE0 > 1e-4
        np.testing.assert_allclose(ke[ok] ** 2 + OTab.ETOK * dE0, k[ok] ** 2, rtol=1e-12)
        np.testing.assert_allclose(ke[~ok], 0.01)
        assert np.all(np.diff(ke) >= 0)
    # 10 eV is k^2 = 2.6247 A^-2 (2 m_e 10 eV / hbar^2)
    assert abs(OTab.k_exp_axis(np.array([0.0]), -10.0)[0] ** 2 - 2.624682843) < 1e-9


def test_transform_resample_nonuniform_axis_is_linear_interp():
    """A transform against a non-uniform model axis (an E0-shifted k_exp) resamples by
    linear interpolation: equal to np.interp inside the axis, zero outside."""
    k = np.linspace(0.0, 16.0, 321)
    ke = OTab.k_exp_axis(k, 3.0)
    t = OT.Transform(ke, 3.0, 14.0, 2, 1.0)
    chi = np.sin(1.7 * k) * np.exp(-0.1 * k)
    got = t.resample(chi)
    inside = (t.k_uniform >= ke[0]) & (t.k_uniform <= ke[-1])
    np.testing.assert_allclose(got[inside], np.interp(t.k_uniform[inside], ke, chi), rtol=0,
                               atol=1e-14)
    assert np.all(got[~inside] == 0.0)


def _small_displaced(seed=0, nsp=2):
    cell = OL.Supercell("fcc", (2, 2, 2), 3.59)
    i, j, sh = cell.neighbour_list(3.0)                   # first shell: 12 per site
    zs = [24, 26, 28][:nsp]
    k, f, d, lam, s02 = OTab.synthetic_ss(OTab.k_grid(), zs)
    rng = np.random.default_rng(seed)
    pos = cell.ideal + rng.normal(0, 0.05, cell.ideal.shape)
    sig2 = 0.004 + 0.002 * rng.random(i.size)
    return cell, i, j, sh, k, f, d, lam, s02, pos, sig2, rng


def test_grad_pos_finite_differences():
    """grad_pos (kernels.py:166-237) against central differences of chi_pairs, with
    cumulant tables set."""
    cell, i, j, sh, k, f, d, lam, s02, pos, sig2, rng = _small_displaced(1, 3)
    sp = rng.integers(0, 3, cell.natoms)
    shell = np.zeros(i.size, int)
    dr = rng.normal(0, 0.01, (1, 3, 3))
    c3 = np.array([3e-4])
    s2t = np.abs(rng.normal(0, 5e-4, (1, 3, 3)))
    g = rng.normal(0, 1, (3, k.size))
    kw = dict(shell=shell, dr=dr, c3=c3, s2types=s2t)
    G = OF.grad_pos(k, f, d, lam, s02, pos, sp, i, j, sh, sig2, g, **kw)
    h = 1e-6
    for atom, ax in [(0, 0), (5, 2), (17, 1), (31, 0)]:
        pp, pm = pos.copy(), pos.copy()
        pp[atom, ax] += h
        pm[atom, ax] -= h
        Lp = (g * OF.chi_pairs(k, f, d, lam, s02, pp, sp, i, j, sh, sig2, normalise=False, **kw)).sum()
        Lm = (g * OF.chi_pairs(k, f, d, lam, s02, pm, sp, i, j, sh, sig2, normalise=False, **kw)).sum()
        fd = (Lp - Lm) / (2 * h)
        assert abs(G[atom, ax] - fd) <= 1e-6 * max(1.0, abs(fd)), (atom, ax, G[atom, ax], fd)
    np.testing.assert_allclose(G.sum(0), 0.0, atol=1e-9 * np.abs(G).max())   # Newton's third law


def test_chi_soft_one_hot_and_grad_soft_finite_differences():
    """chi_soft (kernels.py:514-589) equals chi_pairs for one-hot weights; grad_soft
    (kernels.py:592-660 + the normalisation term of forward.py:252-258) against central
    differences in positions and weights."""
    cell, i, j, sh, k, f, d, lam, s02, pos, sig2, rng = _small_displaced(2, 2)
    sp = rng.integers(0, 2, cell.natoms)
    w1 = np.eye(2)[sp]
    np.testing.assert_allclose(OF.chi_soft(k, f, d, lam, s02, pos, w1, i, j, sh, sig2),
                               OF.chi_pairs(k, f, d, lam, s02, pos, sp, i, j, sh, sig2),
                               rtol=0, atol=1e-12)
    w = rng.dirichlet([1.0, 1.0], cell.natoms)
    g = rng.normal(0, 1, (2, k.size))

    def L(p_, w_):
        return (g * OF.chi_soft(k, f, d, lam, s02, p_, w_, i, j, sh, sig2)).sum()

    Gp, Gw = OF.grad_soft(k, f, d, lam, s02, pos, w, i, j, sh, sig2, g)
    h = 1e-6
    for atom, ax in [(0, 1), (9, 0), (30, 2)]:
        pp, pm = pos.copy(), pos.copy()
        pp[atom, ax] += h
        pm[atom, ax] -= h
        fd = (L(pp, w) - L(pm, w)) / (2 * h)
        assert abs(Gp[atom, ax] - fd) <= 1e-6 * max(1.0, abs(fd))
    for atom, s in [(0, 0), (3, 1), (22, 0)]:
        wp, wm = w.copy(), w.copy()
        wp[atom, s] += h
        wm[atom, s] -= h
        fd = (L(pos, wp) - L(pos, wm)) / (2 * h)
        assert abs(Gw[atom, s] - fd) <= 1e-6 * max(1.0, abs(fd)), (atom, s, Gw[atom, s], fd)


def _fdl(zs):
    def fdl(a, seq, p):
        F = np.ones_like(p)
        for s in seq:
            F = F * (zs[s] / 26.0) * np.exp(-(p - 4.0) ** 2 / 18.0)
        delta = -0.9 * p + 1.2 + 0.05 * np.mean([zs[s] for s in seq]) + 0.1 * a
        return F, delta, 6.0 + 0.3 * p
    return fdl


def test_feff_convention_round_trip(tmp_path):
    """FEFF columns written from chosen f(p), delta(p), lambda(p) through the documented
    convention (k from the Fermi level, amp = mag*red, pha = 2phc + phase) read back to the
    same functions on the p axis (exafs_gpu/scattering.py:158-190)."""
    from tests.feffio import write_library
    cl = OM.enumerate_classes("fcc", 3.59, 4.0)
    zs = [24, 28]
    root = write_library(str(tmp_path / "lib"), ("Cr", "Ni"), cl, 3.59 / 2, 3.59 / np.sqrt(2),
                         _fdl(zs), p0=0.6)
    fn = sorted((tmp_path / "lib" / "Ni").iterdir())[3]
    nleg, deg, reff, cols = OTab.read_feff_columns(str(fn))
    p, f, delta, lam = OTab.feff_columns_to_path(cols, reff)
    np.testing.assert_allclose(p ** 2 - cols[cols[:, 0] > 0.3, 0] ** 2, 0.36, rtol=1e-9)
    # entry 3 of the Ni (index 1) block: first class, slot species from the entry code
    g, sp = OM.entry_list(cl, 2)[3]
    seq = [1 if t < 0 else sp[t] for t in cl[g].slot_of]
    F, D, L = _fdl(zs)(1, seq, p)
    np.testing.assert_allclose(f, F, rtol=1e-9)
    np.testing.assert_allclose(delta, D, rtol=0, atol=1e-8)   # 13 digits in the file
    np.testing.assert_allclose(lam, L, rtol=1e-11)
    assert nleg == len(seq) + 1 and abs(reff - cl[g].reff_units() * 3.59 / 2) < 1e-7
    # the tabulated table at the synthetic functions equals the synthetic closed form
    k = OTab.k_grid()
    Fk, Dk, Lk = (np.interp(k, p, v)[None] for v in (f, delta, lam))
    T = OTab.tabulated_table(k, Fk, Dk, Lk, [reff], 0.9, 0.006, 2.54)
    nz = (k > 0) & (k >= p[0])
    kk = k[nz]
    r = reff
    want = (0.9 * Fk[0, nz] / (kk * r * r) * np.exp(-2 * r / Lk[0, nz] - 2 * kk * kk * 1.3 * 0.006 * r / (2 * 2.54))
            * np.sin(2 * kk * r + Dk[0, nz]))
    np.testing.assert_allclose(T[0, nz], want, rtol=1e-12, atol=1e-15)


def test_posterior_cov_qr_equals_normal_equations():
    """oracle.shellmodel.posterior_cov (QR of [J; sqrt(P)]) equals (J^T J + P)^-1."""
    from oracle import shellmodel as OS
    rng = np.random.default_rng(5)
    J = rng.normal(size=(200, 6))
    P = rng.uniform(0.1, 2.0, 6)
    np.testing.assert_allclose(OS.posterior_cov(J, P), np.linalg.inv(J.T @ J + np.diag(P)),
                               rtol=1e-10, atol=1e-14)


def test_cpu_worker_runs(tmp_path):
    """The CPU-baseline worker (a separate process, NumPy only) runs the plain sampler on a
    pickled problem and reports its replica moves and wall time."""
    import json
    import pickle
    import subprocess
    import sys
    P = _toy_problem()
    path = tmp_path / "p.pkl"
    with open(path, "wb") as fh:
        pickle.dump((P, OL.random_config(108, [1, 1], 5)), fh)
    root = str(__import__("pathlib").Path(__file__).resolve().parents[1])
    out = subprocess.run([sys.executable, "-m", "oracle.cpu_worker", str(path), "0.3", "3", "1"],
                         cwd=root, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert r["moves"] == 3 * r["steps"] and r["seconds"] > 0


# ------------------------------------------------------------------ fit scales, R-factor
def _fit_data(seed, complex_=True, nchan=6, nR=40):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((nchan, nR)) + (1j * rng.standard_normal((nchan, nR)) if complex_ else 0)
    d = rng.standard_normal((nchan, nR)) + (1j * rng.standard_normal((nchan, nR)) if complex_ else 0)
    mask = np.ones(nR, bool)
    mask[:5] = mask[-3:] = False
    return m, d, mask


@pytest.mark.parametrize("complex_", [True, False])
def test_fit_scales_minimise_the_residual(complex_):
    """chimera/model.py:497-512: the fitted amplitude (per channel and per edge) is the
    minimum of the windowed residual; any perturbation raises the R-factor."""
    from oracle import fitting as OFIT
    m, d, mask = _fit_data(1, complex_)
    edge = np.repeat(np.arange(3), 2)
    for e in (None, edge):
        s = OFIT.fit_scales(m, d, mask, e)
        r0 = OFIT.r_factor(m, d, mask, s)
        rng = np.random.default_rng(2)
        for _ in range(20):
            dv = rng.standard_normal(s.size) * 1e-3
            if e is not None:
                dv = dv[e]                    # one amplitude per edge
            assert OFIT.r_factor(m, d, mask, s + dv) > r0


def test_r_factor_closed_forms():
    """R = 0 for data = s0 * model (and the fit returns s0); with data = s0 m + n, n
    orthogonal to m on the window, the fit returns s0 and R = |n|^2/(s0^2|m|^2 + |n|^2);
    data scaled by c scales the amplitude by c and leaves R; per-edge fitting equals a
    fit over the edge's channels laid end to end (chimera/model.py:497-540)."""
    from oracle import fitting as OFIT
    m, d, mask = _fit_data(3)
    s0 = 0.83
    s = OFIT.fit_scales(m, s0 * m, mask)
    np.testing.assert_allclose(s, s0, rtol=1e-14)
    assert OFIT.r_factor(m, s0 * m, mask, s) < 1e-28
    # noise orthogonal to each channel's model on the window (Re<m, n> = 0)
    rng = np.random.default_rng(4)
    n = rng.standard_normal(m.shape) + 1j * rng.standard_normal(m.shape)
    n[:, ~mask] = 0
    for c in range(m.shape[0]):
        mc = np.where(mask, m[c], 0)
        n[c] -= np.real(np.vdot(mc, n[c])) / np.real(np.vdot(mc, mc)) * mc
    data = s0 * m + n
    s = OFIT.fit_scales(m, data, mask)
    np.testing.assert_allclose(s, s0, rtol=1e-12)
    nm = np.sum(np.abs(n[:, mask]) ** 2)
    mm = np.sum(np.abs(m[:, mask]) ** 2)
    np.testing.assert_allclose(OFIT.r_factor(m, data, mask, s), nm / (s0 ** 2 * mm + nm), rtol=1e-12)
    # scale covariance
    s1 = OFIT.fit_scales(m, d, mask)
    np.testing.assert_allclose(OFIT.fit_scales(m, 2.5 * d, mask), 2.5 * s1, rtol=1e-13)
    np.testing.assert_allclose(OFIT.r_factor(m, 2.5 * d, mask, 2.5 * s1),
                               OFIT.r_factor(m, d, mask, s1), rtol=1e-12)
    # per edge == end-to-end channels of the edge
    edge = np.repeat(np.arange(3), 2)
    se = OFIT.fit_scales(m, d, mask, edge)
    for e in range(3):
        cat_m = np.concatenate([m[c][mask] for c in np.flatnonzero(edge == e)])[None]
        cat_d = np.concatenate([d[c][mask] for c in np.flatnonzero(edge == e)])[None]
        one = OFIT.fit_scales(cat_m, cat_d, np.ones(cat_m.shape[1], bool))
        np.testing.assert_allclose(se[edge == e], one[0], rtol=1e-12)


@pytest.mark.parametrize("complex_", [True, False])
def test_host_fit_scales_and_r_factor_equal_the_oracle(complex_):
    """chimera_amd.workloads.fit_scales / r_factor (the host post-processing of a fit)
    against the oracle on random channels, per channel and per edge."""
    from chimera_amd import workloads as W
    from oracle import fitting as OFIT
    m, d, mask = _fit_data(5, complex_)
    edge = np.repeat(np.arange(3), 2)
    for e in (None, edge):
        s_w, s_o = W.fit_scales(m, d, mask, e), OFIT.fit_scales(m, d, mask, e)
        np.testing.assert_allclose(s_w, s_o, rtol=1e-13)
        np.testing.assert_allclose(W.r_factor(m, d, mask, s_w), OFIT.r_factor(m, d, mask, s_o),
                                   rtol=1e-13)
