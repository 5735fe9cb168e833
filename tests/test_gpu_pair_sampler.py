"""The pair-space device sampler (every move of the reference) against the oracle's
pair-space sampler, move by move, and its tracked state against recomputes from scratch."""
import sys

import numpy as np
import pytest

from oracle import pair_rmc as PR

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(name, nc, els, ms, nrep=6, a0=3.59, rmax=5.2):
    from chimera_amd import workloads as W
    _say(f"[build] {name} {nc} ms={ms} nrep={nrep}")
    return W.Problem(W.Workload(name, "fcc", nc, a0, els, rmax, ms, nrep))


@pytest.fixture(scope="module")
def tern108():
    from oracle.build import problem_from
    P = _problem("pt108", (3, 3, 3), ("Cr", "Co", "Ni"), True)
    return P, problem_from(P)


def _sampler(P, **kw):
    from chimera_amd.method_a_rmc import RMCSampler
    args = dict(nrep=P.sampler.nrep, ms_library=P.ms)
    args.update(kw)
    tgt = args.pop("target", P.target)
    return RMCSampler(P.fm, P.ft, tgt, P.mask, P.sigma, **args)


def _compare(P, O, S, ref, nsteps, exch, start_pos=None, tie_tol=1e-6):
    """Run the device sampler S and the oracle sampler ref from the same start, then
    compare move by move (site pair, kind, decision, trial chi^2) and the final state.
    A differing decision is allowed only at a near tie (the comparison stops there)."""
    pos0 = P.cell.ideal if start_pos is None else start_pos
    S.set_recording(nsteps)
    S.init_state(P.cfg0, pos0)
    _say(f"[run] device {nsteps} steps nrep={S.nrep} swap_frac={S.swap_frac}")
    S.run(nsteps, exchange_every=exch, log_every=10 ** 9)
    log = S.move_log(nsteps)
    lc2 = S.trial_chi2_log(nsteps)
    ref.init_state(P.cfg0, pos0)
    _say("[run] oracle")
    ref.run(nsteps, exchange_every=exch)
    kinds = set()
    for t in range(nsteps):
        for rung in range(S.nrep):
            a, b, sw, acc, c2, x, u = ref.log[t][rung][:7]
            g = log[t, rung]
            assert (int(g[0]), int(g[1]), int(g[2]) & 1) == (a, b, sw), f"step {t} rung {rung}"
            np.testing.assert_allclose(lc2[t, rung], c2, rtol=1e-9, err_msg=f"step {t} rung {rung}")
            if (int(g[2]) >> 1) != acc:
                assert abs(x + np.log(u)) < tie_tol, f"step {t} rung {rung}: x {x} u {u}"
                return None
            kinds.add((sw, acc))
    assert np.array_equal(S.species, ref.species)
    np.testing.assert_allclose(S.pos, ref.pos, rtol=0, atol=1e-12)
    np.testing.assert_allclose(S.chi2, ref.chi2, rtol=1e-9)
    np.testing.assert_allclose(S.best_chi2, ref.best_chi2, rtol=1e-9)
    chi = S.chi
    np.testing.assert_allclose(chi, ref.chi, rtol=0, atol=1e-10 * np.abs(ref.chi).max())
    return kinds


def test_dispatch_and_defaults(tern108):
    """The reference's defaults (swap_frac 0.75) give the pair-space engine; swap-only on
    a complex target stays on count space and turns into pair space when started off the
    lattice."""
    from chimera_amd.method_a_rmc import RMCSampler
    from chimera_amd.pair_rmc import PairRMCSampler
    P, _ = tern108
    S = RMCSampler(P.fm, P.ft, P.target, P.mask, P.sigma, nrep=4)
    assert isinstance(S, PairRMCSampler) and S.engine == "pair" and S.swap_frac == 0.75
    T = RMCSampler(P.fm, P.ft, P.target, P.mask, P.sigma, nrep=4, swap_frac=1.0)
    assert T.engine == "count"
    rng = np.random.default_rng(1)
    pos = P.cell.ideal + rng.normal(0, 0.02, P.cell.ideal.shape)
    T.init_state(P.cfg0, pos)
    assert T.engine == "pair" and isinstance(T, PairRMCSampler)
    np.testing.assert_allclose(T.pos[0], pos, rtol=0, atol=0)
    T.run(20, exchange_every=10, log_every=10)
    assert T.accepted_per_rung().sum() > 0
    # swap-only: no position changed
    np.testing.assert_array_equal(T.pos, np.repeat(pos[None], 4, 0))
    for bad in (dict(precision="bf16x3"),):
        with pytest.raises(ValueError):
            RMCSampler(P.fm, P.ft, P.target, P.mask, P.sigma, nrep=4, **bad)


def test_trajectory_swaps_and_displacements_ms(tern108):
    """Complex target, multiple scattering, swaps and displacements (swap_frac 0.6)."""
    P, O = tern108
    S = _sampler(P, swap_frac=0.6, disp_sigma=0.03, seed=5)
    ref = PR.PairSampler(O, nrep=S.nrep, seed=5, swap_frac=0.6, disp_sigma=0.03)
    kinds = _compare(P, O, S, ref, 50, 10)
    if kinds is not None:
        assert (1, 1) in kinds and (0, 1) in kinds
    np.testing.assert_allclose(S.best_verified, S.best_chi2.min(), rtol=1e-9)


def test_trajectory_magnitude_prior_off_lattice(tern108):
    """|chi(R)| target, the energy prior, an off-lattice start, displacements only part
    of the time (swap_frac 0.5), the clamp active (max_disp 0.05)."""
    P, O = tern108
    rng = np.random.default_rng(3)
    V = rng.normal(0, 0.03, (2, 3, 3))
    V = 0.5 * (V + V.transpose(0, 2, 1))
    tabs = np.abs(P.target)
    S = _sampler(P, target=tabs, swap_frac=0.5, disp_sigma=0.04, max_disp=0.05, seed=9,
                 energy_prior=(V, 900.0))
    assert S.engine == "pair" and not S.complex and S.nfit == S.nwin * S.nchan
    ref = PR.PairSampler(O, nrep=S.nrep, seed=9, swap_frac=0.5, disp_sigma=0.04, max_disp=0.05,
                         energy_prior=(V, 900.0), magnitude=True, target_abs=tabs)
    pos0 = P.cell.ideal + rng.normal(0, 0.01, P.cell.ideal.shape)
    _compare(P, O, S, ref, 40, 10, start_pos=pos0)
    off = np.linalg.norm(S.pos - P.cell.ideal[None], axis=-1)
    assert off.max() <= 0.05 + 1e-12          # moved atoms clamped; the start is within it
    E = S.energy()
    want = [ref.energy(sp) for sp in S.species]
    np.testing.assert_allclose(E, want, rtol=1e-12, atol=1e-12)


def test_state_consistency_4000():
    """4000 sites, 16 replicas, 300 steps of the default move mix with exchanges: every
    replica's tracked chi equals the path sum of its species and positions (+ MS), its
    pair distances and types equal a recount, chi^2 equals chi^2 of that chi, and the
    best state's chi^2 is reproduced from scratch."""
    from chimera_amd import workloads as W
    _say("[build] fcc_crconi_4000 nrep=16")
    P = W.build("fcc_crconi_4000", nrep=16)
    S = _sampler(P, nrep=16, seed=2)
    S.init_state(P.cfg0, P.cell.ideal)
    _say("[run] 300 steps")
    S.run(300, exchange_every=25, log_every=100)
    assert S.accepted_per_rung().sum() > 0
    sp, pos = S.species, S.pos
    chi = S.chi
    want = S.chi_k_of(sp, pos)
    np.testing.assert_allclose(chi, want, rtol=0, atol=1e-9 * np.abs(want).max())
    r, _ = P.fm.distances(pos)
    np.testing.assert_allclose(S.r, r, rtol=1e-14)
    assert np.array_equal(S.ptype, P.fm.ptypes(sp))
    np.testing.assert_allclose(S.chi2, S.chi2_of(sp, pos), rtol=1e-9)
    np.testing.assert_allclose(S.best_verified, S.best_chi2.min(), rtol=1e-9)
    assert np.abs(pos - P.cell.ideal[None]).max() > 0
    assert np.abs(pos - P.cell.ideal[None]).max() <= 0.30 + 1e-9


def test_adapt_retunes_displacement(tern108):
    """adapt_every: the displacement step follows the window acceptance
    (method_a_rmc.py:376-398) and reaches the device (the chain still matches a recount)."""
    P, _ = tern108
    S = _sampler(P, swap_frac=0.3, disp_sigma=0.001, adapt_every=20, seed=4)
    S.init_state(P.cfg0, P.cell.ideal)
    S.run(100, exchange_every=10, log_every=50)
    assert S.disp_sigma > 0.001            # a tiny step is almost always accepted: it grows
    want = S.chi_k_of(S.species, S.pos)
    np.testing.assert_allclose(S.chi, want, rtol=0, atol=1e-10 * np.abs(want).max())


def test_trajectory_binary_unprofiled_prior():
    """Two species without multiple scattering, the scale not profiled
    (profile_scale=False), the energy prior, swap_frac 0.4, 500 sites: move by move
    against the oracle; energy_delta() holds the last proposals' energy change."""
    from oracle.build import problem_from
    P = _problem("pb500", (5, 5, 5), ("Fe", "Ni"), False, nrep=4)
    O = problem_from(P)
    V = np.array([[[0.0, -0.025], [-0.025, 0.0]], [[0.0, 0.01], [0.01, 0.0]]])
    S = _sampler(P, nrep=4, swap_frac=0.4, disp_sigma=0.02, seed=11, profile_scale=False,
                 energy_prior=(V, 700.0), ms_library=None)
    ref = PR.PairSampler(O, nrep=4, seed=11, swap_frac=0.4, disp_sigma=0.02, profile_scale=False,
                         energy_prior=(V, 700.0))
    kinds = _compare(P, O, S, ref, 35, 10)                   # no exchange after the last step
    if kinds is not None:
        want = [r[8] for r in ref.log[-1]]                    # rung order
        np.testing.assert_allclose(S.energy_delta(), want, rtol=1e-12, atol=1e-15)


def test_trajectory_four_species():
    """Four species (the NSP = 4 kernels), complex target, no MS, default move mix."""
    from oracle.build import problem_from
    P = _problem("pq108", (3, 3, 3), ("Cr", "Mn", "Fe", "Ni"), False, nrep=4)
    O = problem_from(P)
    S = _sampler(P, nrep=4, seed=3, ms_library=None)
    assert S.engine == "pair" and S.swap_frac == 0.75
    ref = PR.PairSampler(O, nrep=4, seed=3)
    _compare(P, O, S, ref, 30, 10)
