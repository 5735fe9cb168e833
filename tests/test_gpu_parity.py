"""Parity of the HIP path (through the C ABI) with the FP64 oracle on the same inputs."""
import sys

import numpy as np
import pytest

from oracle import forward as OF
from oracle import fourier as OT
from oracle import rmc as OR

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(name, lat, nc, a0, els, rmax, ms, nrep, seed=0):
    from chimera_amd import workloads as W
    _say(f"[build] {name} {nc} ms={ms} nrep={nrep}")
    return W.Problem(W.Workload(name, lat, nc, a0, els, rmax, ms, nrep), seed=seed)


def _oracle_problem(P):
    from oracle.build import problem_from
    return problem_from(P)


@pytest.fixture(scope="module")
def tern108():
    P = _problem("t108", "fcc", (3, 3, 3), 3.59, ("Cr", "Co", "Ni"), 5.2, True, 6)
    return P, _oracle_problem(P)


@pytest.fixture(scope="module")
def bin500():
    P = _problem("b500", "fcc", (5, 5, 5), 3.59, ("Fe", "Ni"), 5.2, False, 8)
    return P, _oracle_problem(P)


def test_path_sum_f64_f32(tern108):
    P, O = tern108
    _say("[test] path sum fp64/fp32 vs oracle")
    rng = np.random.default_rng(0)
    sp = np.stack([P.cfg0, P.ref_species, rng.permutation(P.cfg0)]).astype(np.int32)
    pos = np.repeat(P.cell.ideal[None], 3, 0)
    pos[2] += rng.normal(0, 0.03, pos[2].shape)          # displaced configuration too
    ref = np.stack([OF.chi_pairs(P.k, P.tables.f, P.tables.delta, P.tables.lam, P.tables.s02,
                                 pos[c], sp[c], P.i, P.j, P.shift, P.sig2, normalise=False)
                    for c in range(3)])
    scale = np.abs(ref).max()
    got = P.fm.chi(pos, sp, normalise=False)
    assert np.abs(got - ref).max() / scale < 1e-12
    got32 = P.fm.chi(pos, sp, normalise=False, precision="f32")
    assert np.abs(got32 - ref).max() / scale < 1e-5
    # normalised by atoms of the absorber species (forward.py:183-187)
    gn = P.fm.chi(pos[:1], sp[:1])
    cnt = np.bincount(sp[0], minlength=3)
    np.testing.assert_allclose(gn[0], ref[0] / cnt[:, None], rtol=1e-12, atol=1e-14 * scale)


def test_cumulants_dr_c3_s2(tern108):
    P, O = tern108
    _say("[test] cumulant tables")
    ns, nsp = P.fm.nshell, P.nsp
    rng = np.random.default_rng(1)
    dr = rng.normal(0, 0.01, (ns, nsp, nsp))
    c3 = rng.normal(0, 1e-3, ns)
    s2 = np.abs(rng.normal(0, 1e-3, (ns, nsp, nsp)))
    P.fm.set_dr(dr).set_c3(c3).set_sig2_types(s2)
    try:
        sp = P.cfg0[None]
        got = P.fm.chi(P.cell.ideal[None], sp, normalise=False)[0]
        ref = OF.chi_pairs(P.k, P.tables.f, P.tables.delta, P.tables.lam, P.tables.s02,
                           P.cell.ideal, P.cfg0, P.i, P.j, P.shift, P.sig2, shell=P.shell, dr=dr,
                           c3=c3, s2types=s2, normalise=False)
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
        B = P.fm.ss_basis(P.radii, np.full(ns, P.w.sigma2))
        Bo = OF.ss_basis(P.k, P.tables.f, P.tables.delta, P.tables.lam, P.tables.s02, P.radii,
                         np.full(ns, P.w.sigma2), dr=dr, c3=c3, s2types=s2)
        assert np.abs(B - Bo).max() / np.abs(Bo).max() < 1e-13
    finally:
        P.fm.set_dr(np.zeros((ns, nsp, nsp))).set_c3(np.zeros(ns)).set_sig2_types(
            np.zeros((ns, nsp, nsp)))


def test_distances_ptypes(tern108):
    P, O = tern108
    pos = P.cell.ideal[None] + np.random.default_rng(2).normal(0, 0.02, (1,) + P.cell.ideal.shape)
    r, ev = P.fm.distances(pos)
    d = (pos[0][P.j] + P.shift) - pos[0][P.i]
    np.testing.assert_allclose(r[0], np.linalg.norm(d, axis=1), rtol=1e-15)
    np.testing.assert_allclose(ev[0], d / np.linalg.norm(d, axis=1)[:, None], rtol=1e-14, atol=1e-15)
    pt = P.fm.ptypes(P.cfg0[None])
    assert np.array_equal(pt[0], P.cfg0[P.i] * 3 + P.cfg0[P.j])


def test_transform_fft_grid_and_window_grid(tern108):
    P, O = tern108
    _say("[test] transform")
    from chimera_amd.fourier import FourierTransform
    rng = np.random.default_rng(3)
    chi = rng.standard_normal((4, P.k.size))
    ft = FourierTransform(P.k, 3.0, 14.0, 2, 1.0)              # FFT grid, 1024 points
    got = ft.to_R(chi)
    ref = OT.Transform(P.k, 3.0, 14.0, 2, 1.0).to_R(chi)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
    chi3 = rng.standard_normal((2, P.nsp, P.k.size))
    got = P.ft.to_R(chi3)
    ref = np.stack([OT.multi_to_R(O.transforms, P.chan_edge, c, R=P.R) for c in chi3])
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12


def test_entry_counts_exact(tern108):
    P, O = tern108
    _say("[test] entry counts")
    S = P.sampler
    rng = np.random.default_rng(4)
    sps = np.stack([P.cfg0, P.ref_species, rng.permutation(P.cfg0)]).astype(np.int32)
    C = S.counts_of(sps)
    nss = P.nsp * P.fm.nshell
    nrow_e = S.nrow_e
    for c in range(3):
        Co = O.ms.counts(sps[c])
        N = OF.pair_counts(sps[c], P.i, P.j, P.shell, P.nsp, P.fm.nshell)
        Cg = C[c].reshape(P.nsp, nrow_e)
        assert np.array_equal(Cg[:, :nss].reshape(N.shape), N)
        assert np.array_equal(Cg[:, nss:].reshape(-1), Co)


def test_chi_of_counts_matches_pair_space(tern108):
    P, O = tern108
    S = P.sampler
    got = S.chi_k_of(P.ref_species[None])[0]
    ref = O.chi_full(P.ref_species.astype(np.int64))
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
    S.init_state(P.cfg0, P.cell.ideal)
    cnt = np.bincount(P.cfg0, minlength=3).astype(float)
    c2_ref = O.chi2(O.chi_full(P.cfg0.astype(np.int64)), cnt)
    np.testing.assert_allclose(S.chi2, c2_ref, rtol=1e-10)
    np.testing.assert_allclose(S.chi, np.repeat(O.chi_full(P.cfg0)[None], S.nrep, 0),
                               rtol=0, atol=1e-12 * np.abs(ref).max())


def _compare_trajectory(P, O, nsteps, exch, seed, tie_tol=0.0, c2_rtol=1e-9, adapt=0):
    """Device and oracle move by move.  With tie_tol > 0 a differing decision is
    allowed only at a near tie (|x + log u| < tie_tol); the comparison stops there."""
    S = P.sampler
    S.seed = seed
    S.set_recording(log_steps=nsteps)
    S.init_state(P.cfg0, P.cell.ideal)
    _say(f"[run] device {nsteps} steps nrep={S.nrep} exchange_every={exch} adapt_every={adapt}")
    S.run(nsteps, exchange_every=exch, log_every=10 ** 9)
    log = S.move_log(nsteps)                                    # (steps, slot, 4)
    R = OR.PlainSampler(O, nrep=S.nrep, seed=seed)
    R.init_state(P.cfg0)
    _say("[run] oracle")
    R.run(nsteps, exchange_every=exch, adapt_every=adapt)
    for t in range(nsteps):
        by_rung = {int(r[3]): r for r in log[t]}
        for rung in range(S.nrep):
            a, b, acc, c2, x, u = R.log[t][rung]
            g = by_rung[rung]
            assert (int(g[0]), int(g[1])) == (a, b), f"step {t} rung {rung}"
            if int(g[2]) != acc:
                assert tie_tol > 0 and abs(x + np.log(u)) < tie_tol, f"step {t} rung {rung}"
                return S, R
    assert np.array_equal(S.species, R.species)
    np.testing.assert_allclose(S.chi2, R.chi2, rtol=c2_rtol)
    np.testing.assert_allclose(S.best_chi2, R.best_chi2, rtol=c2_rtol)
    return S, R


def test_sampler_trajectory_ternary_ms(tern108):
    P, O = tern108
    S, R = _compare_trajectory(P, O, 60, 10, seed=5)
    assert S.accepted_per_rung().sum() > 0
    # counts tracked incrementally equal a recount from scratch
    C = S.counts
    assert np.array_equal(C, S.counts_of(S.species))
    np.testing.assert_allclose(S.best_verified, S.best_chi2.min(), rtol=1e-9)


def test_sampler_trajectory_bf16x3(tern108):
    """Split-BF16 incremental transform with FP64 running chi(R) and re-anchoring."""
    P, O = tern108
    S = P.sampler
    S.precision, S.reanchor_every = "bf16x3", 4
    try:
        S2, R = _compare_trajectory(P, O, 60, 10, seed=5, tie_tol=1e-4, c2_rtol=2e-6)
        # running accumulator vs the exact FP64 transform of the same counts
        m = S.running_chiR()
        C = S.counts_slot_order()
        exact = np.stack([S.chi2_of(S.species)])  # noqa: F841 (exercises the FP64 path)
        c2 = S.chi2
        np.testing.assert_allclose(c2, S.chi2_of(S.species), rtol=2e-6)
        assert S.reanchors().sum() > 0
        assert np.all(np.isfinite(m)) and C.shape[0] == S.nrep
    finally:
        S.precision, S.reanchor_every = "f64", 512
        S.init_state(P.cfg0, P.cell.ideal)


@pytest.mark.parametrize("prec,acc,rean,tol", [("bf16x3", "f64", 4, 2e-6), ("f32", "f64", 4, 2e-6),
                                              ("bf16x3", "f32", 0, 1e-4), ("f32", "f32", 0, 1e-4)])
def test_precision_modes_running_chiR(tern108, prec, acc, rean, tol):
    """Split-BF16 and FP32 transforms, FP64 and FP32 accumulators: the running chi(R) stays
    within tol of the exact FP64 transform of the integer counts, and the chain follows
    the FP64 oracle move by move up to a near tie."""
    P, O = tern108
    S = P.sampler
    S.precision, S.accumulator, S.reanchor_every = prec, acc, rean
    try:
        _say(f"[test] precision {prec} acc {acc} reanchor {rean}")
        _compare_trajectory(P, O, 60, 10, seed=9, tie_tol=1e-3, c2_rtol=tol)
        m, ex = S.running_chiR(), S.exact_chiR()
        assert np.abs(m - ex).max() <= tol * np.abs(ex).max()
        if rean:
            assert S.reanchors().sum() > 0
    finally:
        S.precision, S.accumulator, S.reanchor_every = "f64", "f64", 512
        S.init_state(P.cfg0, P.cell.ideal)


def test_sampler_trajectory_binary_ss(bin500):
    P, O = bin500
    _compare_trajectory(P, O, 40, 5, seed=11)


def test_ladder_adaptation(tern108):
    """adapt_every (method_a_rmc.py:376-425): the device chain with its ladder retuned
    from the exchange rates of every window follows the oracle move by move, and both
    end on the same ladder (the same host arithmetic on the same rates)."""
    P, O = tern108
    S = P.sampler
    beta0 = S.beta
    S.adapt_every = 20
    try:
        S2, R = _compare_trajectory(P, O, 80, 5, seed=13, adapt=20)
        assert not np.allclose(S.beta, beta0, rtol=1e-6)          # the ladder moved
        np.testing.assert_allclose(S.beta[[0, -1]], beta0[[0, -1]], rtol=1e-13)
        np.testing.assert_allclose(S.beta, R.beta, rtol=1e-14)
    finally:
        S.adapt_every = 0
        S._beta_rung = beta0
        S.init_state(P.cfg0, P.cell.ideal)


def test_null_moves_when_candidates_repeat(bin500):
    P, O = bin500
    S = P.sampler
    sp = np.zeros(P.cell.natoms, np.int32)
    sp[:3] = 1                                   # almost every proposal finds no partner
    S.set_recording(log_steps=20)
    S.init_state(sp, P.cell.ideal)
    S.run(20, exchange_every=0, log_every=10 ** 9)
    R = OR.PlainSampler(O, nrep=S.nrep, seed=S.seed)
    R.init_state(sp)
    R.run(20, exchange_every=0)
    assert np.array_equal(S.species, R.species)
    log = S.move_log(20)
    nulls = sum(1 for t in range(20) for r in R.log[t] if r[0] == r[1])
    assert nulls > 0 and (log[..., 2] == 0).sum() >= nulls
    S.init_state(P.cfg0, P.cell.ideal)


def test_replay_determinism(tern108):
    P, O = tern108
    S = P.sampler
    S.set_recording(log_steps=40)
    logs, sps = [], []
    for _ in range(2):
        S.init_state(P.cfg0, P.cell.ideal)
        S.run(40, exchange_every=7, log_every=10 ** 9)
        logs.append(S.move_log(40).copy())
        sps.append(S.species.copy())
    assert np.array_equal(logs[0], logs[1]) and np.array_equal(sps[0], sps[1])
    # replaying the recorded moves on the CPU reproduces the species bit-exactly
    sp = np.repeat(P.cfg0[None], S.nrep, 0).astype(np.int64)
    for t in range(40):
        for slot in range(S.nrep):
            a, b, acc, rung = logs[0][t, slot]
            if acc:
                sp[slot, a], sp[slot, b] = sp[slot, b], sp[slot, a]
    rung_of = S._read(5, (S.nrep,), np.int32)
    assert np.array_equal(sp[np.argsort(rung_of)], sps[0])


def test_launch_split_invariance(tern108):
    """The trajectory does not depend on how many workgroups share a replica's gather
    (the autotuner's knob): every split takes the same decisions and the integer count
    changes merge exactly."""
    from chimera_amd import autotune
    P, O = tern108
    S = P.sampler
    S.set_recording(log_steps=30)
    outs = []
    for ns in (1, 3, 8):
        S.init_state(P.cfg0, P.cell.ideal)
        from chimera_amd import _lib
        _lib.check(_lib.load().chm_sampler_set_tuning(S._S, ns))
        assert autotune.current(S) == ns
        S.run(30, exchange_every=7, log_every=10 ** 9)
        outs.append((S.move_log(30).copy(), S.species.copy(), S.chi2.copy(), S.counts.copy()))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
    rec = autotune.tune(S, candidates=(1, 2), iters=5, use_cache=False)
    assert rec["nsplit"] in (1, 2) and set(rec["times_ms"]) == {"1", "2"}


def test_workload_4000_consistency():
    """Full-size workload: incremental counts == recount, chi^2 from counts == chi^2 from
    the pair-space path sum + entry counts, composition conserved."""
    from chimera_amd import workloads as W
    _say("[build] fcc_crconi_4000")
    P = W.build("fcc_crconi_4000")
    S = P.sampler
    c2_0 = S.chi2.copy()
    _say("[run] 2000 steps")
    S.run(2000, exchange_every=25, log_every=10 ** 9)
    sp = S.species
    assert all(np.array_equal(np.bincount(s, minlength=3), np.bincount(P.cfg0, minlength=3))
               for s in sp)
    assert np.array_equal(S.counts, S.counts_of(sp))
    np.testing.assert_allclose(S.chi2, S.chi2_of(sp), rtol=1e-9)
    assert S.best_chi2.min() < c2_0.min()


def test_collection_and_posterior_wc(bin500):
    """Posterior collection (method_a_rmc.py:401-491): every collect_every steps after
    collect_after the rungs nearest collect_beta are stored; posterior_wc averages the
    Warren-Cowley parameters of the stored configurations.  Checked against the oracle
    chain snapshotted at the same steps and the oracle's Warren-Cowley by shell index."""
    from oracle import lattice as OL
    P, O = bin500
    S = P.sampler
    S.collect_after, S.collect_every, S.n_collect = 0, 10, 3
    try:
        S.seed = 17
        S.init_state(P.cfg0, P.cell.ideal)
        _say("[run] device with collection")
        S.run(40, exchange_every=5, log_every=10 ** 9)
        R = OR.PlainSampler(O, nrep=S.nrep, seed=17, n_collect=3)
        R.init_state(P.cfg0)
        rungs = R.collect_rungs()
        assert np.array_equal(np.sort(S.collect_rungs()), np.sort(rungs))
        snaps = []
        for _ in range(4):
            R.run(10, exchange_every=5)
            snaps += [R.species[r].copy() for r in S.collect_rungs()]
        assert len(S.samples) == len(snaps) == 12
        for (sp, c2), ref in zip(S.samples, snaps):
            assert np.array_equal(sp, ref)
        mean, std, n = S.posterior_wc(float(P.radii[0]))
        iu = np.triu_indices(2)
        ref = np.array([OL.warren_cowley(sp, P.i, P.shell, P.j, 0, 2)[iu] for sp in snaps])
        assert n == 12
        np.testing.assert_allclose(mean, ref.mean(0), rtol=0, atol=1e-12)
        np.testing.assert_allclose(std, ref.std(0), rtol=0, atol=1e-12)
        # the radius form of the reference (pair lists and cell given explicitly)
        m2, _, _ = S.posterior_wc(float(P.radii[1]), P.i, P.j, P.shift, P.cell, nsp=2)
        ref1 = np.array([OL.warren_cowley(sp, P.i, P.shell, P.j, 1, 2)[iu] for sp in snaps])
        np.testing.assert_allclose(m2, ref1.mean(0), rtol=0, atol=1e-12)
    finally:
        S.collect_after, S.collect_every, S.n_collect = None, 25, 4
        S.samples = []


def test_target_preparation(bin500):
    """Target and model-error sigma of the workload (chimera/fitrun.py:267-288): the data's
    chi(R) over the random start configuration's per-edge amplitude, sigma the RMS residual
    of that configuration at its amplitude, per component, over the amplitude.  Recomputed
    with the oracle's transform and pair sum."""
    P, O = bin500
    t = P.tables
    chi_d = P.chi_data
    trs = O.transforms
    tgt = OT.multi_to_R(trs, P.chan_edge, chi_d, R=P.R)
    cnt = np.bincount(P.cfg0, minlength=P.nsp)
    chi0 = OF.chi_pairs(P.k, t.f, t.delta, t.lam, t.s02, P.cell.ideal, P.cfg0, P.i, P.j,
                        P.shift, P.sig2, normalise=False) / np.maximum(cnt, 1)[:, None]
    m0 = OT.multi_to_R(trs, P.chan_edge, chi0, R=P.R)
    sc = np.empty(P.nchan)
    for e in range(P.nsp):
        ce = P.chan_edge == e
        sc[ce] = (m0[ce].conj() * tgt[ce]).real.sum() / (np.abs(m0[ce]) ** 2).sum()
    np.testing.assert_allclose(P.scales0, sc, rtol=1e-10)
    np.testing.assert_allclose(P.target, tgt / sc[:, None], rtol=1e-10, atol=1e-14)
    sig = np.sqrt((np.abs(m0 * sc[:, None] - tgt) ** 2).mean(1) / 2.0) / sc
    np.testing.assert_allclose(P.sigma[:, 0], sig, rtol=1e-9)
