"""Learned swap proposal on the device (through the C ABI) against the oracle:
network output, gradients, move records and the sampler trajectory step by step."""
import sys

import numpy as np
import pytest

from oracle import learned as LP
from oracle import rmc as OR

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tern108():
    from chimera_amd import workloads as W
    from oracle.build import problem_from
    _say("[build] t108 fcc 3x3x3 ternary MS nrep=6")
    P = W.Problem(W.Workload("t108", "fcc", (3, 3, 3), 3.59, ("Cr", "Co", "Ni"), 5.2, True, 6), seed=0)
    P.sampler.init_state(P.cfg0, P.cell.ideal)
    return P, problem_from(P)


def _stencil(P):
    return LP.Stencil(P.cell.ideal, P.i, P.j, P.shift, P.shell, P.cell.box)


def _random_params(lp, seed, w3_scale=2.0):
    rng = np.random.default_rng(seed)
    flat = lp.initial_params(seed)
    o = lp.offsets()
    for name, sc in (("w3", w3_scale / np.sqrt(256)), ("b1", 0.2), ("b2", 0.2), ("b3", 0.5)):
        off, n = o[name]
        flat[off:off + n] = (rng.standard_normal(n) * sc).astype(np.float32)
    return flat


_LP = {}


def _lp(P, seed=0):
    from chimera_amd.learned import LearnedProposal
    S = P.sampler
    if id(S) not in _LP:
        _LP[id(S)] = LearnedProposal(S, seed=seed)
    lp = _LP[id(S)]
    lp.disable()
    return lp


def test_network_output_matches_oracle(tern108):
    P, O = tern108
    lp = _lp(P)
    flat = _random_params(lp, 1)
    lp.set_params(flat)
    rng = np.random.default_rng(2)
    occ = rng.integers(0, P.nsp, (100, lp.nocc)).astype(np.uint8)
    _say(f"[test] MFMA scorer on {occ.shape} occupancy rows vs oracle")
    g_dev = lp.eval(occ).astype(np.float64)
    g_or, _ = LP.forward(flat, occ, P.nsp, scorer=True)
    err = np.abs(g_dev - g_or)
    assert np.median(err) < 1e-4 * max(1.0, np.abs(g_or).max())
    assert np.all(err <= 3e-3 * np.maximum(1.0, np.abs(g_or)))


def test_gradients_match_oracle(tern108):
    P, O = tern108
    lp = _lp(P)
    flat = _random_params(lp, 3)
    lp.set_params(flat)
    rng = np.random.default_rng(4)
    n = 64
    occ = rng.integers(0, P.nsp, (n, lp.nocc)).astype(np.uint8)
    y = rng.standard_normal(n).astype(np.float32) * 3
    w = (rng.random(n) < 0.8).astype(np.float32)
    _say("[test] trainer forward/backward vs oracle gradient")
    loss_d, g_d = lp.grad(occ, y, w)
    loss_o, g_o = LP.loss_and_grad(flat, occ, y.astype(np.float64), w.astype(np.float64), P.nsp)
    assert abs(loss_d - loss_o) <= 1e-3 * abs(loss_o) + 1e-6
    for name, (off, m) in lp.offsets().items():
        a, b = g_d[off:off + m].astype(np.float64), g_o[off:off + m]
        scale = np.abs(b).max() + 1e-12
        assert np.abs(a - b).max() <= 5e-3 * scale, name
        if m > 1:
            assert np.corrcoef(a, b)[0, 1] > 0.9999, name


def test_move_records_match_oracle(tern108):
    """Records of the uniform sampler: the proposal's occupancy before the move, the
    energy change nfit * dchi2, the rung and the outcome, against the oracle's run."""
    P, O = tern108
    S = P.sampler
    lp = _lp(P)
    nsteps = 12
    S.seed = 5
    S.set_recording(log_steps=nsteps)
    S.init_state(P.cfg0, P.cell.ideal)
    lp.record(nsteps * S.nrep)
    _say(f"[run] device {nsteps} steps with records")
    S.run(nsteps, exchange_every=0, log_every=10 ** 9)
    nv, occ, y, meta = lp.records()
    assert nv == nsteps * S.nrep
    R = OR.PlainSampler(O, nrep=S.nrep, seed=5)
    R.init_state(P.cfg0)
    st = _stencil(P)
    bad = []
    for t in range(nsteps):
        before = R.species.copy()
        c2_before = R.chi2.copy()
        R.step()
        for rung in range(S.nrep):                 # no exchange: slot == rung
            a, b, acc, c2, x, u = R.log[t][rung]
            idx = t * S.nrep + rung
            if a == b:
                assert meta[idx] == -1
                continue
            want, _, _ = LP.occupancy(before[rung], st, a, b)
            assert np.array_equal(occ[idx], want), (t, rung)
            assert meta[idx] == rung | (acc << 16)
            dE = (c2 - c2_before[rung]) * O.nfit
            if abs(y[idx] - dE) > 1e-5 * max(1.0, abs(dE)):
                bad.append((t, rung, float(y[idx]), dE, acc))
    np.testing.assert_allclose(S.chi2, R.chi2, rtol=1e-9)
    lp.record(0)
    assert not bad, bad[:12]


_TILT_W = np.array([[0, 1, -2], [1, 0, 3], [-2, 3, 0]])


@pytest.mark.parametrize("lift,tilt_w", [(False, None), (True, None), (True, _TILT_W)])
def test_learned_trajectory_matches_oracle(tern108, lift, tilt_w):
    """Device sampler with the learned proposal (eps 0.3 uniform mixture, tilt on, the two
    coldest rungs uniform; symmetric or lifted tilt, unlike-bond or other bond weights)
    move by move against the oracle's restatement with the same parameters: the
    speculative scoring (both outcomes of the pending step), the per-site layer-1 sums and
    their lag reproduce the proposal drawn from the state each step reaches.  A differing
    choice is allowed only at a near tie of the selecting uniform or of the acceptance;
    the comparison stops there."""
    P, O = tern108
    S = P.sampler
    lp = _lp(P)
    flat = _random_params(lp, 7, w3_scale=0.05)
    lp.set_params(flat)
    nsteps, exch, seed = 16, 5, 11
    S.seed = seed
    S.set_recording(log_steps=nsteps)
    S.init_state(P.cfg0, P.cell.ideal)
    lp.set_tilt_weights(LP.unlike_weights(P.nsp) if tilt_w is None else tilt_w)
    lp.enable(ncand=64, eps=0.3, lam=1.0, kappa=0.2, beta_max=10.0, lift=lift)
    _say(f"[run] device learned {nsteps} steps lift={lift} weights={tilt_w is not None}")
    S.run(nsteps, exchange_every=exch, log_every=10 ** 9)
    log = S.move_log(nsteps)
    lp.disable()
    lp.set_tilt_weights(LP.unlike_weights(P.nsp))
    R = OR.PlainSampler(O, nrep=S.nrep, seed=seed)
    R.init_state(P.cfg0)
    L = LP.LearnedSampler(R, flat, _stencil(P), ncand=64, eps=0.3, lam=1.0, kappa=0.2,
                          beta_max=10.0, lift=lift, tilt_w=tilt_w)
    _say("[run] oracle learned")
    L.run(nsteps, exchange_every=exch)
    margins = {(s, c): m for s, c, m in L.margins}
    learned_steps = 0
    for t in range(nsteps):
        by_rung = {int(r[3]): r for r in log[t]}
        for rung in range(S.nrep):
            a, b, acc, c2, x, u = R.log[t][rung]
            gdev = by_rung[rung]
            m = margins[(t + 1, rung)]
            learned_steps += m < 1.0
            if (int(gdev[0]), int(gdev[1])) != (a, b):
                assert m < 1e-5, f"step {t} rung {rung}: different pair, margin {m}"
                return
            if int(gdev[2]) != acc:
                assert abs(x + np.log(u)) < 1e-4, f"step {t} rung {rung}"
                return
    assert learned_steps > nsteps * S.nrep // 3
    assert np.array_equal(S.species, R.species)
    np.testing.assert_allclose(S.chi2, R.chi2, rtol=1e-9)


def _scratch_sums(P, lp, flat, sp):
    """Per-site layer-1 sums (2, natoms, 256) of one configuration from scratch, in units
    of 2^-16, with the oracle's fixed-point weights."""
    kp, off, _ = LP.layout(lp.nfeat)
    o, m = off["W1"]
    W = np.rint(LP.w1_fixed(LP.bf16(flat[o:o + m].reshape(kp, LP.HID))) * 65536.0).astype(np.int64)
    st = _stencil(P)
    nc, nsp = st.ncoord, P.nsp
    out = np.zeros((2, sp.size, LP.HID), np.int64)
    for c in range(sp.size):
        nb = st.neighbours(c)
        for role in (0, 1):
            f = [(role * nc + q) * nsp + int(sp[x]) for q, x in enumerate(nb)]
            f.append((2 * nc + role) * nsp + int(sp[c]))
            out[role, c] = W[f].sum(0)
    return out


def test_site_sums_follow_the_chain(tern108):
    """After a learned run with exchanges, every slot's per-site layer-1 sums kept by
    increments (each accepted swap folded in by the per-replica kernel) equal the sums from
    scratch at that slot's configuration, exactly."""
    P, O = tern108
    S = P.sampler
    lp = _lp(P)
    flat = _random_params(lp, 9, w3_scale=0.05)
    lp.set_params(flat)
    S.seed = 4
    S.init_state(P.cfg0, P.cell.ideal)
    lp.enable(ncand=64, eps=0.2, lam=1.0, kappa=0.3, beta_max=10.0, lift=True)
    _say("[run] device learned 60 steps, exchange every 7")
    S.run(60, exchange_every=7, log_every=10 ** 9)
    sums = lp.site_sums()
    lp.disable()
    rung_of = S._read(5, (S.nrep,), np.int32)
    sp_rung = S.species
    assert S.accepted_per_rung().sum() > 0
    for s in range(S.nrep):
        want = _scratch_sums(P, lp, flat, sp_rung[rung_of[s]])
        np.testing.assert_array_equal(sums[s].astype(np.int64), want, err_msg=f"slot {s}")


def test_training_reduces_loss(tern108):
    P, O = tern108
    S = P.sampler
    lp = _lp(P)
    lp.set_params(lp.initial_params(0))
    S.seed = 3
    S.init_state(P.cfg0, P.cell.ideal)
    lp.record(200 * S.nrep)
    _say("[run] 200 uniform steps with records, then training")
    S.run(200, exchange_every=25, log_every=10 ** 9)
    l0 = lp.train(iters=1, batch=256, lr=1e-3, seed=1)
    l1 = lp.train(iters=300, batch=256, lr=1e-3, seed=2)
    _say(f"[test] loss {l0:.4g} -> {l1:.4g}")
    assert np.isfinite(l1) and l1 < 0.5 * l0
    lp.record(0)


def test_training_with_null_move_records(tern108):
    """Records of null moves carry no occupancy (meta -1): the ring starts zeroed and the
    trainer gives such rows species 0 and weight 0."""
    P, O = tern108
    S = P.sampler
    lp = _lp(P)
    sp = np.zeros(P.cell.natoms, np.int32)
    sp[:4], sp[4:8] = 1, 2                       # most proposals find no partner
    try:
        S.init_state(sp, P.cell.ideal)
        lp.record(100 * S.nrep)
        _say("[run] 100 steps of a 100:4:4 composition with records, then training")
        S.run(100, exchange_every=25, log_every=10 ** 9)
        nv, occ, y, meta = lp.records()
        nulls = meta[:nv] == -1
        assert nulls.sum() > 0
        assert np.all(occ[:nv][nulls] == 0)
        loss = lp.train(iters=20, batch=256, lr=1e-3, seed=3)
        assert np.isfinite(loss)
    finally:
        lp.record(0)
        S.init_state(P.cfg0, P.cell.ideal)


# ---------------------------------------------------------------- the C4 workload size
@pytest.fixture(scope="module")
def crconi4000():
    """The 4000-site CrCoNi workload (SS + 3-/4-leg MS to 5.5 A, three edges); 4 replicas
    for the oracle comparison."""
    from chimera_amd import workloads as W
    from oracle.build import problem_from
    _say("[build] fcc_crconi_4000 nrep=4")
    P = W.Problem(W.WORKLOADS["fcc_crconi_4000"], nrep=4, seed=0)
    _say("[build] oracle problem")
    return P, problem_from(P)


def test_learned_4000_sites_256_candidates_matches_oracle(crconi4000):
    """At the learned proposal's own size (4000 sites, 256 candidates per replica and
    step, split-BF16 transform as in production) the device chain equals the oracle's
    learned sampler move by move for a few steps: proposals, Hastings ratios and
    decisions (a differing choice only at a near tie), then species and chi^2."""
    P, O = crconi4000
    S = P.sampler
    lp = _lp(P)
    flat = _random_params(lp, 3, w3_scale=0.05)
    lp.set_params(flat)
    nsteps, seed = 4, 21
    S.seed = seed
    S.precision = "bf16x3"
    try:
        S.set_recording(log_steps=nsteps)
        S.init_state(P.cfg0, P.cell.ideal)
        lp.enable(ncand=256, eps=0.2, lam=1.0, kappa=0.5, beta_max=1e300)
        _say(f"[run] device learned 4000 sites, 256 candidates, {nsteps} steps")
        S.run(nsteps, exchange_every=0, log_every=10 ** 9)
        log = S.move_log(nsteps)
        lp.disable()
        R = OR.PlainSampler(O, nrep=S.nrep, seed=seed)
        R.init_state(P.cfg0)
        L = LP.LearnedSampler(R, flat, _stencil(P), ncand=256, eps=0.2, lam=1.0, kappa=0.5)
        _say("[run] oracle learned 4000 sites")
        L.run(nsteps, exchange_every=0)
        margins = {(st, c): m for st, c, m in L.margins}
        learned = 0
        for t in range(nsteps):
            by_rung = {int(r[3]): r for r in log[t]}
            for rung in range(S.nrep):
                a, b, acc, c2, x, u = R.log[t][rung]
                gdev = by_rung[rung]
                m = margins[(t + 1, rung)]
                learned += m < 1.0
                if (int(gdev[0]), int(gdev[1])) != (a, b):
                    assert m < 1e-5, f"step {t} rung {rung}: different pair, margin {m}"
                    return
                if int(gdev[2]) != acc:
                    assert abs(x + np.log(u)) < 1e-4, f"step {t} rung {rung}"
                    return
        assert learned >= nsteps * S.nrep // 2
        assert np.array_equal(S.species, R.species)
        np.testing.assert_allclose(S.chi2, R.chi2, rtol=2e-6)
    finally:
        S.set_recording(0)
        S.precision = "f64"


def test_trained_network_gains_over_untrained_weights():
    """The 4000-site workload with 32 replicas (the learned proposal's configuration):
    after training on the uniform sampler's own move records, the network predicts the
    energy change of held-out records, and on a fixed seed the learned proposal accepts
    more often at the collection rung and reaches a larger alpha_1 ESS per step than the
    same proposal with the untrained weights (whose score is the same for every
    candidate, leaving only the first-shell tilt)."""
    import bench
    from chimera_amd import workloads as W
    from chimera_amd.learned import LearnedProposal
    _say("[build] fcc_crconi_4000 nrep=32")
    P = W.build("fcc_crconi_4000")
    S = P.sampler
    S.precision = "bf16x3"
    S.init_state(P.cfg0, P.cell.ideal)
    lp = LearnedProposal(S, seed=0)
    untrained = lp.get_params().copy()
    lp.record(10000 * S.nrep)
    _say("[run] 10000 uniform steps with records, training")
    S.run(10000, exchange_every=25, log_every=10 ** 9)
    lp.train(iters=1500, batch=1024, lr=1e-3, seed=1)
    lp.train(iters=1500, batch=1024, lr=3e-4, seed=2)
    nv, occ, y, meta = lp.records(60000)
    ok = meta >= 0
    g = lp.eval(occ[ok][-20000:])
    corr = float(np.corrcoef(g, np.clip(y[ok][-20000:], -30, 30))[0, 1])
    lp.record(0)
    trained = lp.get_params().copy()
    settings = dict(ncand=256, lam=1.0, kappa=1.0, eps=0.05, beta_max=3.0)
    res = {}
    for name, par in (("untrained", untrained), ("trained", trained)):
        lp.set_params(par)
        S.seed = 5
        _say(f"[run] ESS run, {name} weights")
        res[name] = bench.ess_run(P, S, 2000, 30000, 25, lambda: lp.enable(**settings))
    lp.disable()
    _say(f"[test] fit corr {corr:.3f}; " + "; ".join(
        f"{k}: ESS {v['alpha1_ess']} accept {v['accept_collect_rung']}" for k, v in res.items()))
    assert corr > 0.5
    assert res["trained"]["accept_collect_rung"] > res["untrained"]["accept_collect_rung"] + 0.05
    assert res["trained"]["alpha1_ess"] > res["untrained"]["alpha1_ess"]
    # the posterior is the same: alpha_1 means agree within the runs' scatter
    for k in res["trained"]["alpha1_mean"]:
        assert abs(res["trained"]["alpha1_mean"][k] - res["untrained"]["alpha1_mean"][k]) < 0.01
