"""Failure handling and resumption of the device sampler, forced from the host: a
checkpointed chain resumes bit for bit, a corrupted running chi(R) is repaired by the FP64
re-anchor, NaN tables make every move rejected without touching the state, and the run
history taken on the device equals the one read with host round trips."""
import sys

import numpy as np
import pytest

from chimera_amd import _lib
from chimera_amd.method_a_rmc import state_sections

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(name="t108", nrep=6, ms=True):
    from chimera_amd import workloads as W
    _say(f"[build] {name} nrep={nrep} ms={ms}")
    return W.Problem(W.Workload(name, "fcc", (3, 3, 3), 3.59, ("Cr", "Co", "Ni"), 5.2, ms, nrep))


@pytest.fixture(scope="module")
def tern():
    return _problem()


def _state(S):
    return (S.species.copy(), S.chi2.copy(), S.best_chi2.copy(), S.counts.copy(),
            S.running_chiR().copy(), S.accepted_per_rung().copy(), S.reanchors().copy(),
            S._read(8, (2, S.nrep - 1), np.int64).copy(), S.beta.copy())


@pytest.mark.parametrize("prec,paired", [("f64", False), ("bf16x3", True), ("bf16x3", False)])
def test_checkpoint_resume_bit_exact(tern, prec, paired):
    """Save mid-run (mid exchange period), continue; restore the blob into a freshly
    initialised sampler (and into a second sampler of the same problem) and continue
    the same number of steps: species, counts, chi^2, best, running chi(R), counters
    and ladder are identical, and so is the move log of the resumed segment."""
    P = tern
    S = P.sampler
    S.precision, S.paired, S.reanchor_every = prec, paired, 3
    try:
        _say(f"[test] checkpoint {prec} paired={paired}")
        S.init_state(P.cfg0, P.cell.ideal)
        S.run(37, exchange_every=10, log_every=10 ** 9)
        blob = S.save_state()
        S.set_recording(log_steps=41)
        S.run(41, exchange_every=10, log_every=10 ** 9)
        want, want_log = _state(S), S.move_log(41).copy()
        # resume in the same sampler after a fresh start
        S.init_state(P.cfg0, P.cell.ideal)
        S.run(5, exchange_every=10, log_every=10 ** 9)        # unrelated progress, overwritten
        S.load_state(blob)
        S.set_recording(log_steps=41)
        S.run(41, exchange_every=10, log_every=10 ** 9)
        got, got_log = _state(S), S.move_log(41)
        for w, g in zip(want, got):
            assert np.array_equal(w, g)
        assert np.array_equal(want_log, got_log)
        assert S.accepted_per_rung().sum() > 0
    finally:
        S.set_recording(0, 0)
        S.precision, S.paired, S.reanchor_every = "f64", True, 512
        S.init_state(P.cfg0, P.cell.ideal)


def test_checkpoint_into_new_sampler_and_shape_checks(tern):
    """A blob restores into a second sampler built from the same problem description;
    a blob of another shape or precision is refused with ValueError."""
    P = tern
    S = P.sampler
    S.init_state(P.cfg0, P.cell.ideal)
    S.run(23, exchange_every=10, log_every=10 ** 9)
    blob = S.save_state()
    S.run(30, exchange_every=10, log_every=10 ** 9)
    want = _state(S)
    Q = _problem()
    T = Q.sampler
    T.init_state(P.cfg0, P.cell.ideal)
    T.load_state(blob)
    T.run(30, exchange_every=10, log_every=10 ** 9)
    for w, g in zip(want, _state(T)):
        assert np.array_equal(w, g)
    T.precision = "bf16x3"
    T.init_state(P.cfg0, P.cell.ideal)
    with pytest.raises(ValueError):
        T.load_state(blob)                      # saved in the exact FP64 mode
    T.precision = "f64"
    R = _problem("t108r4", nrep=4)
    with pytest.raises(ValueError):
        R.sampler.load_state(blob)              # 6 replicas into 4
    bad = bytearray(blob)
    sec = state_sections(bad)
    off, _ = sec["species"]
    bad[off] = 7                                # species index out of range
    with pytest.raises(ValueError):
        T.load_state(bytes(bad))


def test_reanchor_repairs_corrupted_accumulator(tern):
    """Corrupt the running chi(R) from the host (through a checkpoint blob); every
    replica's FP64 re-anchor, due after reanchor_every accepted moves, restores it to
    the exact transform of its integer counts."""
    P = tern
    S = P.sampler
    S.precision, S.paired, S.reanchor_every = "bf16x3", True, 2
    try:
        S.init_state(P.cfg0, P.cell.ideal)
        S.run(20, exchange_every=10, log_every=10 ** 9)
        exact = S.exact_chiR()
        assert np.abs(S.running_chiR() - exact).max() < 1e-7 * np.abs(exact).max()
        blob = bytearray(S.save_state())
        off, nb = state_sections(blob)["running_chiR"]
        m = np.frombuffer(bytes(blob[off:off + nb]), np.float64).copy()
        rng = np.random.default_rng(3)
        m += rng.normal(0, 0.05 * np.abs(m).max(), m.shape)
        blob[off:off + nb] = m.tobytes()
        S.load_state(bytes(blob))
        assert np.abs(S.running_chiR() - S.exact_chiR()).max() > 1e-3 * np.abs(exact).max()
        n0 = S.reanchors().copy()
        acc0 = S.accepted_per_rung().sum()
        _say("[test] corrupted accumulator: running until every replica re-anchored")
        for _ in range(40):
            S.run(20, exchange_every=10, log_every=10 ** 9)
            if np.all(S.reanchors() > n0 + 1):
                break
        assert np.all(S.reanchors() > n0 + 1), "a replica accepted too few moves to re-anchor"
        assert S.accepted_per_rung().sum() > acc0
        ex = S.exact_chiR()
        assert np.abs(S.running_chiR() - ex).max() < 1e-7 * np.abs(ex).max()
        np.testing.assert_allclose(S.chi2, S.chi2_of(S.species), rtol=1e-6)
    finally:
        S.precision, S.paired, S.reanchor_every = "f64", True, 512
        S.init_state(P.cfg0, P.cell.ideal)


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_nan_tables_reject_every_move(prec):
    """A NaN in the scattering tables makes every chi^2 NaN: every swap and every
    replica exchange is rejected (the reference's `u < exp(-min(x, 60))` is False for NaN,
    methods/method_a_rmc.py:280-285, 349-350) and the configuration and integer counts are
    untouched; with the tables restored the same sampler moves again."""
    P = _problem("tnan", nrep=6)
    S, fm = P.sampler, P.fm
    S.precision = prec
    lib = _lib.load()
    f = fm.f.copy()
    f[1, 150] = np.nan                           # one k point of one scatterer
    _lib.check(lib.chm_fm_set_tables(fm.ctx.h, fm.nsp, fm.nk, fm.k, f,
                                     _lib.f64(fm.delta.reshape(fm.nsp * fm.nsp, -1)), fm.lam,
                                     fm.s02))
    S.set_target(P.target, P.sigma)               # rebuilds the basis from the tables
    S.init_state(P.cfg0, P.cell.ideal)
    assert np.all(np.isnan(S.chi2))
    _say(f"[test] NaN tables, {prec}")
    S.run(60, exchange_every=10, log_every=10 ** 9)
    assert S.accepted_per_rung().sum() == 0
    ea = S._read(8, (2, S.nrep - 1), np.int64)
    assert ea[0].sum() > 0 and ea[1].sum() == 0
    assert np.array_equal(S.species, np.repeat(P.cfg0[None], S.nrep, 0))
    assert np.array_equal(S.counts, S.counts_of(S.species))
    # tables restored: the chain moves, and its counts stay exact
    _lib.check(lib.chm_fm_set_tables(fm.ctx.h, fm.nsp, fm.nk, fm.k, fm.f,
                                     _lib.f64(fm.delta.reshape(fm.nsp * fm.nsp, -1)), fm.lam,
                                     fm.s02))
    S.set_target(P.target, P.sigma)
    S.init_state(P.cfg0, P.cell.ideal)
    S.run(60, exchange_every=10, log_every=10 ** 9)
    assert S.accepted_per_rung().sum() > 0
    assert np.array_equal(S.counts, S.counts_of(S.species))
    assert np.all(np.isfinite(S.chi2))


def test_history_from_device_marks(tern):
    """run()'s history from device snapshots (no callback) equals the one read with host
    round trips (callback), and best_verified equals chi2_of(best)."""
    P = tern
    S = P.sampler
    S.precision, S.paired = "bf16x3", True
    try:
        seen = []
        S.init_state(P.cfg0, P.cell.ideal)
        h_sync = S.run(57, exchange_every=10, log_every=20,
                       callback=lambda st, b, a: seen.append((st, b, a)))
        sp_sync = S.species
        S.init_state(P.cfg0, P.cell.ideal)
        h_dev = S.run(57, exchange_every=10, log_every=20)
        assert [h[0] for h in h_dev] == [1, 20, 40] == [h[0] for h in h_sync]
        assert [h[0] for h in seen] == [1, 20, 40]
        for a, b in zip(h_sync, h_dev):
            assert a[1] == b[1] and a[2] == b[2]
        times = [h[3] for h in h_dev]
        assert all(t >= 0 for t in times) and times == sorted(times)
        assert np.array_equal(S.species, sp_sync)
        bsp, _, bc2 = S.best()
        want = S.chi2_of(bsp[None])[0]
        np.testing.assert_allclose(want, bc2, rtol=1e-6)
        # best_verified is the value at the end of the run (method_a_rmc.py:439-448), also
        # when it is first read after later steps
        for _ in range(3):
            S.step()
        np.testing.assert_allclose(S.best_verified, want, rtol=1e-12)
    finally:
        S.precision, S.paired = "f64", True
        S.init_state(P.cfg0, P.cell.ideal)


@pytest.mark.parametrize("paired,nsplit", [(True, 24), (False, 16)])
def test_best_state_consistent_with_many_workgroups(paired, nsplit):
    """At 4000 sites with more per-replica workgroups than stay resident together (the
    launch's later workgroups start after earlier ones, the lead among them, have
    finished), every replica's best configuration still has the best chi^2 it is stored
    with: the workgroups decide 'new best' from the copy the trial transform made, which
    the launch's lead does not overwrite."""
    from chimera_amd import workloads as W
    _say(f"[build] fcc_crconi_4000 paired={paired} nsplit={nsplit}")
    P = W.build("fcc_crconi_4000")
    S = P.sampler
    S.precision, S.paired = "bf16x3", paired
    S.init_state(P.cfg0, P.cell.ideal)
    _lib.check(_lib.load().chm_sampler_set_tuning(S._S, nsplit))
    S.run(1500, exchange_every=25, log_every=10 ** 9)
    best_sp = S._read(4, (S.nrep, S.natoms), np.int32)
    rec = S.chi2_of(best_sp)
    # the running chi(R) is split-BF16 with an FP64 accumulator: a recompute agrees to ~1e-7
    assert np.all(np.abs(rec - S.best_chi2) <= 1e-5 * S.best_chi2)
