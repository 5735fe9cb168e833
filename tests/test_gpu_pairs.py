"""Paired steps against single steps.  In the split-BF16 mode the sampler makes two
Metropolis steps per transform + per-replica launch: the transform computes the trial of
step n+1 and the trials of step n+2 after either outcome of step n+1 (pre-fetching), and
the per-replica launch decides both in order.  The chain must be the one made step by
step, bit for bit: move log, species, integer counts, chi^2, best chi^2 and best
configuration, the running chi(R), the pair-count series, acceptance and re-anchor
counts."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(name, lat, nc, a0, els, rmax, ms, nrep, seed=0):
    from chimera_amd import workloads as W
    _say(f"[build] {name} {nc} ms={ms} nrep={nrep}")
    return W.Problem(W.Workload(name, lat, nc, a0, els, rmax, ms, nrep), seed=seed)


@pytest.fixture(scope="module")
def tern108():
    return _problem("t108", "fcc", (3, 3, 3), 3.59, ("Cr", "Co", "Ni"), 5.2, True, 6)


@pytest.fixture(scope="module")
def tern108_40():
    # more replica slots than one 32-slot transform row block (the reference's default is 48)
    return _problem("t108r40", "fcc", (3, 3, 3), 3.59, ("Cr", "Co", "Ni"), 5.2, True, 40)


@pytest.fixture(scope="module")
def bin500():
    return _problem("b500", "fcc", (5, 5, 5), 3.59, ("Fe", "Ni"), 5.2, False, 8)


def _run(P, paired, nsteps, exch, seed, rean=4, series=False, species0=None):
    S = P.sampler
    S.precision, S.reanchor_every, S.paired = "bf16x3", rean, paired
    S.seed = seed
    S.set_recording(log_steps=nsteps, series_steps=nsteps if series else 0)
    S.init_state(P.cfg0 if species0 is None else species0, P.cell.ideal)
    _say(f"[run] paired={paired} {nsteps} steps exchange_every={exch} seed={seed}")
    S.run(nsteps, exchange_every=exch, log_every=10 ** 9)
    out = {"log": S.move_log(nsteps), "species": S.species, "chi2": S.chi2,
           "best": S.best_chi2, "best_sp": S._read(4, (S.nrep, S.natoms), np.int32),
           "counts": S.counts, "acc": S.accepted_per_rung(), "mcur": S.running_chiR(),
           "reanchors": S.reanchors()}
    if series:
        out["series"] = S.pair_series(nsteps)
    return out


def _restore(P):
    S = P.sampler
    S.precision, S.reanchor_every, S.paired = "f64", 512, True
    S.set_recording(0, 0)
    S.init_state(P.cfg0, P.cell.ideal)


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


# (steps, exchange period, seed): odd and even periods (segments of 1, 3..25 steps, a
# single step closing odd segments), no exchange (one long segment)
CASES = [(61, 25, 3), (40, 6, 5), (33, 7, 11), (20, 0, 2), (12, 3, 8)]


@pytest.mark.parametrize("nsteps,exch,seed", CASES)
def test_pairs_equal_single_steps(tern108, nsteps, exch, seed):
    P = tern108
    try:
        single = _run(P, False, nsteps, exch, seed, series=True)
        paired = _run(P, True, nsteps, exch, seed, series=True)
        assert single["acc"].sum() > 0
        if exch:   # re-anchoring happens at exchange-period starts
            assert single["reanchors"].sum() > 0
        _same(single, paired)
    finally:
        _restore(P)


def test_pairs_two_row_blocks(tern108_40):
    """40 replicas: two row blocks in every transform and in the draws column."""
    P = tern108_40
    try:
        single = _run(P, False, 33, 7, 12, series=True)
        paired = _run(P, True, 33, 7, 12, series=True)
        assert single["acc"].sum() > 0
        _same(single, paired)
    finally:
        _restore(P)


def test_pairs_single_scattering_binary(bin500):
    """No path records (single scattering only): the gather is the neighbour stencil."""
    P = bin500
    try:
        _same(_run(P, False, 30, 5, 4), _run(P, True, 30, 5, 4))
    finally:
        _restore(P)


def test_pairs_with_null_moves(bin500):
    """A composition where most proposals find no partner: null moves in either step of
    a pair, and second-step proposals that differ between the two outcomes."""
    P = bin500
    sp = np.zeros(P.cell.natoms, np.int32)
    sp[:6] = 1
    try:
        _same(_run(P, False, 24, 0, 1, species0=sp), _run(P, True, 24, 0, 1, species0=sp))
    finally:
        _restore(P)


def test_pairs_launch_split_invariance(tern108):
    from chimera_amd import _lib
    P = tern108
    try:
        ref = _run(P, True, 30, 7, 6)
        for ns in (1, 5):
            S = P.sampler
            S.precision, S.reanchor_every, S.paired = "bf16x3", 4, True
            S.seed = 6
            S.set_recording(log_steps=30)
            S.init_state(P.cfg0, P.cell.ideal)
            _lib.check(_lib.load().chm_sampler_set_tuning(S._S, ns))
            S.run(30, exchange_every=7, log_every=10 ** 9)
            assert np.array_equal(S.move_log(30), ref["log"]) and np.array_equal(S.counts, ref["counts"])
    finally:
        _restore(P)


def test_pairs_workload_4000():
    """Full-size workload (4000 atoms, 3 edges with multiple scattering, 32 replicas):
    paired and single steps agree bit for bit over 120 steps with exchanges."""
    from chimera_amd import workloads as W
    _say("[build] fcc_crconi_4000")
    P = W.build("fcc_crconi_4000")
    S = P.sampler
    S.reanchor_every = 16
    outs = []
    for paired in (False, True):
        S.precision, S.paired = "bf16x3", paired
        S.seed = 21
        S.set_recording(log_steps=120)
        S.init_state(P.cfg0, P.cell.ideal)
        _say(f"[run] 4000 atoms paired={paired}")
        S.run(120, exchange_every=25, log_every=10 ** 9)
        outs.append((S.move_log(120), S.species, S.counts, S.chi2, S.best_chi2, S.running_chiR()))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    assert np.array_equal(S.counts, S.counts_of(S.species))
