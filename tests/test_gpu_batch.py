"""Batched stepping of uneven problems (one launch per kernel per step for all of them)
against the same problems run one by one: identical trajectories."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def problems():
    from chimera_amd import workloads as W
    specs = [
        W.Workload("bt_fcc256", "fcc", (4, 4, 4), 3.30, ("Cr", "Co", "Ni"), 1.25 * 3.30, True, 16,
                   seed=11, edges=(0, 2)),
        W.Workload("bt_bcc256", "bcc", (8, 4, 4), 2.95, ("Fe", "Ni"), 1.45 * 2.95, True, 8,
                   seed=12, edges=(1,)),
        W.Workload("bt_fcc500", "fcc", (5, 5, 5), 3.55, ("Cr", "Co", "Ni"), 1.25 * 3.55, True, 16,
                   seed=13),
    ]
    out = []
    for w in specs:
        _say(f"[build] {w.name}")
        out.append(W.Problem(w))
    return out


@pytest.mark.parametrize("precision", ["f64", "bf16x3"])
def test_batch_matches_single_runs(problems, precision):
    from chimera_amd.batch import Batch
    nsteps, exch = 40, 7
    singles = []
    for P in problems:
        S = P.sampler
        S.precision = precision
        S.set_recording(log_steps=nsteps)
        S.init_state(P.cfg0, P.cell.ideal)
        _say(f"[run] single {P.w.name} {precision}")
        S.run(nsteps, exchange_every=exch, log_every=10 ** 9)
        singles.append((S.move_log(nsteps).copy(), S.species.copy(), S.chi2.copy(),
                        S.counts.copy()))
    B = Batch(problems, precision=precision, tune=False)
    try:
        _say(f"[run] batch of {len(problems)} ({B.replicas} replicas) {precision}")
        B.run(nsteps, exchange_every=exch)
        for P, (log, sp, c2, cn) in zip(problems, singles):
            S = P.sampler
            assert np.array_equal(S.move_log(nsteps), log), P.w.name
            assert np.array_equal(S.species, sp), P.w.name
            assert np.array_equal(S.counts, cn), P.w.name
            np.testing.assert_array_equal(S.chi2, c2)
            # the batched chain is a valid chain: chi^2 recomputed from scratch agrees
            np.testing.assert_allclose(S.chi2, S.chi2_of(S.species), rtol=2e-6 if precision != "f64" else 1e-9)
    finally:
        B.close()
        for P in problems:
            P.sampler.precision = "f64"
            P.sampler.set_recording(0)


def test_edges_subset_weights(problems):
    """A species that is not a measured edge contributes no chi^2 (zero-weight channels)."""
    P = problems[1]                      # Fe-Ni, only Ni measured
    S = P.sampler
    S.init_state(P.cfg0, P.cell.ideal)
    assert np.all(P.sigma[P.chan_edge == 0] > 1e20)
    assert np.all(P.sigma[P.chan_edge == 1] < 1e3)


@pytest.fixture(scope="module")
def uneven():
    """One problem of every shape in the benchmark's 64-problem batch (fcc 256 / 500 /
    864 / 2048 / 4000 and bcc 256 / 500 / 864 / 4000 sites), with the batch's own
    lattice constants, alloys and measured edges."""
    from chimera_amd import workloads as W
    from chimera_amd.batch import batch_specs
    specs, seen = [], set()
    for w in batch_specs(64):
        if (w.lattice, w.ncell) not in seen:
            seen.add((w.lattice, w.ncell))
            specs.append(w)
    out = []
    for w in specs:
        _say(f"[build] {w.name}")
        out.append(W.Problem(w))
    return out


def test_uneven_batch_of_the_benchmark_matches_single_runs(uneven):
    """The batched split-BF16 chains of every problem shape of the 64-problem batch are
    the single runs move by move; counts equal a recount and chi^2 a from-scratch
    recompute for every problem."""
    from chimera_amd.batch import Batch
    nsteps, exch = 30, 7
    sizes = sorted({int(P.cell.natoms) for P in uneven})
    assert {864, 2048, 4000} <= set(sizes)
    singles = []
    for P in uneven:
        S = P.sampler
        S.precision = "bf16x3"
        S.set_recording(log_steps=nsteps)
        S.init_state(P.cfg0, P.cell.ideal)
        _say(f"[run] single {P.w.name}")
        S.run(nsteps, exchange_every=exch, log_every=10 ** 9)
        singles.append((S.move_log(nsteps).copy(), S.species.copy(), S.chi2.copy(), S.counts.copy()))
    B = Batch(uneven, precision="bf16x3", tune=False)
    try:
        _say(f"[run] batch of {len(uneven)} ({B.replicas} replicas, sizes {sizes})")
        B.run(nsteps, exchange_every=exch)
        for P, (log, sp, c2, cn) in zip(uneven, singles):
            S = P.sampler
            assert np.array_equal(S.move_log(nsteps), log), P.w.name
            assert np.array_equal(S.species, sp), P.w.name
            assert np.array_equal(S.counts, cn), P.w.name
            np.testing.assert_array_equal(S.chi2, c2)
            assert np.array_equal(S.counts, S.counts_of(S.species)), P.w.name
            np.testing.assert_allclose(S.chi2, S.chi2_of(S.species), rtol=2e-6)
    finally:
        B.close()
        for P in uneven:
            P.sampler.precision = "f64"
            P.sampler.set_recording(0)
