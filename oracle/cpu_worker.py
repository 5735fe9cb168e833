"""One CPU-baseline worker (oracle; run by bench.py's CPU baseline only).

    python -m oracle.cpu_worker <problem.pkl> <seconds> <nrep> <seed> [<warmup steps>]

Loads the oracle Problem the bench pickled (its own file), runs the plain reference
sampler (PlainSampler: per-pair delta terms over all k, MS delta through the incidence
lists, full windowed transform of the trial chi, profiled chi^2, Metropolis, exchange
every 25 steps) on `nrep` replicas for about `seconds` after an untimed warm-up (200
steps by default), and prints one JSON line {"moves": replica moves, "seconds": wall time,
"steps": timed steps, "warmup": warm-up steps}.  Thread counts come from the
environment (OMP_NUM_THREADS / OPENBLAS_NUM_THREADS), set by the caller.
"""
from __future__ import annotations

import json
import pickle
import sys
import time


def main(argv):
    path, seconds, nrep, seed = argv[0], float(argv[1]), int(argv[2]), int(argv[3])
    warm = int(argv[4]) if len(argv) > 4 else 200
    from . import rmc as OR
    with open(path, "rb") as fh:
        prob, cfg0 = pickle.load(fh)
    S = OR.PlainSampler(prob, nrep=nrep, seed=seed)
    S.init_state(cfg0)
    t0 = time.perf_counter()
    S.run(max(1, warm), exchange_every=25)
    per_step = (time.perf_counter() - t0) / max(1, warm)
    n = max(1, int(seconds / max(per_step, 1e-6)))
    t0 = time.perf_counter()
    S.run(n, exchange_every=25)
    dt = time.perf_counter() - t0
    print(json.dumps({"moves": nrep * n, "seconds": dt, "steps": n, "warmup": max(1, warm)}),
          flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
