"""Many independent RMC problems on one GPU, stepped together (no counterpart in the
reference, whose fit driver loops over problems and seeds on the host,
chimera/fitrun.py:302-333).

Each problem is an ordinary device sampler (its own supercell, edges, basis rows,
transform columns and replicas).  `Batch` steps all of them with ONE launch per kernel
per step (descriptor tables of problem x tile and problem x replica slot, built once)
and one captured hipGraph per exchange period.  Each problem's launch split (workgroups
per replica in the gather) is chosen by the autotuner and cached by problem shape.

`batch_specs` draws the uneven batch of the benchmark: fcc or bcc supercells of 256,
500, 864, 2048 or 4000 sites, a0 ~ U(2.85, 3.62) A, Fe-Ni or Cr-Co-Ni with 1..nsp
measured edges, series of 2-6 points with a0 x U(0.98, 1.02), 16 replicas each,
single scattering to the third shell plus the 3- and 4-leg paths within it.
"""
from __future__ import annotations

import ctypes as C
import sys
import time

import numpy as np

from . import _lib
from ._lib import check

CELLS = {
    "fcc": {256: (4, 4, 4), 500: (5, 5, 5), 864: (6, 6, 6), 2048: (8, 8, 8), 4000: (10, 10, 10)},
    "bcc": {256: (8, 4, 4), 500: (10, 5, 5), 864: (12, 6, 6), 2048: (16, 8, 8), 4000: (20, 10, 10)},
}
RMAX_OVER_A0 = {"fcc": 1.25, "bcc": 1.45}   # third neighbour shell
ALLOYS = (("Fe", "Ni"), ("Cr", "Co", "Ni"))


def batch_specs(n=64, seed=2024, nrep=16):
    """The workloads of an uneven batch of n problems (series of 2-6 related points)."""
    from .workloads import Workload
    rng = np.random.default_rng(seed)
    out = []
    series = 0
    while len(out) < n:
        npts = min(int(rng.integers(2, 7)), n - len(out))
        lat = ("fcc", "bcc")[int(rng.integers(0, 2))]
        natoms = (256, 500, 864, 2048, 4000)[int(rng.integers(0, 5))]
        els = ALLOYS[int(rng.integers(0, 2))]
        nedge = int(rng.integers(1, len(els) + 1))
        edges = tuple(sorted(int(e) for e in rng.choice(len(els), nedge, replace=False)))
        a0 = float(rng.uniform(2.85, 3.62))
        alpha = float(rng.uniform(-0.12, -0.02))
        for k in range(npts):
            a = a0 * float(rng.uniform(0.98, 1.02))
            idx = len(out)
            out.append(Workload(f"batch{idx:02d}_s{series}_{lat}{natoms}", lat, CELLS[lat][natoms], a,
                                els, RMAX_OVER_A0[lat] * a, True, nrep, alpha_target=alpha,
                                seed=1000 + idx, edges=edges))
        series += 1
    return out


ITEMS_PER_LANE = (2, 4, 8, 16, 32, 64)   # gather items per lane: the batch tuner's candidates


def _items(S):
    """Gather items of one swap: two neighbour lists and two path-record stencils."""
    return 2 * S.ncoord + 2 * (0 if S.ms is None else S.ms.nstencil)


class Batch:
    """Step many problems (workloads.Problem objects) together on the device.

    tune=True picks, by measurement on the live batch (HIP events; cached by the batch's
    shape), two launch shapes together: the path-list partitioning (every problem's
    replicas split their gather over ceil(items / (256 * ipl)) workgroups, items per lane
    ipl) and the transform's tile shape (K blocks each wavefront holds: 1 when every
    problem's rows fit four K blocks, or 4)."""

    def __init__(self, problems, precision="bf16x3", reanchor_every=512, tune=True, log=None):
        self.P = list(problems)
        self.ctx = _lib.Context.default()
        self.h = None
        self.tuned = None
        for P in self.P:
            S = P.sampler
            S.precision, S.reanchor_every = precision, reanchor_every
            S.init_state(P.cfg0, P.cell.ideal)
        if tune:
            self.tuned = self._tune(log)
        self._create()
        if log:
            for P in self.P:
                log(f"[batch] {P.w.name}: {P.cell.natoms} sites, {P.nsp} species, "
                    f"{P.sampler.nrow_e} rows/edge, nsplit {self._split(P.sampler)}")

    def _split(self, S):
        from . import autotune
        return autotune.current(S)

    def _set_ipl(self, ipl):
        for P in self.P:
            S = P.sampler
            n = max(1, min(64, -(-_items(S) // (256 * ipl))))
            check(_lib.load().chm_sampler_set_tuning(S._S, n))

    def _create(self, kb=None):
        self.close()
        handles = (C.c_void_p * len(self.P))(*[P.sampler._S.value for P in self.P])
        self._handles = handles
        h = _lib.vp()
        check(_lib.load().chm_batch_create(self.ctx.h, len(self.P), C.cast(handles, C.c_void_p),
                                           C.byref(h)))
        self.h = h
        if kb is None:
            kb = (self.tuned or {}).get("kb_per_wave", 0)
        check(_lib.load().chm_batch_set_transform(self.h, int(kb)))

    def _max_nkb(self):
        n = C.c_int(0)
        check(_lib.load().chm_batch_get_shape(self.h, C.byref(n)))
        return int(n.value)

    def _tune(self, log=None, iters=20):
        from . import autotune
        key = "batch|" + "|".join(autotune._key(P.sampler) for P in self.P)
        cache = autotune._load()
        if key in cache and "kb_per_wave" in cache[key]:
            ipl = int(cache[key]["items_per_lane"])
            self._set_ipl(ipl)
            return cache[key]
        times = {}
        self._create(0)
        kbs = (1, 4) if self._max_nkb() <= 4 else (4,)
        for kb in kbs:
            for ipl in ITEMS_PER_LANE:
                self._set_ipl(ipl)
                self._create(kb)
                times[(kb, ipl)] = self.profile(iters)["step_ms"]
                if log:
                    log(f"[batch] tune: {kb} K blocks/wave, {ipl} items/lane -> "
                        f"{times[(kb, ipl)]:.4f} ms/step")
        kb_best, best = min(times, key=times.get)
        self._set_ipl(best)
        self.close()
        for P in self.P:                          # the timing ran steps: start again
            P.sampler.init_state(P.cfg0, P.cell.ideal)
        rec = {"items_per_lane": int(best), "kb_per_wave": int(kb_best),
               "step_ms": {f"kb{k}_ipl{i}": v for (k, i), v in times.items()}}
        cache[key] = rec
        autotune._save(cache)
        return rec

    @property
    def replicas(self):
        return sum(P.sampler.nrep for P in self.P)

    def run(self, nsteps, exchange_every=25):
        """Run every problem nsteps steps; returns the wall time (s)."""
        t0 = time.perf_counter()
        check(_lib.load().chm_batch_run(self.h, int(nsteps), int(exchange_every)))
        return time.perf_counter() - t0

    def profile(self, iters=50):
        """Device ms of the batched transform, per-replica kernel and whole step."""
        ms = np.zeros(3)
        check(_lib.load().chm_batch_profile(self.h, int(iters), ms.ctypes.data))
        return {"transform_ms": float(ms[0]), "slot_step_ms": float(ms[1]), "step_ms": float(ms[2])}

    def close(self):
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib._lib.chm_batch_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return
        self.close()


def build(specs, log=None):
    from . import workloads as W
    out = []
    for w in specs:
        t0 = time.perf_counter()
        out.append(W.Problem(w))
        if log:
            log(f"[batch] built {w.name} in {time.perf_counter() - t0:.1f} s")
    return out
