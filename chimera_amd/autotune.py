"""Launch-shape autotuner of the device sampler, with a per-problem cache.

The per-replica accept/propose/gather kernel splits each replica's gather over `nsplit`
workgroups: more workgroups fill more CUs but add redundant work (every workgroup of a
replica makes the same decision and proposal) and more integer atomics when the
workgroups merge their count changes.  The best split depends on the problem (sites,
path records, replicas), so it is measured: each candidate runs real sampler steps
timed with HIP events on the sampler's stream (chm_sampler_profile), and the fastest is
kept and cached under a key of the problem's shape and the device.

No counterpart in the reference (its launch shapes are fixed constants).
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import _lib

CANDIDATES = (1, 2, 4, 8, 12, 16, 24, 32, 48, 64)


def _cache_path():
    root = os.environ.get("CHM_TUNE_CACHE")
    if root:
        return root
    return os.path.join(os.path.expanduser("~"), ".cache", "chimera_amd", "tune.json")


def _key(S):
    ms = S.ms
    return "|".join(str(v) for v in (
        "v2", S.natoms, S.nsp, S.ncoord, S.nrow_e, S.nkw, S.nwin, S.nrep,
        0 if ms is None else ms.nstencil, S.precision,
        int(bool(getattr(S, "paired", False)) and S.precision == "bf16x3")))


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(cache):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def step_time_ms(S, iters=100):
    """Device time (ms) of one sampler step at the current launch shape."""
    ms = np.zeros(4)
    _lib.check(_lib.load().chm_sampler_profile(S._S, int(iters), ms))
    return float(ms[3])


def tune(S, candidates=CANDIDATES, iters=100, use_cache=True):
    """Pick the per-replica workgroup count of sampler S (its state must be initialised).
    Returns {"nsplit": best, "times_ms": {candidate: step ms}}."""
    lib = _lib.load()
    cache = _load() if use_cache else {}
    key = _key(S)
    if use_cache and key in cache:
        best = int(cache[key]["nsplit"])
        _lib.check(lib.chm_sampler_set_tuning(S._S, best))
        return cache[key]
    times = {}
    for c in candidates:
        _lib.check(lib.chm_sampler_set_tuning(S._S, int(c)))
        step_time_ms(S, 10)                      # warm the new shape
        times[int(c)] = step_time_ms(S, iters)
    best = min(times, key=times.get)
    _lib.check(lib.chm_sampler_set_tuning(S._S, best))
    rec = {"nsplit": int(best), "times_ms": {str(k): v for k, v in times.items()}}
    if use_cache:
        cache[key] = rec
        _save(cache)
    return rec


def current(S):
    v = C.c_int(0)
    _lib.check(_lib.load().chm_sampler_get_tuning(S._S, C.byref(v)))
    return int(v.value)
