"""Precision study of the device sampler (no counterpart in the reference, whose
transform is an FP64 FFT of the full trial chi every step).

For one workload it reports, against the exact FP64 path (FP64 pair-space path sum;
chi(R) recomputed in FP64 from the integer counts; the FP64 path is checked against the
FP64 NumPy oracle by the test suite):
  * chi(k) of the FP32 path sum (FP32 terms on the hardware sine / exponential units,
    summed in FP32 per run of one absorber species and in FP64 across runs);
  * Re/Im chi(R) of the running accumulator after a run, chi^2 and the R-factor computed
    from it, for split-BF16 and FP32 transforms, FP64 and FP32 accumulators, with and
    without the FP64 re-anchor;
  * alpha_1..alpha_4 (means at the likelihood-temperature rung) and the fraction of
    acceptance decisions equal to the exact chain's, same seed.

    python -c "from chimera_amd import precision_study as p; p.main()"   # or bench.py --precision-study
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np

MODES = [
    # (label, precision, accumulator, reanchor_every)
    ("f64 exact", "f64", "f64", 0),
    ("bf16x3, f64 acc, re-anchor 512", "bf16x3", "f64", 512),
    ("bf16x3, f64 acc, no re-anchor", "bf16x3", "f64", 0),
    ("bf16x3, f32 acc, re-anchor 512", "bf16x3", "f32", 512),
    ("bf16x3, f32 acc, no re-anchor", "bf16x3", "f32", 0),
    ("f32, f64 acc, re-anchor 512", "f32", "f64", 512),
    ("f32, f32 acc, no re-anchor", "f32", "f32", 0),
]


def _err(a, b):
    d = np.abs(a - b)
    return {"max_abs": float(d.max()), "max_rel": float(d.max() / max(np.abs(b).max(), 1e-300))}


def chik_errors(P, nconf=4, seed=0):
    """FP32 vs FP64 pair-space path sum on random configurations of the workload."""
    from .lattice import random_config
    comp = np.ones(P.nsp) / P.nsp
    sp = np.stack([random_config(P.cell.natoms, comp, seed + i) for i in range(nconf)])
    pos = np.repeat(P.cell.ideal[None], nconf, 0)
    c64 = P.fm.chi(pos, sp, normalise=False, precision="f64")
    c32 = P.fm.chi(pos, sp, normalise=False, precision="f32")
    return _err(c32, c64)


def run_mode(P, label, precision, accumulator, reanchor, steps, seed, exchange_every=25):
    S = P.sampler
    S.precision, S.accumulator, S.reanchor_every = precision, accumulator, reanchor
    S.seed = seed
    S.set_recording(log_steps=steps, series_steps=steps)
    S.init_state(P.cfg0, P.cell.ideal)
    t0 = time.perf_counter()
    S.run(steps, exchange_every=exchange_every, log_every=10 ** 9)
    P.fm.ctx.sync()
    dt = time.perf_counter() - t0
    out = {"mode": label, "steps": steps, "us_per_step": round(1e6 * dt / steps, 3)}
    exact = S.exact_chiR()
    if precision != "f64":
        run = S.running_chiR()
        out["chiR_re"] = _err(run[:, 0::2], exact[:, 0::2])
        out["chiR_im"] = _err(run[:, 1::2], exact[:, 1::2])
        out["reanchors"] = int(S.reanchors().sum())
    # chi^2 the chain carries vs a from-scratch FP64 recompute of its states
    sp = S.species
    c2_exact = S.chi2_of(sp)
    out["chi2"] = _err(S.chi2, c2_exact)
    # R-factor of the collection rung from the chi(R) the chain carries vs from scratch
    from .workloads import fit_scales, r_factor
    rung_of = S._read(5, (S.nrep,), np.int32)
    cr = int(S.collect_rungs()[0])
    slot = int(np.nonzero(rung_of == cr)[0][0])
    m_chain = (S.running_chiR() if precision != "f64" else exact)[slot].reshape(P.nchan, -1, 2)
    m_exact = exact[slot].reshape(P.nchan, -1, 2)
    rf = []
    for m in (m_chain, m_exact):
        z = m[..., 0] + 1j * m[..., 1]
        rf.append(r_factor(z, P.target[:, P.mask], np.ones(z.shape[1], bool),
                           fit_scales(z, P.target[:, P.mask], np.ones(z.shape[1], bool), P.chan_edge)))
    out["r_factor"] = {"chain": rf[0], "exact": rf[1], "abs_err": abs(rf[0] - rf[1])}
    al = np.stack([S.alpha_series(steps, shell=s) for s in range(min(4, S.nshell))], 0)
    iu = np.triu_indices(P.nsp, 1)
    out["alpha"] = [[round(float(al[s, :, a, b].mean()), 5) for a, b in zip(*iu)]
                    for s in range(al.shape[0])]
    out["_log"] = S.move_log(steps)
    return out


def study(workload="fcc_crconi_4000", steps=20000, seed=7, modes=MODES, problem=None):
    from . import workloads as W
    P = W.build(workload) if problem is None else problem
    saved = (P.sampler.precision, P.sampler.accumulator, P.sampler.reanchor_every, P.sampler.seed)
    res = {"workload": workload, "chik_f32_vs_f64": chik_errors(P)}
    rows = []
    ref_log = None
    for label, prec, accm, rean in modes:
        print(f"[precision] {label}", file=sys.stderr, flush=True)
        r = run_mode(P, label, prec, accm, rean, steps, seed)
        log = r.pop("_log")
        if ref_log is None:
            ref_log = log
        by = lambda L: {(t, int(x[3])): int(x[2]) for t in range(L.shape[0]) for x in L[t]}
        a, b = by(ref_log), by(log)
        r["decisions_equal_to_exact"] = float(np.mean([a[k] == b.get(k, -1) for k in a]))
        # alpha_1..alpha_4 means against the exact chain's (same seed and start)
        r["alpha_abs_err_vs_exact"] = float(np.abs(np.asarray(r["alpha"]) -
                                                   np.asarray(rows[0]["alpha"] if rows else r["alpha"])).max())
        rows.append(r)
    res["modes"] = rows
    P.sampler.precision, P.sampler.accumulator, P.sampler.reanchor_every, P.sampler.seed = saved
    P.sampler.set_recording(0, 0)
    P.sampler.init_state(P.cfg0, P.cell.ideal)
    return res


def table(res):
    lines = ["| mode | Re χ(R) max abs / rel | Im χ(R) max abs / rel | χ² max rel | R-factor abs err | "
             "α₁ (01, 02, 12) | decisions = exact |", "|---|---|---|---|---|---|---|"]
    for r in res["modes"]:
        re_ = r.get("chiR_re", {"max_abs": 0.0, "max_rel": 0.0})
        im_ = r.get("chiR_im", {"max_abs": 0.0, "max_rel": 0.0})
        lines.append(f"| {r['mode']} | {re_['max_abs']:.2e} / {re_['max_rel']:.2e} | "
                     f"{im_['max_abs']:.2e} / {im_['max_rel']:.2e} | {r['chi2']['max_rel']:.2e} | "
                     f"{r['r_factor']['abs_err']:.2e} | {r['alpha'][0]} | "
                     f"{100 * r['decisions_equal_to_exact']:.2f} % |")
    return "\n".join(lines)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="fcc_crconi_4000")
    ap.add_argument("--steps", type=int, default=20000)
    args = ap.parse_args(argv)
    res = study(args.workload, args.steps)
    print(json.dumps(res))
    print(table(res), file=sys.stderr)


if __name__ == "__main__":
    main()
