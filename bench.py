"""Benchmark: parallel-tempering RMC swap moves/s on one MI355X.

Workload (default): fcc CrCoNi 10x10x10 cells = 4000 atoms, a0 3.56 A, single
scattering + 3-/4-leg multiple scattering to 5.5 A, three edges, k^1..3 channels,
complex chi(R) on R 1.4-5.0 A (step 0.02), 32 parallel-tempering replicas, swap moves,
exchange every 25 steps.  A "step" is one swap proposal per replica; `value` is
replica swap moves/s over the timed steps (inputs resident in HBM).

Prints ONE JSON line with the roofline of the dominant kernel (timed with HIP events
on the sampler's stream) and the CPU baseline (the FP64 oracle's plain reference
algorithm timed on this host on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MIN_ESS_STEPS = 2000          # shorter alpha series give no usable autocorrelation estimate
FP64_PEAK_TFLOPS = 78.6


def host_cores():
    """(cores this process may run on, cores the host has, the host CPU's model name)."""
    import subprocess
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    model = ""
    try:
        model = next((ln.split(":", 1)[1].strip() for ln in
                      subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines()
                      if ln.startswith("Model name")), "")
    except OSError:
        pass
    return usable, os.cpu_count() or usable, model


# the host cores that go with one GPU on a shared node: worker pools are sized to this share
CPU_SHARE = 16


def cpu_baseline(P, seconds=15.0, warmup=200):
    """The oracle's plain sampler (the reference algorithm step for step, FP64 NumPy) on this
    host, in the two arrangements of the baseline protocol: (i) one process with every
    core's BLAS threads running all replicas; (ii) one single-threaded process per core,
    each running its share of the replicas (replicas are independent between exchanges).
    Each process makes `warmup` untimed steps, then a bounded timed sample of about
    `seconds`; the better arrangement is the baseline.  The cores used are the host's
    share for one GPU (CPU_SHARE, or fewer if this process may use fewer).  The workers
    are separate Python processes that import only NumPy and the oracle (no GPU)."""
    import pickle
    import subprocess
    import tempfile
    from oracle.build import problem_from
    O = problem_from(P)
    nrep = P.sampler.nrep
    usable, total, model = host_cores()
    cores = max(1, min(CPU_SHARE, usable))
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "problem.pkl")
        with open(path, "wb") as fh:
            pickle.dump((O, P.cfg0), fh)

        def arrangement(threads, reps):
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS=str(threads),
                       MKL_NUM_THREADS=str(threads))
            procs = [subprocess.Popen([sys.executable, "-m", "oracle.cpu_worker", path,
                                       str(seconds), str(r), str(1 + q), str(warmup)], cwd=ROOT,
                                      env=env, stdout=subprocess.PIPE, text=True)
                     for q, r in enumerate(reps)]
            outs = []
            for pr in procs:
                out, _ = pr.communicate()
                if pr.returncode != 0:
                    raise RuntimeError(f"CPU baseline worker failed with code {pr.returncode}")
                outs.append(json.loads(out.strip().splitlines()[-1]))
            return sum(o["moves"] / o["seconds"] for o in outs), outs

        v1, o1 = arrangement(cores, [nrep])
        nproc = min(cores, nrep)
        reps = [nrep // nproc + (1 if q < nrep % nproc else 0) for q in range(nproc)]
        v2, o2 = arrangement(1, reps)
    best = max((v1, cores, "one process, %d BLAS threads, %d replicas" % (cores, nrep)),
               (v2, nproc, "%d single-threaded processes, %s replicas each" % (nproc, "/".join(map(str, sorted(set(reps)))))))
    return {"value": best[0], "unit": "replica swap moves/s", "cores": best[1], "kind": "port",
            "host": {"cpu_model": model, "host_cores": total, "usable_cores": usable,
                     "cores_per_gpu_share": CPU_SHARE, "numpy": np.__version__,
                     "threads_env": "OMP_NUM_THREADS=OPENBLAS_NUM_THREADS=MKL_NUM_THREADS set per arrangement"},
            "arrangements": {"one_process_blas_threads": {"threads": cores, "moves_per_s": round(v1, 1),
                                                          "steps": o1[0]["steps"], "warmup": o1[0]["warmup"]},
                             "processes_single_thread": {"processes": nproc, "moves_per_s": round(v2, 1),
                                                         "steps": o2[0]["steps"], "warmup": o2[0]["warmup"]}},
            "sample": (f"{best[2]}; (i) 1 process x {cores} threads: {v1:.1f} moves/s over "
                       f"{o1[0]['steps']} steps, (ii) {nproc} processes x 1 thread: {v2:.1f} moves/s "
                       f"over {o2[0]['steps']} steps each, each after {warmup} untimed warm-up steps "
                       f"(about {seconds:.0f} s timed per arrangement, same workload, FP64 NumPy oracle "
                       f"of the reference algorithm; host {model}, {total} cores, {usable} usable, "
                       f"{CPU_SHARE} per GPU)")}


def ess_run(P, S, burn, steps, exchange_every, setup):
    """Equilibrate from the workload's start configuration, then time `steps` steps and
    take the alpha_1 ESS (smallest over the unlike pairs) of the likelihood-temperature
    rung."""
    from chimera_amd import workloads as W
    from chimera_amd.ess import ess
    S.init_state(P.cfg0, P.cell.ideal)
    setup()
    # recording is set before the burn-in, whose first run captures the step graphs
    S.set_recording(log_steps=0, series_steps=burn + steps)
    S.run(burn, exchange_every=exchange_every, log_every=10 ** 9)
    a0 = S.accepted_per_rung().copy()
    P.fm.ctx.sync()
    t0 = time.perf_counter()
    S.run(steps, exchange_every=exchange_every, log_every=10 ** 9)
    P.fm.ctx.sync()
    dt = time.perf_counter() - t0
    cr = int(S.collect_rungs()[0])
    acc = (S.accepted_per_rung() - a0) / steps
    al = S.alpha_series(burn + steps, shell=0)[burn:]
    iu = np.triu_indices(P.nsp, 1)
    e = min(ess(al[:, a, b]) for a, b in zip(*iu))
    bsp, _, bc2 = S.best()
    cnt = np.bincount(bsp, minlength=P.nsp)
    chib = S.chi_k_of(bsp[None])[0] / np.maximum(cnt, 1)[:, None]
    mb = P.ft.to_R(chib[None])[0]
    rfac = W.r_factor(mb, P.target, P.mask, W.fit_scales(mb, P.target, P.mask, P.chan_edge))
    return {"us_per_step": round(1e6 * dt / steps, 3), "alpha1_ess": round(e, 1),
            "alpha1_ess_per_s": round(e / dt, 2), "accept_collect_rung": round(float(acc[cr]), 4),
            "alpha1_mean": {f"{a}{b}": round(float(al[:, a, b].mean()), 5) for a, b in zip(*iu)},
            "best_chi2": round(float(bc2), 6), "best_r_factor": float(f"{rfac:.4g}")}


def learned_vs_uniform(P, S, args):
    """Train the learned proposal on move records of the uniform sampler, then compare the
    two samplers' alpha_1 ESS/s under the same protocol."""
    from chimera_amd.learned import LearnedProposal
    lp = LearnedProposal(S, seed=0)
    S.init_state(P.cfg0, P.cell.ideal)
    lp.record(args.train_steps * S.nrep)
    t0 = time.perf_counter()
    S.run(args.train_steps, exchange_every=args.exchange_every, log_every=10 ** 9)
    P.fm.ctx.sync()
    t_rec = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss1 = lp.train(iters=args.train_iters // 2, batch=1024, lr=1e-3, seed=1)
    loss2 = lp.train(iters=args.train_iters - args.train_iters // 2, batch=1024, lr=3e-4, seed=2)
    t_train = time.perf_counter() - t0
    nv, occ, y, meta = lp.records(min(lp.rec_cap, 200000))
    ok = meta >= 0
    g = lp.eval(occ[ok][:20000])
    yy = np.clip(y[ok][:20000], -30, 30)
    corr = float(np.corrcoef(g, yy)[0, 1])
    lp.record(0)
    settings = dict(ncand=args.lp_ncand, lam=1.0, kappa=args.lp_kappa, eps=0.05,
                    beta_max=args.lp_beta_max, lift=not args.lp_symmetric)
    print(f"[bench] learned: trained {args.train_iters} its, loss {loss2:.3f}, corr {corr:.3f}; "
          "ESS runs", file=sys.stderr, flush=True)
    uni = ess_run(P, S, args.ess_burn, args.ess_steps, args.exchange_every, lambda: lp.disable())
    lrn = ess_run(P, S, args.ess_burn, args.ess_steps, args.exchange_every,
                  lambda: lp.enable(**settings))
    prof = lp.profile(50)
    lp.disable()
    # MFMA work of one scoring launch (every slot learned): layer 2 and the output of each
    # slot's candidates at the committed state (layer 1 is a sum of two per-site rows; the
    # re-scored rows near the pending swap are not counted)
    flops = 2.0 * S.nrep * args.lp_ncand * (256 * 256 + 256)
    return {"settings": settings,
            "training": {"records": int(nv), "record_steps": args.train_steps,
                         "record_s": round(t_rec, 2), "iters": args.train_iters, "batch": 1024,
                         "train_s": round(t_train, 2), "loss": round(loss2, 4),
                         "loss_first_half": round(loss1, 4), "fit_corr": round(corr, 4)},
            "uniform": uni, "learned": lrn,
            "alpha1_ess_per_s_ratio": round(lrn["alpha1_ess_per_s"] / uni["alpha1_ess_per_s"], 3),
            "alpha1_ess_per_step_ratio": round(lrn["alpha1_ess"] / uni["alpha1_ess"], 3),
            "kernels_ms": {k: round(v, 5) for k, v in prof.items()},
            "scorer_roofline": {"bound": "mfma", "achieved": round(flops / (prof["score_ms"] * 1e-3) / 1e12, 3),
                                "peak": 2500.0, "unit": "TFLOP/s",
                                "frac": round(flops / (prof["score_ms"] * 1e-3) / 1e12 / 2500.0, 6)}}


def replica_sweep(args, counts, steps, warmup):
    """The same workload and step at more replicas per GPU: at 32 replicas both kernels of a
    step are dependent latency chains, so the step time grows much more slowly than the
    work.  Each count is its own sampler (built, autotuned, warmed up, then timed)."""
    from chimera_amd import _lib, autotune
    from chimera_amd import workloads as W
    out = []
    for n in counts:
        P = W.build(args.workload, nrep=n)
        S = P.sampler
        S.precision, S.reanchor_every = args.precision, args.reanchor_every
        S.paired = not args.single_steps
        S.init_state(P.cfg0, P.cell.ideal)
        tuned = autotune.tune(S, use_cache=False)
        S.run(warmup, exchange_every=args.exchange_every, log_every=10 ** 9)
        P.fm.ctx.sync()
        a0 = float(S.accepted_per_rung().sum())
        t0 = time.perf_counter()
        S.run(steps, exchange_every=args.exchange_every, log_every=10 ** 9)
        P.fm.ctx.sync()
        dt = time.perf_counter() - t0
        acc = (float(S.accepted_per_rung().sum()) - a0) / (steps * n)
        ms = np.zeros(4)
        _lib.check(_lib.load().chm_sampler_profile(S._S, 20, ms))
        out.append({"replicas": n, "steps": steps, "nsplit": tuned["nsplit"],
                    "us_per_step": round(1e6 * dt / steps, 3),
                    "replica_moves_per_s": round(n * steps / dt, 1),
                    "transform_us_per_step": round(1e3 * ms[1], 3),
                    "per_replica_us_per_step": round(1e3 * ms[2], 3),
                    "accept": round(acc, 4)})
        print(f"[bench] {n} replicas: {out[-1]['us_per_step']} us/step", file=sys.stderr, flush=True)
        del S, P
    return out


def forward_bench(P, S, iters=10):
    """Full chi(k) evaluations of the replicas' configurations: the pair-space path sum
    (FP64 and FP32, with distances and pair types), its adjoint dL/dr, and the
    relaxed-occupancy path sum and its adjoint (dL/dr, dL/dw), timed with HIP events on the forward model's stream.
    Work unit: one path term = one (pair, k) evaluation (exp, sin, amplitude)."""
    fm = P.fm
    sp = S.species
    nconf = sp.shape[0]
    pos = np.repeat(P.cell.ideal[None], nconf, 0)
    c64 = fm.chi(pos, sp, normalise=False)
    c32 = fm.chi(pos, sp, normalise=False, precision="f32")
    g = np.ones((nconf, P.nsp, P.k.size))
    fm.grad_pos(pos, sp, g)
    w = np.eye(P.nsp)[sp]
    fm.chi_soft(pos, w)
    fm.grad_soft(pos, w, g)
    terms = float(nconf) * fm.npairs * int(np.count_nonzero(P.k > 0))
    err = np.abs(c32 - c64)
    out = {"configs": nconf, "pairs": int(fm.npairs), "k_points": int(P.k.size),
           "terms_per_launch": terms,
           "f32_vs_f64": {"max_abs": float(f"{err.max():.3g}"),
                          "max_rel": float(f"{err.max() / np.abs(c64).max():.3g}")}}
    for key, what in (("path_sum_f64", "chi"), ("path_sum_f32", "chi_f32"),
                      ("adjoint_dr_f64", "adjoint"), ("soft_path_sum_f64", "soft"),
                      ("soft_adjoint_f64", "soft_adjoint")):
        ms = fm.profile(what, nconf, iters)
        out[key] = {"ms": round(ms, 4), "terms_per_s": float(f"{terms / (ms * 1e-3):.4g}")}
    out["chi_evals_per_s_f32"] = round(nconf / (out["path_sum_f32"]["ms"] * 1e-3), 1)
    out["chi_evals_per_s_f64"] = round(nconf / (out["path_sum_f64"]["ms"] * 1e-3), 1)
    if P.ms is not None:
        # full chi(k) with multiple scattering, end to end from the host (species in,
        # chi(k) out): FP64 pair-space sum of the single-scattering pairs + the entry
        # counts of every 3-/4-leg path from scratch + counts x entry table.  The work
        # unit is the survey's: (ordered SS pairs + reversal-merged MS paths) x nk terms.
        best = float("inf")
        for _ in range(5):
            fm.ctx.sync()
            t0 = time.perf_counter()
            S.chi_k_of(sp)
            fm.ctx.sync()
            best = min(best, time.perf_counter() - t0)
        p_ms = int(P.cell.natoms) * P.ms.merged_per_absorber
        terms_ms = float(nconf) * (fm.npairs + p_ms) * int(P.k.size)
        out["full_chi_with_ms_f64"] = {
            "ms": round(1e3 * best, 4), "configs": nconf, "ms_paths_merged": p_ms,
            "chi_evals_per_s": round(nconf / best, 1),
            "path_term_equivalents_per_s": float(f"{terms_ms / best:.4g}"),
            "note": "host to host: species upload, pair-space FP64 SS sum, MS entry counts "
                    "from scratch and the entry-table product, chi(k) download"}
    # the transcendental bound of an FP32 term (exp, sin and a reciprocal on the
    # quarter-rate unit: 256 CUs x 16 lanes x 2.4 GHz / 3 = 3.28e12 terms/s)
    out["path_sum_f32"]["frac_of_transcendental_bound"] = round(
        out["path_sum_f32"]["terms_per_s"] / 3.28e12, 4)
    return out


def pair_leg(P, steps, warmup, exchange_every, seed=1):
    """The reference sampler's own default move mix (methods/method_a_rmc.py:64-66:
    swap_frac 0.75, disp_sigma 0.015, max_disp 0.30) on the pair-space engine, same
    workload and replica count: replica moves/s over `steps` timed steps (swaps and
    displacements counted alike), the acceptance, and a check that the tracked chi^2 of
    every replica equals a recompute from its species and positions."""
    from chimera_amd.method_a_rmc import RMCSampler
    S = RMCSampler(P.fm, P.ft, P.target, P.mask, P.sigma, nrep=P.sampler.nrep, seed=seed,
                   ms_library=P.ms)
    S.init_state(P.cfg0, P.cell.ideal)
    S.run(warmup, exchange_every=exchange_every, log_every=10 ** 9)
    P.fm.ctx.sync()
    a0 = S.accepted_per_rung().sum()
    t0 = time.perf_counter()
    S.run(steps, exchange_every=exchange_every, log_every=10 ** 9)
    P.fm.ctx.sync()
    dt = time.perf_counter() - t0
    acc = (S.accepted_per_rung().sum() - a0) / (steps * S.nrep)
    sp, pos = S.species, S.pos
    chk = S.chi2_of(sp, pos)
    err = float(np.max(np.abs(chk - S.chi2) / np.abs(chk)))
    out = {"engine": S.engine, "swap_frac": S.swap_frac, "disp_sigma": S.disp_sigma,
           "max_disp": S.max_disp, "steps": steps, "warmup": warmup,
           "ms_per_step": round(1e3 * dt / steps, 5),
           "replica_moves_per_s": round(S.nrep * steps / dt, 1),
           "acceptance": round(float(acc), 4), "max_displacement_A": round(float(
               np.linalg.norm(pos - P.cell.ideal[None], axis=-1).max()), 4),
           "chi2_tracked_vs_recomputed_max_rel": float(f"{err:.3g}")}
    S._destroy()
    return out


def batch_leg(nprob, steps, warmup, exchange_every, precision, reanchor_every=512,
              sequential=False, log=None):
    """The uneven batch (batch_specs(nprob): fcc/bcc supercells of 256-4000 sites, 1-3
    edges, series of 2-6 points, 16 replicas each) stepped together: one launch per kernel
    per step, one hipGraph per exchange period; aggregate replica swap moves/s over
    `steps` timed steps.  With `sequential`, also the same problems stepped one after
    another for comparison."""
    from chimera_amd import batch as BT
    specs = BT.batch_specs(nprob)
    t0 = time.perf_counter()
    probs = BT.build(specs)
    t_build = time.perf_counter() - t0
    if log:
        log(f"[bench] built {len(probs)} problems in {t_build:.1f} s")
    B = BT.Batch(probs, precision=precision, reanchor_every=reanchor_every, log=log)
    B.run(warmup, exchange_every=exchange_every)
    dt = B.run(steps, exchange_every=exchange_every)
    nrep = B.replicas
    moves = nrep * steps / dt
    prof = B.profile(50)
    sites = [int(P.cell.natoms) for P in probs]
    out = {"value": round(moves, 1), "unit": "replica swap moves/s", "steps": steps,
           "warmup": warmup, "ms_per_step": round(1e3 * dt / steps, 5),
           "problems": len(probs), "replicas": nrep, "sites_total": int(sum(sites)),
           "sites": {str(n): sites.count(n) for n in sorted(set(sites))},
           "lattices": {lat: sum(P.w.lattice == lat for P in probs) for lat in ("fcc", "bcc")},
           "edges": {str(k): sum(len(P.w.edges or range(P.nsp)) == k for P in probs) for k in (1, 2, 3)},
           "kernels_ms": {k: round(v, 5) for k, v in prof.items()},
           "autotune": B.tuned, "nsplit": [B._split(P.sampler) for P in probs],
           "build_s": round(t_build, 1)}
    B.close()
    if sequential:
        seq_steps = max(200, steps // 10)
        t_seq = 0.0
        for P in probs:
            S = P.sampler
            S.run(exchange_every, exchange_every=exchange_every, log_every=10 ** 9)   # graphs
            P.fm.ctx.sync()
            t1 = time.perf_counter()
            S.run(seq_steps, exchange_every=exchange_every, log_every=10 ** 9)
            P.fm.ctx.sync()
            t_seq += time.perf_counter() - t1
        out["sequential_moves_per_s"] = round(nrep * seq_steps / t_seq, 1)
        out["batched_over_sequential"] = round(moves / out["sequential_moves_per_s"], 2)
    del probs, B
    return out


def batch_bench(args):
    """Uneven problems stepped together (one launch per kernel per step, one hipGraph per
    exchange period): aggregate replica swap moves/s."""
    log = lambda m: print(m, file=sys.stderr, flush=True)       # noqa: E731
    r = batch_leg(args.batch, args.steps, args.warmup, args.exchange_every, args.precision,
                  args.reanchor_every, sequential=True, log=log)
    moves, steps, dt = r["value"], args.steps, r["ms_per_step"] * args.steps * 1e-3
    nrep = r["replicas"]
    out = {
        "metric": "aggregate replica swap moves/s, batched uneven RMC problems",
        "value": round(moves, 1), "unit": "replica swap moves/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(1e3 * dt / steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 chi^2; " + {"bf16x3": "bf16x3 MFMA delta + f64 accumulator",
                                  "f32": "f32 delta + f64 accumulator", "f64": "f64 transform"}[args.precision],
        "data": "synthetic (seeded tables, seeded reference configurations + noise)",
        "config": {"workload": f"batch of {r['problems']} uneven problems", "problems": r["problems"],
                   "replicas": nrep, "sites_total": r["sites_total"], "sites": r["sites"],
                   "lattices": r["lattices"], "edges": r["edges"],
                   "exchange_every": args.exchange_every},
        "kernels_ms": r["kernels_ms"],
        "sequential_moves_per_s": r["sequential_moves_per_s"],
        "batched_over_sequential": r["batched_over_sequential"],
        "autotune": r["autotune"],
        "nsplit": r["nsplit"],
        "build_s": r["build_s"],
    }
    print(json.dumps(out))


def recorded_traffic(kernel, nsplit, workload):
    """HBM-side bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by a plain copy of known size), when it was
    measured on this workload at this launch split; else None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.txt")
    try:
        lines = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith("#")]
    except OSError:
        return None
    scale = {ln[0]: float(ln[1]) for ln in lines if len(ln) == 2}
    for ln in lines:
        if len(ln) > 2 and ln[0] == "kernel" and ln[1] == kernel:
            f = dict(zip(ln[2::2], ln[3::2]))
            if int(f["nsplit"]) != nsplit:
                return None
            kib = float(f["fetch_kib"]) / scale["fetch_scale"] + float(f["write_kib"]) / scale["write_scale"]
            return round(kib * 1024.0)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--workload", default="fcc_crconi_4000")
    ap.add_argument("--exchange-every", type=int, default=25)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-warmup", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-iters", type=int, default=200)
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "f64", "f32"])
    ap.add_argument("--reanchor-every", type=int, default=512)
    ap.add_argument("--batch", type=int, default=0, metavar="NPROB",
                    help="only run the batched benchmark over NPROB uneven problems and exit")
    ap.add_argument("--precision-study", type=int, default=0, metavar="STEPS",
                    help="only run the precision study (STEPS steps per mode) and exit")
    ap.add_argument("--no-forward", action="store_true",
                    help="skip the full chi(k) path-sum / adjoint timings")
    ap.add_argument("--forward-only", action="store_true",
                    help="only time the forward-model kernels (path sums, adjoint) and exit")
    ap.add_argument("--learned-only", action="store_true",
                    help="only the learned-vs-uniform comparison (training, both ESS runs, the "
                         "scorer profile), as one JSON line")
    ap.add_argument("--learned-timing", type=int, default=0, metavar="STEPS",
                    help="time STEPS steps of the learned sampler (the bench's learned settings, "
                         "the network's initial weights, no recording) and exit")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-vs-uniform proposal comparison")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the short precision study (every transform / accumulator mode)")
    ap.add_argument("--precision-steps", type=int, default=2000)
    ap.add_argument("--no-batch-leg", action="store_true",
                    help="skip the short batch of uneven problems")
    ap.add_argument("--batch-leg-problems", type=int, default=64)
    ap.add_argument("--no-pair-leg", action="store_true",
                    help="skip the pair-space engine leg (swap_frac 0.75 with displacements)")
    ap.add_argument("--pair-leg-steps", type=int, default=2000)
    ap.add_argument("--no-replica-sweep", action="store_true",
                    help="skip the same workload at 64, 128 and 256 replicas")
    ap.add_argument("--replica-sweep-only", action="store_true")
    ap.add_argument("--pair-only", action="store_true",
                    help="only the pair-space engine leg (--pair-leg-steps timed after --warmup)")
    ap.add_argument("--batch-leg-steps", type=int, default=2000)
    ap.add_argument("--ess-steps", type=int, default=100000)
    ap.add_argument("--ess-burn", type=int, default=10000)
    ap.add_argument("--train-steps", type=int, default=312500,
                    help="uniform steps whose move records train the learned proposal (x replicas records)")
    ap.add_argument("--train-iters", type=int, default=24000)
    ap.add_argument("--lp-ncand", type=int, default=256)
    ap.add_argument("--lp-kappa", type=float, default=0.5)
    ap.add_argument("--lp-symmetric", action="store_true",
                    help="symmetric tilt kappa |dU| instead of the lifted kappa d dU")
    ap.add_argument("--lp-beta-max", type=float, default=2.0)
    ap.add_argument("--nsplit", type=int, default=0,
                    help="workgroups per replica in the slot kernel (0: autotune)")
    ap.add_argument("--single-steps", action="store_true",
                    help="one Metropolis step per launch pair (paired steps are the default)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the known-size "
                         "kernel that scales hardware byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (replicas shard across GPUs by "
              "design, but no multi-GPU path is built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))
        return
    if args.batch:
        batch_bench(args)
        return
    if args.precision_study:
        from chimera_amd import precision_study
        precision_study.main(["--steps", str(args.precision_study)])
        return
    if args.replica_sweep_only:
        print(json.dumps({"workload": args.workload,
                          "replica_sweep": replica_sweep(args, (32, 64, 128, 256), 4000, 500)}))
        return
    if args.pair_only:
        from chimera_amd import workloads as W
        P = W.build(args.workload)
        print(json.dumps({"workload": args.workload,
                          "pair_engine_default_moves": pair_leg(P, args.pair_leg_steps, args.warmup,
                                                                args.exchange_every)}))
        return
    if args.learned_timing:
        from chimera_amd import workloads as W
        from chimera_amd.learned import LearnedProposal
        P = W.build(args.workload)
        S = P.sampler
        S.precision, S.reanchor_every = args.precision, args.reanchor_every
        S.paired = not args.single_steps
        S.init_state(P.cfg0, P.cell.ideal)
        lp = LearnedProposal(S, seed=0)
        lp.enable(ncand=args.lp_ncand, lam=1.0, kappa=args.lp_kappa, eps=0.05, beta_max=args.lp_beta_max,
                  lift=not args.lp_symmetric)
        S.run(args.warmup, exchange_every=args.exchange_every, log_every=10 ** 9)
        P.fm.ctx.sync()
        t0 = time.perf_counter()
        S.run(args.learned_timing, exchange_every=args.exchange_every, log_every=10 ** 9)
        P.fm.ctx.sync()
        dt = time.perf_counter() - t0
        print(json.dumps({"workload": args.workload, "learned_steps": args.learned_timing,
                          "us_per_step": round(1e6 * dt / args.learned_timing, 3),
                          "kernels_ms": {k: round(v, 5) for k, v in lp.profile(50).items()}}))
        return
    if args.learned_only:
        from chimera_amd import workloads as W
        P = W.build(args.workload)
        S = P.sampler
        S.precision, S.reanchor_every = args.precision, args.reanchor_every
        S.paired = not args.single_steps
        S.init_state(P.cfg0, P.cell.ideal)
        if args.nsplit:
            from chimera_amd import _lib
            _lib.check(_lib.load().chm_sampler_set_tuning(S._S, args.nsplit))
        print(json.dumps({"workload": args.workload, "learned_vs_uniform": learned_vs_uniform(P, S, args)}))
        return
    if args.forward_only:
        from chimera_amd import workloads as W
        P = W.build(args.workload)
        print(json.dumps({"workload": args.workload, "forward_model": forward_bench(P, P.sampler)}))
        return

    from chimera_amd import workloads as W
    print(f"[bench] building {args.workload}", file=sys.stderr, flush=True)
    t_build = time.time()
    P = W.build(args.workload)
    S = P.sampler
    S.precision, S.reanchor_every = args.precision, args.reanchor_every
    S.paired = not args.single_steps
    S.init_state(P.cfg0, P.cell.ideal)
    from chimera_amd import autotune
    if args.nsplit:
        from chimera_amd import _lib
        _lib.check(_lib.load().chm_sampler_set_tuning(S._S, args.nsplit))
        tuned = {"nsplit": args.nsplit, "times_ms": {}}
    else:
        tuned = autotune.tune(S, use_cache=False)
    print(f"[bench] tuned: {tuned}", file=sys.stderr, flush=True)
    print(f"[bench] built in {time.time() - t_build:.1f} s; warmup {args.warmup}",
          file=sys.stderr, flush=True)
    # pair counts of the likelihood-temperature rung after every warm-up and timed step (the
    # alpha series); set before the warm-up, whose first run captures the step graphs
    S.set_recording(log_steps=0, series_steps=args.warmup + args.steps)
    S.run(args.warmup, exchange_every=args.exchange_every, log_every=10 ** 9)
    P.fm.ctx.sync()
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    S.run(args.steps, exchange_every=args.exchange_every, log_every=10 ** 9)
    P.fm.ctx.sync()
    dt = time.perf_counter() - t0
    moves = S.nrep * args.steps / dt

    # per-kernel device times (HIP events on the sampler's stream)
    from chimera_amd import _lib
    ms = np.zeros(4)
    _lib.check(_lib.load().chm_sampler_profile(S._S, args.profile_iters, ms))
    _, k2, k_slot, kstep = ms          # per step (a paired launch makes two steps)
    paired = bool(getattr(S, "paired", False)) and args.precision == "bf16x3"
    nrow, ncol_e, nrep, nsp = S.nrow_e, S.nkw * S.nwin * 2, S.nrep, P.nsp
    ncol_tot = nsp * ncol_e
    # algorithmic work of one launch of each kernel (per replica: species, path records,
    # tile sums, counts + count-change rows, running chi(R) copy)
    flops_k2 = 2.0 * nrep * nsp * nrow * ncol_e                     # one real GEMM
    ms_stencil = 0 if P.ms is None else P.ms.nstencil
    nct = -(-ncol_e // 16)
    if paired:
        # three trials per launch (m1, m2r, m2a); two decisions, two or three gathers over
        # one read of the path records, three tile-sum sets and count-change rows
        flops_k2 *= 3
        bytes_slot = nrep * (P.cell.natoms + 2 * ms_stencil * 16 + 3 * 8 * nsp * nct * 4
                             + 4 * nsp * nrow * 7 + 2 * 8 * ncol_tot)
        k2_launch, kslot_launch = 2 * k2, 2 * k_slot
        kname, sname = "k_gemm_pair_bf16", "k_slot_pair"
    else:
        bytes_slot = nrep * (P.cell.natoms + 2 * ms_stencil * 16 + 8 * nsp * nct * 4
                             + 4 * nsp * nrow * 3 + 2 * 8 * ncol_tot)
        k2_launch, kslot_launch = k2, k_slot
        kname, sname = {"bf16x3": "k_gemm_delta_bf16", "f32": "k_gemm_delta_f32",
                        "f64": "k_gemm_abs_f64"}[args.precision], "k_slot_step"
    kern = {"slot_step_ms": k_slot, "trial_transform_ms": k2, "step_ms": kstep,
            "slot_launch_ms": kslot_launch, "transform_launch_ms": k2_launch}
    peak2 = {"bf16x3": 2500.0 / 3.0,                 # split-BF16: BF16 dense / 3
             "f32": 157.3,                           # FP32 (vector + matrix) peak
             "f64": FP64_PEAK_TFLOPS}[args.precision]
    rooflines = {
        sname: {"bound": "hbm", "achieved": bytes_slot / (kslot_launch * 1e-3) / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s"},
        kname: {"bound": "mfma", "achieved": flops_k2 / (k2_launch * 1e-3) / 1e12, "peak": peak2,
                "unit": "TFLOP/s"},
    }
    dom = sname if k_slot >= k2 else kname
    r = rooflines[dom]
    from chimera_amd import autotune as _at
    roof = {"bound": r["bound"], "achieved": round(r["achieved"], 3), "peak": r["peak"],
            "unit": r["unit"], "frac": round(r["achieved"] / r["peak"], 5),
            "traffic": recorded_traffic(dom, _at.current(S), args.workload),
            "kernel": dom,
            "all": {k: {"achieved": round(v["achieved"], 3), "unit": v["unit"],
                        "frac": round(v["achieved"] / v["peak"], 5)} for k, v in rooflines.items()}}

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(P, args.cpu_seconds, args.cpu_warmup)

    # alpha_1 ESS/s of the likelihood-temperature rung over the timed steps (only when the
    # series is long enough for the autocorrelation estimate), and the R-factor of the best
    # state (reported beside the rate)
    from chimera_amd.ess import ess
    al = S.alpha_series(args.warmup + args.steps, shell=0)[args.warmup:]
    iu = np.triu_indices(P.nsp, 1)
    ess_min = None
    if args.steps >= MIN_ESS_STEPS:
        ess_min = float(min(ess(al[:, a, b]) for a, b in zip(*iu)))
    bsp, _, bc2 = S.best()
    cnt = np.bincount(bsp, minlength=P.nsp)
    chib = S.chi_k_of(bsp[None])[0] / np.maximum(cnt, 1)[:, None]
    mb = P.ft.to_R(chib[None])[0]
    rfac = W.r_factor(mb, P.target, P.mask, W.fit_scales(mb, P.target, P.mask, P.chan_edge))
    nanchor = int(S.reanchors().sum()) if args.precision != "f64" else 0
    fwd = None
    if not args.no_forward:
        print("[bench] forward model", file=sys.stderr, flush=True)
        fwd = forward_bench(P, S)
    c3 = None
    if not args.no_precision:
        print("[bench] precision study (short)", file=sys.stderr, flush=True)
        from chimera_amd import precision_study
        res = precision_study.study(args.workload, steps=args.precision_steps, problem=P)
        c3 = {"steps_per_mode": args.precision_steps, "chik_f32_vs_f64": res["chik_f32_vs_f64"],
              "modes": [{k: v for k, v in r.items() if k not in ("steps",)} for r in res["modes"]]}
        S.precision, S.reanchor_every = args.precision, args.reanchor_every
        S.init_state(P.cfg0, P.cell.ideal)
    lvu = None
    if not args.no_learned:
        lvu = learned_vs_uniform(P, S, args)
        print(f"[bench] learned/uniform alpha_1 ESS/s: {lvu['alpha1_ess_per_s_ratio']}",
              file=sys.stderr, flush=True)
    pl = None
    if not args.no_pair_leg:
        print("[bench] pair-space engine, the reference's default move mix", file=sys.stderr, flush=True)
        pl = pair_leg(P, args.pair_leg_steps, 200, args.exchange_every)
    rs = None
    if not args.no_replica_sweep:
        print("[bench] replica-count sweep", file=sys.stderr, flush=True)
        rs = replica_sweep(args, (64, 128, 256), 2000, 300)
    c2 = None
    if not args.no_batch_leg:
        print("[bench] batch of uneven problems (short)", file=sys.stderr, flush=True)
        c2 = batch_leg(args.batch_leg_problems, args.batch_leg_steps, 100, args.exchange_every,
                       args.precision, args.reanchor_every)
    out = {
        "metric": "RMC swap moves/s & alpha_1 ESS/s at fixed R-factor, learned vs uniform proposal",
        "value": round(moves, 1),
        "unit": "replica swap moves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 chi^2; " + {"bf16x3": "bf16x3 MFMA delta + f64 accumulator",
                                  "f32": "f32 delta + f64 accumulator",
                                  "f64": "f64 transform"}[args.precision],
        "data": "synthetic (seeded tables, seeded reference configuration + noise)",
        "config": {"workload": args.workload, "natoms": int(P.cell.natoms),
                   "replicas": int(S.nrep), "edges": int(P.nsp), "channels": int(P.nchan),
                   "R_points": int(S.nwin), "ms_paths": P.ms is not None,
                   "exchange_every": args.exchange_every, "paired_steps": paired},
        "kernels_ms": {k: round(v, 5) for k, v in kern.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "gpu_over_cpu": None if cpu is None else round(moves / cpu["value"], 1),
        "best_chi2": round(float(bc2), 5),
        "best_r_factor": float(f"{rfac:.4g}"),
        "alpha1_ess": None if ess_min is None else round(ess_min, 1),
        "alpha1_ess_per_s": None if ess_min is None else round(ess_min / dt, 2),
        "alpha1_mean": {f"{a}{b}": round(float(al[:, a, b].mean()), 5) for a, b in zip(*iu)},
        "fp64_reanchors": nanchor,
        "forward_model": fwd,
        "learned_vs_uniform": lvu,
        "precision_study": c3,
        "batch_uneven": c2,
        "pair_engine_default_moves": pl,
        "replica_sweep": rs,
        "autotune": {"nsplit": tuned["nsplit"],
                     "step_ms": {k: round(v, 5) for k, v in tuned["times_ms"].items()}},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
