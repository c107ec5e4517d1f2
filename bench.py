#!/usr/bin/env python3
"""Free-entropy + gradient throughput on one MI355X.

Workload (one step = one objective+gradient evaluation of the whole lattice,
inputs resident in HBM): FCC Pd 40x40x40 cells = 256,000 Pd + 256,000 H
sites, a = 3.89 A, periodic, H occupancy iid U(0, 0.6), displacements
N(0, 0.03 A), sigma U(0.08, 0.15) A, T = 500 K, rc = 5.35 A, 12-point rule
(--quad gh3: the 27-node Gauss-Hermite rule).

Prints one JSON line:  value = sites * steps / seconds of the timed region
(HIP events on the context stream, bracketed by device synchronisation),
plus the roofline of the dominant kernel, the FP64 comparator, the error of
the measured path against the FP64 kernel, the relaxation half of the metric
-- "batched" (64 uneven lattices relaxed in lockstep: relaxations/s, relaxed
objectives against FP64-only relaxations) and "serving" (Poisson request
streams over a sweep of rates: relaxations/s, p50/p99 latency, the FP64
one-at-a-time comparator) -- and the CPU baseline (the FP64 reference-
faithful oracle in FD mode on the host cores, on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "mixed": 157.3}  # TFLOP/s, MI355X vector FP64 / FP32 (published)
# Algorithmic flops per unordered pair per quadrature point, density pass and
# gradient pass (FMA = 2, sqrt/exp/rcp/log = 1): Pd-Pd, Pd-H, H-H.
FLOP_DENS = (62, 65, 27)
FLOP_GRAD = (180, 104, 72)
POINTS = {"r12": 12, "gh3": 27}  # quadrature points per pair


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000, help="timed evaluations (the workload's 1000 repeats)")
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--cells", type=int, default=40)
    p.add_argument("--precision", choices=["mixed", "fp64"], default=os.environ.get("A2C_BENCH_PREC", "mixed"))
    p.add_argument("--quad", choices=["r12", "gh3"], default="r12",
                   help="quadrature rule of the single-lattice workload")
    p.add_argument("--no-gh3", action="store_true", help="skip the Gauss-Hermite line beside the 12-point one")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--calib", action="store_true",
                   help="also run a 256 MiB copy kernel (counter calibration under rocprofv3 --pmc)")
    p.add_argument("--workload", choices=["single", "batched", "dmd", "plain"], default="single",
                   help="single: 40^3 evaluation throughput (default); batched: 64 uneven lattices relaxed "
                        "in lockstep; dmd: coupled diffusion + re-minimization on a slab; plain: the 6^3-cell "
                        "lattice evaluated once and relaxed")
    p.add_argument("--lattices", type=int, default=64)
    p.add_argument("--dmd-steps", type=int, default=100)
    p.add_argument("--dmd-cells", type=int, default=64)
    p.add_argument("--fp64-sample", type=int, default=8, help="batched: lattices relaxed one at a time in FP64")
    p.add_argument("--no-batched", action="store_true", help="single: skip the batched-relaxation block")
    p.add_argument("--no-serving", action="store_true", help="single: skip the serving block")
    p.add_argument("--serve-rates", default="5,10,20,40", help="serving: Poisson request rates (req/s)")
    p.add_argument("--serve-requests", type=int, default=200, help="serving: requests per rate")
    p.add_argument("--serve-fp64-sample", type=int, default=None,
                   help="serving: requests relaxed one at a time in FP64 (the comparator; default all)")
    p.add_argument("--serve-policy", default="deadline", choices=("deadline", "fcfs", "fair", "lockstep"),
                   help="serving: which running lattices a tick steps (a2c_amd.serve)")
    p.add_argument("--lmvm-m", type=int, default=5, help="batched: stored pairs of the LMVM (TAO's default 5)")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(fe, X, steps, warmup):
    N = fe.N
    dX, dG, df = fe.dev_alloc(X.nbytes), fe.dev_alloc(X.nbytes), fe.dev_alloc(8)
    try:
        fe.upload(dX, X)
        for _ in range(warmup):
            fe.eval_device(dX, dG, df)
        fe.sync()
        t0 = time.perf_counter()
        tm = fe.time_eval(dX, dG, df, steps)  # syncs at the end
        wall = time.perf_counter() - t0
        G = fe.download(np.zeros(4 * N), dG)
        f = fe.download(np.zeros(1), df)[0]
    finally:
        for p in (dX, dG, df):
            fe.dev_free(p)
    return tm, wall, f, G


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, mk):
    """Reference-faithful FP64 oracle (FD derivatives recomputed as the
    reference does) on a smaller lattice of the same recipe: per-site work is
    the same under periodic images, so sites/s carries over.  Timed on every
    core of this process's affinity mask up to 16 (the CPU share of one GPU
    on the box), then on one core; plus the plain method's relaxation
    (6^3 cells, GH3 rule) by scipy's L-BFGS-B (m = 5) on the oracle in FD
    mode, to ||g|| <= 1e-8 or where its line search stalls."""
    import oracle as O
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, avail))
    X, x, box, npd, nh = mk(6)

    def rate(nthr, budget):
        orc = O.Oracle(npd, nh, box, (1, 1, 1), T=500.0, deriv="fd", threads=nthr, quad=args.quad)
        orc.build_neighbors(X)
        orc.eval(X, x)  # warm-up
        n, t0 = 0, time.perf_counter()
        while True:
            orc.eval(X, x)
            n += 1
            if time.perf_counter() - t0 >= budget:
                break
        dt = time.perf_counter() - t0
        return (npd + nh) * n / dt, n, dt

    v, n, dt = rate(threads, args.cpu_seconds)
    v1, n1, dt1 = rate(1, max(2.0, args.cpu_seconds / 2))
    out = {"value": v, "unit": "atom-evals/s", "cores": threads, "kind": "port",
           "sample": "%d full FD-mode evaluations of the 6x6x6-cell lattice of the same recipe "
                     "(1,728 sites), %.1f s, %d OpenMP threads (%d cores in the affinity mask, at most 16 used)"
                     % (n, dt, threads, avail),
           "cpu_model": cpu_model(), "cores_available": avail,
           "one_core": {"value": v1, "evals": n1, "seconds": dt1}}
    out["plain_relaxation"] = cpu_relaxation(threads)
    return out


def cpu_relaxation(threads):
    """The plain method's case relaxed on the CPU: scipy L-BFGS-B (5 pairs)
    on the oracle's FD-mode objective and gradient (the reference's
    evaluation), from the same start as the GPU's plain workload."""
    import oracle as O
    from scipy.optimize import minimize
    import a2c_amd as A
    from a2c_amd import workloads
    s = workloads.plain()
    orc = O.Oracle(s["n_pd"], s["n_h"], s["box"], s["periodic"], T=s["T"], deriv="fd", threads=threads,
                   quad=s["quad"])
    orc.build_neighbors(s["X"])
    nfev = [0]

    def fg(v):
        nfev[0] += 1
        return orc.eval(v, s["x"])

    bounds = [(None, None)] * 3 + [(1e-6, None)]
    t0 = time.perf_counter()
    r = minimize(fg, s["X"], jac=True, method="L-BFGS-B", bounds=bounds * (s["n_pd"] + s["n_h"]),
                 options={"maxiter": 2000, "maxfun": 4000, "gtol": 1e-8 / np.sqrt(4 * (s["n_pd"] + s["n_h"])),
                          "ftol": 0.0, "maxcor": 5})
    dt = time.perf_counter() - t0
    g = np.linalg.norm(fg(r.x)[1])
    return {"seconds": dt, "evals": nfev[0] - 1, "iterations": int(r.nit), "f": float(r.fun), "gnorm": float(g),
            "threads": threads, "solver": "scipy L-BFGS-B m=5 on the FD-mode oracle",
            "stopped": str(r.message)}


def batched(args):
    """64 lattices, edge n ~ U{8..24} cells, T ~ U(300,700) K, mean H fraction
    ~ U(0.01, 0.6) with iid noise U(-0.05, 0.05), displacements N(0, 0.02 A),
    sigma 0.10 A, periodic; all relaxed in lockstep (one fused launch per
    evaluation round), mixed precision with the FP64 endgame, to ||Pg|| <= 1e-8.
    Comparator: the first --fp64-sample lattices relaxed one at a time in FP64."""
    import a2c_amd as A
    from a2c_amd import workloads
    specs = workloads.batched(args.lattices)
    sites = sum(s["n_pd"] + s["n_h"] for s in specs)
    opts = dict(gatol=1e-8, grtol=0.0, m=args.lmvm_m)
    log("batched: %d lattices, %d sites" % (len(specs), sites))
    with A.Batch(precision="mixed", max_lattices=len(specs)) as B:
        t0 = time.perf_counter()
        slots = [B.add(**s) for s in specs]
        t_adm = time.perf_counter() - t0
        for sl in slots:
            B.start(sl, **opts)
        t1 = time.perf_counter()
        B.run(progress=lambda it, n: log("batched: %d iterations, %d lattices running" % (it, n)))
        t_relax = time.perf_counter() - t1
        reps = [B.report(sl) for sl in slots]
    wall = t_adm + t_relax
    conv = sum(1 for r in reps if r["reason"] > 0)
    # one at a time in FP64 (sample)
    k = min(args.fp64_sample, len(specs))
    f64 = []
    t0 = time.perf_counter()
    for s in specs[:k]:
        with A.FreeEntropy(s["n_pd"], s["n_h"], s["box"], s["periodic"], T=s["T"], precision="fp64") as fe:
            fe.set_occupancy(s["x"])
            fe.findNeighbors(s["X"])
            f64.append(fe.minimizeFreeEntropy(s["X"], **opts)[1]["f"])
    t64 = time.perf_counter() - t0
    # the batched (mixed, FP64 endgame) relaxations against the FP64-only ones
    frel = max((abs(r["f"] - f) / abs(f) for r, f in zip(reps, f64)), default=None)
    sample_sites = sum(s["n_pd"] + s["n_h"] for s in specs[:k])
    return {
        "metric": "relaxations/s", "value": len(specs) / wall, "unit": "relaxations/s", "n_gpus": 1,
        "steps": len(specs), "warmup": 0, "ms_per_step": wall * 1e3 / len(specs), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32-bulk/f64-reduce", "data": "synthetic",
        "config": {"workload": "64 lattices, edge U{8..24} cells, T U(300,700) K, x U(0.01,0.6)+-0.05, "
                               "disp N(0,0.02), sigma 0.10, periodic; lockstep BLMVM to ||Pg||<=1e-8",
                   "lattices": len(specs), "sites": sites, "lmvm_m": args.lmvm_m},
        "converged": conv, "admission_s": t_adm, "relax_s": t_relax,
        "iterations": [r["iterations"] for r in reps], "evals": sum(r["evals"] for r in reps),
        "evals_fp64": sum(r["evals_fp64"] for r in reps),
        "fp64_one_at_a_time": {"lattices": k, "sites": sample_sites, "wall_s": t64,
                               "relaxations_per_s": k / t64,
                               "sites_per_s": sample_sites / t64},
        "relaxed_f_rel_vs_fp64": frel,
        "batched_sites_per_s": sites / wall,
    }


def plain(args):
    """The plain method's case: FCC Pd 6^3 cells (864 Pd + 864 octahedral H,
    occupancy 0.10), periodic, T = 300 K, 3-point-per-axis Gauss-Hermite
    rule, displacements N(0, 0.02 A), sigma 0.10 A: one FP64 objective +
    gradient evaluation, then an LMVM relaxation to ||Pg|| <= 1e-8, in FP64
    and in mixed precision (FP64 endgame).  Reports ms per relaxation (the
    median of --steps repeats from the same start)."""
    import a2c_amd as A
    X, x, box, npd, nh = A.lattice(6, seed=1, disp_std=0.02, sigma=(0.10, 0.10), x=(0.10, 0.10))
    out = {}
    for prec in ("fp64", "mixed"):
        fe = A.FreeEntropy(npd, nh, box, (1, 1, 1), T=300.0, precision=prec, quad="gh3")
        fe.set_occupancy(x)
        fe.findNeighbors(X)
        t0 = time.perf_counter()
        f0, _ = fe.eval(X)
        t_eval = time.perf_counter() - t0
        fe.minimizeFreeEntropy(X, gatol=1e-8, grtol=0.0)  # warm
        times = []
        for _ in range(max(1, min(args.steps, 20))):
            t0 = time.perf_counter()
            Xr, rep = fe.minimizeFreeEntropy(X, gatol=1e-8, grtol=0.0)
            times.append(time.perf_counter() - t0)
        fe.close()
        out[prec] = {"ms": 1e3 * float(np.median(times)), "iterations": rep["iterations"], "evals": rep["evals"],
                     "evals_fp64": rep["evals_fp64"], "reason": rep["reason"], "gnorm": rep["gnorm"],
                     "f0": f0, "f": rep["f"], "first_eval_ms": 1e3 * t_eval}
        log("plain %s: %.2f ms, %d iterations" % (prec, out[prec]["ms"], rep["iterations"]))
    return {
        "metric": "ms/relaxation", "value": out["mixed"]["ms"], "unit": "ms/relaxation", "n_gpus": 1,
        "steps": max(1, min(args.steps, 20)), "warmup": 1, "ms_per_step": out["mixed"]["ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32-bulk/f64-reduce",
        "data": "synthetic",
        "config": {"workload": "plain method: FCC Pd 6^3 cells + 864 H at x=0.10, periodic, T=300K, GH3 rule, "
                               "disp N(0,0.02), sigma 0.10; FP64 eval + LMVM to ||Pg||<=1e-8", "sites": npd + nh},
        "fp64": out["fp64"], "mixed": out["mixed"],
    }


def dmd(args):
    """Coupled diffusive MD: n^3 cells (default 64^3 = 2,097,152 sites), H
    occupancy 0.6 in the first 8 cell layers along x, xLb = 1e-3 elsewhere,
    T = 400 K, nu = 1e13 /s, Qm = 0.23 eV, dt = 5e-13 s (12 rate dt < 0.1);
    each step: relaxation (mixed, FP64 endgame) -> chemical potential ->
    diffusion (FP64 flux).  Reports s/step and the balance of total H per
    step: the net flux of the diffusion update (zero up to the rounding of the
    per-site sums, relative to the summed |edge terms|), the H added by the
    xLb clamp (integrate.cpp:79-81), and the change of total H, which equals
    the clamp lift when the flux balances."""
    import a2c_amd as A
    from a2c_amd import workloads
    n = args.dmd_cells
    s = workloads.dmd(n)
    X, x, npd, nh = s["X"], s["x"], s["n_pd"], s["n_h"]
    fe = A.FreeEntropy(npd, nh, s["box"], s["periodic"], T=s["T"], precision="mixed")
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    lmvm = workloads.dmd_lmvm(npd + nh)
    prm = A.dmd_params(**workloads.DMD_STEP, lmvm=lmvm)
    g = np.zeros(nh)
    fh = np.zeros(nh, np.int32)
    h0 = float(np.sum(x))
    reps = []
    t0 = time.perf_counter()
    for st in range(args.dmd_steps):
        reps.append(A.dmd_step(fe, prm, X, x, g, fh))
        r = reps[-1]
        sg = X.reshape(-1, 4)[:, 3]
        log("dmd step %d: f=%.12g it=%d reason=%d dH=%.6e lift=%.6e flux=%.2e/%.2e sigma %.4f..%.4f" % (
            st, r["f"], r["lmvm_iterations"], r["lmvm_reason"], r["total_h_after"] - r["total_h_before"],
            r["clamp_lift"], r["flux_sum"], r["flux_gross"], sg.min(), sg.max()))
    wall = time.perf_counter() - t0
    # the diffusion kernels alone at this size and the last step's occupancy
    diff_kernels = diffusion_timing(fe, x, nh)
    diff_ms = sum(diff_kernels.values())
    fe.close()
    dH = [r["total_h_after"] - r["total_h_before"] for r in reps]
    flux_rel = [abs(r["flux_sum"] - r["reservoir_supply"]) / r["flux_gross"] for r in reps if r["flux_gross"] > 0]
    resid = [abs(d - r["clamp_lift"] - r["reservoir_supply"]) for d, r in zip(dH, reps)]
    return {
        "metric": "s/step", "value": wall / args.dmd_steps, "unit": "s/step", "n_gpus": 1,
        "steps": args.dmd_steps, "warmup": 0, "ms_per_step": wall * 1e3 / args.dmd_steps,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32-bulk/f64-reduce",
        "data": "synthetic",
        "config": {"workload": "coupled DMD, %d^3 cells, slab x=0.6 in 8 layers, x=1e-3 elsewhere, T=400K, "
                               "dt=5e-13 s, nu=1e13/s, Qm=0.23 eV" % n, "sites": npd + nh},
        "total_H_start": h0, "total_H_end": float(np.sum(x)),
        "balance": {
            "flux_rel_max": max(flux_rel) if flux_rel else None,
            "flux_note": "|sum of applied increments - reservoir supply| / sum of |edge terms| per step "
                         "(exact antisymmetry of the edge terms: zero up to the rounding of the site sums)",
            "clamp_lift_per_step": [r["clamp_lift"] for r in reps],
            "dH_per_step": dH,
            "dH_minus_lift_max": max(resid),
            "dH_note": "total H (compensated sums) changes by the H the xLb clamp adds to sites that lose H; "
                       "no reservoirs in this workload",
        },
        "gatol": lmvm["gatol"], "sigma_ub": lmvm["sigma_ub"],
        "lmvm_iterations": [r["lmvm_iterations"] for r in reps],
        "lmvm_reasons": sorted(set(r["lmvm_reason"] for r in reps)),
        "diffusion_hbm": {"bytes": 76.0 * nh, "ms": diff_ms, "GB/s": 76.0 * nh / (diff_ms * 1e-3) / 1e9,
                          "frac": 76.0 * nh / (diff_ms * 1e-3) / 8e12, "kernels_ms": diff_kernels,
                          "note": "76 B per H site (x and gamma, 12 HH1 indices, new x, full_h) against 8 TB/s"},
    }


def batched_block(args):
    """The batched workload's relaxations/s, reported in the default line."""
    r = batched(args)
    return {"metric": "relaxations/s", "value": r["value"], "unit": "relaxations/s",
            "lattices": r["config"]["lattices"], "sites": r["config"]["sites"], "converged": r["converged"],
            "admission_s": r["admission_s"], "relax_s": r["relax_s"], "evals": r["evals"],
            "evals_fp64": r["evals_fp64"], "iterations_max": max(r["iterations"]),
            "relaxed_f_rel_vs_fp64": r["relaxed_f_rel_vs_fp64"],
            "fp64_one_at_a_time": r["fp64_one_at_a_time"],
            "config": "64 lattices, edge U{8..24} cells, T U(300,700) K, x U(0.01,0.6)+-0.05, disp N(0,0.02), "
                      "sigma 0.10, periodic, 12-point rule; lockstep BLMVM (m=5) to ||Pg||<=1e-8, mixed precision "
                      "with the FP64 endgame; admission (lists, optimizer memory) inside the timed wall"}


def serving_block(args):
    """Continuous-batching serving over a Poisson rate sweep (a2c_amd.serve)."""
    from a2c_amd import serve
    rates = [float(r) for r in args.serve_rates.split(",")]

    def note(r):
        if "summary" in r:
            sm = r["summary"]
            log("serving %g req/s: %.2f relaxations/s, p50 %.3f s, p99 %.3f s" % (
                r["rate"], sm["relaxations_per_s"], sm["p50_s"], sm["p99_s"]))
    out = serve.sweep(rates, args.serve_requests, args.serve_fp64_sample, log=note, policy=args.serve_policy)
    out["metric"] = "relaxations/s + p50/p99 latency"
    out["config"] = ("Poisson arrivals, edge log-uniform [6,48] cells, T U(300,700) K, x U(0,0.6)+-0.05, "
                     "disp N(0,0.02), sigma 0.10, periodic; continuous batching (admission by HBM footprint, "
                     "'%s' scheduling), mixed precision with the FP64 endgame, ||Pg||<=1e-8; comparator: the same "
                     "requests relaxed one at a time in FP64, replayed first-come-first-served on each rate's "
                     "arrivals" % args.serve_policy)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        print(json.dumps(batched(args)), flush=True)
        return
    if args.workload == "dmd":
        print(json.dumps(dmd(args)), flush=True)
        return
    if args.workload == "plain":
        print(json.dumps(plain(args)), flush=True)
        return
    import a2c_amd as A

    from a2c_amd import workloads

    def mk(n):
        s = workloads.single(n)
        return s["X"], s["x"], s["box"], s["n_pd"], s["n_h"]

    X, x, box, npd, nh = mk(args.cells)
    log("workload: %d cells^3, %d sites" % (args.cells, npd + nh))
    m = measure(A, args, X, x, box, npd, nh, args.quad, args.steps, args.warmup, calib=args.calib, paths=True)
    N = npd + nh
    result = {
        "metric": "free-entropy+grad atom-evals/s",
        "value": m["value"],
        "unit": "atom-evals/s",
        "n_gpus": 1,
        "steps": m["steps"],
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-bulk/f64-reduce" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "single large lattice: FCC Pd 40^3 cells + octahedral H, periodic, "
                               "x~U(0,0.6), disp N(0,0.03), sigma U(0.08,0.15), T=500K, %s rule"
                               % {"r12": "R12", "gh3": "GH3"}[args.quad],
                   "cells": args.cells, "sites": N, "precision": args.precision, "quad": args.quad},
        "roofline": m["roofline"],
        "kernels_ms": m["kernels_ms"],
        "fp64_gpu": m["fp64_gpu"],
        "error_vs_fp64": m["error_vs_fp64"],
        "neighbor_build_ms": m["neighbor_build_ms"],
        "neighbor_rebuild_device_ms": m["neighbor_rebuild_device_ms"],
        "neighbor_skin_rebuild_device_ms": m["neighbor_skin_rebuild_device_ms"],
        "density_pass_flop": m["density_pass_flop"],
        "wall_s_timed": m["wall_s_timed"],
        "hbm_paths": m["hbm_paths"],
    }
    if args.quad == "r12" and not args.no_gh3:
        # the other rule on the same lattice, reported beside the headline
        g = measure(A, args, X, x, box, npd, nh, "gh3", max(10, args.steps // 2), args.warmup)
        result["gh3"] = {k: g[k] for k in ("value", "ms_per_step", "steps", "kernels_ms", "fp64_gpu",
                                           "error_vs_fp64")}
        result["gh3"]["roofline"] = g["roofline"]
    if not args.no_batched:
        log("batched block")
        result["batched"] = batched_block(args)
    if not args.no_serving:
        log("serving block")
        result["serving"] = serving_block(args)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(args, mk)
    print(json.dumps(result), flush=True)


def diffusion_timing(fe, x, nh, iters=50):
    """The diffusion step of this lattice's HH1 graph on device arrays, HIP
    events per kernel: gamma near the equilibrium of the occupancy
    (ln(x/(1-x)) + N(0, 0.05)), no reservoirs, the DMD workload's rate and
    time step."""
    dgam, dfull, dxo = fe.dev_alloc(8 * nh), fe.dev_alloc(4 * nh), fe.dev_alloc(8 * nh)
    try:
        xc = np.clip(x, 1e-12, 1.0 - 1e-12)
        gam = np.log(xc / (1.0 - xc)) + np.random.default_rng(12).normal(0.0, 0.05, nh)
        fe.upload(dgam, gam)
        fe.upload(dfull, np.zeros(nh, np.int32))
        args = (5e-13, 1e13, 0.23, 1e-3, 1.0 - 1e-16, dgam, dfull, dxo)
        fe.time_diffuse(*args, 3)  # warm-up
        tm = fe.time_diffuse(*args, iters)
    finally:
        for p in (dgam, dfull, dxo):
            fe.dev_free(p)
    return tm["kernels_ms"]


def hbm_paths(N, nh, P, eval_ms, nbr_ms, diff_kernels):
    """Algorithmic HBM bytes of the three paths against 8 TB/s: evaluation
    8 B x P (two passes over 4-byte neighbor indices) + 56 B x N + 4 B x nH;
    neighbor build 24 B x N + 4 B x P + 4 B x (N + 1); diffusion step 76 B x
    nH (x and gamma, 12 HH1 indices, new x, full_h).  P = directed pairs."""
    diff_ms = sum(diff_kernels.values())
    rows = {"evaluation": (8.0 * P + 56.0 * N + 4.0 * nh, eval_ms, "whole evaluation (HIP events)"),
            "neighbor_rebuild": (24.0 * N + 4.0 * P + 4.0 * (N + 1), nbr_ms,
                                 "filtered rebuild from device positions, wall with its one host round trip"),
            "diffusion": (76.0 * nh, diff_ms, "k_diff_f + k_diff_step + k_diff_balance (HIP events)")}
    out = {}
    for k, (b, ms, what) in rows.items():
        gbs = b / (ms * 1e-3) / 1e9
        out[k] = {"bytes": b, "ms": ms, "GB/s": gbs, "frac": gbs / 8000.0, "timed": what}
    out["diffusion"]["kernels_ms"] = diff_kernels
    return out


def measure(A, args, X, x, box, npd, nh, quad, steps, warmup, calib=False, paths=False):
    """One rule on one lattice: the FP64 kernel (comparator), then the
    measured precision; returns throughput, per-kernel times, roofline of
    the gradient owner kernel and the error against the FP64 kernel (and,
    with paths, the HBM fractions of evaluation, neighbor build and
    diffusion step)."""
    N = npd + nh
    fe = A.FreeEntropy(npd, nh, box, (1, 1, 1), T=500.0, precision="fp64", quad=quad)
    fe.set_occupancy(x)
    t0 = time.perf_counter()
    st = fe.findNeighbors(X)
    nbr_ms = (time.perf_counter() - t0) * 1e3
    U = (st["total"][0] / 2.0, float(st["total"][1]), st["total"][3] / 2.0)  # unordered PdPd, PdH, HH
    K = POINTS[quad]
    flop_grad = K * sum(c * u for c, u in zip(FLOP_GRAD, U))
    flop_dens = K * sum(c * u for c, u in zip(FLOP_DENS, U))
    # rebuilds from device-resident positions (the DMD loop's case), wall
    # clock including the build's host synchronisations: after moves like a
    # DMD step's (every coordinate by N(0, 0.005 A) per rebuild, so the sites
    # stay within skin/2 of the skin lists: the lists are filtered from them),
    # and after a move beyond that (N(0, 0.2 A): the cells are walked at
    # rc + skin first).  Positions are uploaded outside the timed region.
    dXn = fe.dev_alloc(X.nbytes)
    rng = np.random.default_rng(11)
    Xm = X.copy().reshape(-1, 4)
    fe.upload(dXn, Xm.reshape(-1))
    fe.findNeighbors_device(dXn)  # the skin lists (second build)
    reb, skin_reb = [], []
    for _ in range(5):
        Xm[:, :3] += rng.normal(0.0, 0.005, size=(N, 3))
        fe.upload(dXn, Xm.reshape(-1))
        t0 = time.perf_counter()
        sd = fe.findNeighbors_device(dXn)
        reb.append((time.perf_counter() - t0) * 1e3)
        assert sd["refreshed"]
    for _ in range(3):
        Xj = Xm.copy()
        Xj[:, :3] += rng.normal(0.0, 0.2, size=(N, 3))
        fe.upload(dXn, Xj.reshape(-1))
        t0 = time.perf_counter()
        fe.findNeighbors_device(dXn)
        skin_reb.append((time.perf_counter() - t0) * 1e3)
        fe.upload(dXn, X)
        fe.findNeighbors_device(dXn)  # back to the specimen for the timed evaluations
    fe.dev_free(dXn)
    diff_kernels = diffusion_timing(fe, x, nh) if paths else None
    nbr_dev_ms = float(np.median(reb))
    nbr_skin_ms = float(np.median(skin_reb))
    log("%s neighbors: %.1f ms first build from host arrays; rebuilds from device: %.2f ms filtered from the "
        "skin lists (median of 5), %.2f ms with new skin lists (median of 3); max %d, totals %s" % (
            quad, nbr_ms, nbr_dev_ms, nbr_skin_ms, st["max_neib"], st["total"]))
    if calib:
        nb = 256 << 20
        a, b = fe.dev_alloc(nb), fe.dev_alloc(nb)
        fe.dev_copy(b, a, nb)
        fe.dev_free(a)
        fe.dev_free(b)

    # FP64 comparator (always measured; the reference answer for the error)
    s64 = max(3, steps // 10)
    tm64, wall64, f64, G64 = run_gpu(fe, X, s64, 2)
    ms64 = tm64["total_ms"] / s64
    log("%s fp64: %.3f ms/eval %s" % (quad, ms64, tm64["kernels_ms"]))
    if args.precision == "mixed":
        fe.set_precision("mixed")
        tm, wall, f, G = run_gpu(fe, X, steps, warmup)
        # pair-once default: the owner kernel carries every pair term; full-list mode: grad32
        gkern = "grad_own" if "grad_own" in tm["kernels_ms"] else "grad32"
    else:
        tm, wall, f, G = tm64, wall64, f64, G64
        steps = s64
        gkern = "grad_own64" if "grad_own64" in tm["kernels_ms"] else "grad64"
    fe.close()
    kern = tm["kernels_ms"]
    gms = kern.get(gkern, float("nan"))
    achieved = flop_grad / (gms * 1e-3) / 1e12
    traffic = None  # HBM bytes per dispatch of gkern from the committed PMC summary (FETCH/WRITE_SIZE, scaled)
    traffic_src = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        tr = json.load(open(tpath))
        if tr.get("kernel") == gkern and tr.get("quad", "r12") == quad:
            traffic = tr["bytes_per_dispatch"]
            traffic_src = "copied from %s (rocprofv3 FETCH_SIZE/WRITE_SIZE, scaled by the copy calibration)" % (
                tr.get("source", "profiles/pmc_traffic.json"))
    log("%s %s: %.3f ms/eval %s" % (quad, args.precision, tm["total_ms"] / steps, kern))
    return {
        "value": N * steps / (tm["total_ms"] * 1e-3),
        "steps": steps,
        "ms_per_step": tm["total_ms"] / steps,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK[args.precision], "unit": "TFLOP/s",
                     "frac": achieved / PEAK[args.precision], "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": gkern,
                     "kernel_ms": gms, "algorithmic_flop": flop_grad},
        "kernels_ms": kern,
        "fp64_gpu": {"value": N / (ms64 * 1e-3), "ms_per_step": ms64, "kernels_ms": tm64["kernels_ms"]},
        "error_vs_fp64": {"objective_rel": abs(f - f64) / abs(f64),
                          "gradient_rel_l2": float(np.linalg.norm(G - G64) / np.linalg.norm(G64))},
        "neighbor_build_ms": nbr_ms,
        "neighbor_rebuild_device_ms": nbr_dev_ms,
        "neighbor_skin_rebuild_device_ms": nbr_skin_ms,
        "density_pass_flop": flop_dens,
        "wall_s_timed": wall,
        "hbm_paths": hbm_paths(N, nh, float(sum(st["total"])), tm["total_ms"] / steps, nbr_dev_ms, diff_kernels)
        if paths else None,
    }


if __name__ == "__main__":
    main()
