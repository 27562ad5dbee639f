#!/usr/bin/env python3
"""Throughput of the tonal Farassat 1A hot path on one MI355X.

Workload: the four-blade rotor in forward flight -- 4 blades x 20,000 panels (100 chordwise x 100
spanwise per side), 1024 source steps per revolution, 1024 observers on a 20 m hemisphere,
oM = 250 rad/s, Minf = 0.1 at -5 deg: 8.39e10 source-observer-time evaluations per pass.
Synthetic, seeded input in the reference's layout, preprocessed on the host before timing; the
element data is resident in HBM when the timed region starts.

A step is one full pass of the mixed FP32/FP64 path: clear partials, eval + binning kernel,
ordered Kahan reduction to pT/pL/pA.  The plain FP64 path is timed the same way and reported
beside it.  Accuracy of the timed outputs is checked against the FP64 CPU oracle on the first
observers (that oracle run is also the CPU baseline).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--oracle-obs N] [--no-fp64]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_EVAL = 100.0          # algorithmic count per (element, step, observer)
FLOP_FP32, FLOP_FP64 = 70.0, 30.0
PEAK_FP32, PEAK_FP64 = 157.3e12, 78.6e12   # MI355X vector peaks, FLOP/s
HBM_PEAK = 8.0e12                           # B/s


def mixed_peak():
    # FLOP/s at which 70 FP32 + 30 FP64 FLOP per eval would run at both vector peaks
    return FLOP_PER_EVAL / (FLOP_FP32 / PEAK_FP32 + FLOP_FP64 / PEAK_FP64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--oracle-obs", type=int, default=32)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay the pass as a captured HIP graph")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(observer sharding is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        return 2

    import acum3_periodic_amd as acum

    t0 = time.time()
    run = acum.Run.synth(acum.synth_params(2)).preprocess()
    case = run.tonal_case(acum.MIXED)
    evals = 0
    E = [case.E[s] for s in range(case.nSurf)]
    evals = sum(E) * case.nTime * case.nObs
    host_prep_s = time.time() - t0

    ctx = acum.Context(0)
    t0 = time.time()
    h = ctx.stage([case])
    stage_s = time.time() - t0

    def timed(precision, steps, warmup):
        for _ in range(warmup):
            ctx.run(h, precision, repeats=1, graph=args.graph)
        eval_ms = red_ms = 0.0
        t = time.perf_counter()
        for _ in range(steps):
            tm = ctx.run(h, precision, repeats=1, graph=args.graph)
            eval_ms += tm.eval_ms
            red_ms += tm.reduce_ms
        dt = time.perf_counter() - t
        tm = ctx.timing()
        return dt / steps, eval_ms / steps, red_ms / steps, tm

    step_s, eval_ms, red_ms, tm = timed(acum.MIXED, args.steps, args.warmup)
    mixed = ctx.fetch(h, 0, case.nTime, case.nObs)
    partials = tm.n_partials
    items = tm.n_items

    fp64 = None
    if not args.no_fp64:
        s64, e64, r64, t64 = timed(acum.FP64, max(1, min(args.steps, 3)), 1)
        plain = ctx.fetch(h, 0, case.nTime, case.nObs)
        fp64 = dict(value=evals / s64 / 1e9, ms_per_step=s64 * 1e3, eval_ms=e64, reduce_ms=r64,
                    partials=t64.n_partials)

    # CPU baseline + accuracy: FP64 oracle on the first observers (observers are independent)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O  # tests/oracle.py: the oracle binding (CPU baseline leg only)
    nob = max(1, min(args.oracle_obs, case.nObs))
    sub = acum.TonalCase()
    C.pointer(sub)[0] = case
    sub.nObs = nob
    sub.precision = acum.FP64
    threads = len(os.sched_getaffinity(0))
    threads = min(threads, 16)
    t = time.perf_counter()
    oT, oL, oA = O.tonal_f64(sub, threads)
    cpu_s = time.perf_counter() - t
    cpu_evals = sum(E) * case.nTime * nob

    gA = mixed[2][:, :nob]
    d_oaspl = np.abs(acum.oaspl(gA) - acum.oaspl(oA.astype(np.float32)))
    acc = dict(observers=nob, mixed_vs_oracle_rel_l2_pA=O.rel_l2(gA, oA),
               mixed_vs_oracle_max_doaspl_db=float(d_oaspl.max()))
    if fp64:
        acc["fp64_vs_oracle_rel_l2_pA"] = O.rel_l2(plain[2][:, :nob], oA)
        d = np.abs(acum.oaspl(mixed[2]) - acum.oaspl(plain[2]))
        acc["mixed_vs_fp64_all_obs_max_doaspl_db"] = float(d.max())

    flops = FLOP_PER_EVAL * evals
    achieved = flops / (eval_ms * 1e-3)
    peak = mixed_peak()
    bin_bytes = partials * case.nTime * case.nObs * 2 * 4 + case.nTime * case.nObs * 12
    value = evals / step_s / 1e9
    out = {
        "metric": "G src-obs-time evals/s, FP32-mixed vs FP64 at <=0.01 dB OASPL; %HBM/VALU peak",
        "value": round(value, 3),
        "unit": "G src-obs-time evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "tonal Farassat 1A, 4-blade rotor in forward flight: 4 x 20,000 "
                               "elements, 1024 steps, 1024 observers (8.39e10 evals)",
                   "evals_per_step": evals, "precision": "mixed", "graph": bool(args.graph),
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3),
                     "peak": round(peak / 1e12, 3), "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": None,
                     "kernel": "tonal_eval_kernel<mixed>", "eval_ms": round(eval_ms, 3),
                     "flop_per_eval": FLOP_PER_EVAL},
        "cpu_baseline": {"value": round(cpu_evals / cpu_s / 1e9, 4), "unit": "G src-obs-time evals/s",
                         "cores": threads, "kind": "port",
                         "sample": f"FP64 oracle, same workload, first {nob} observers "
                                   f"({cpu_evals:.3g} evals, {cpu_s:.1f} s)"},
        "binning": {"kernel": "tonal_reduce_kernel<float>", "ms": round(red_ms, 4),
                    "bytes": bin_bytes, "GBps": round(bin_bytes / (red_ms * 1e-3) / 1e9, 1),
                    "frac_hbm": round(bin_bytes / (red_ms * 1e-3) / HBM_PEAK, 4),
                    "partials": partials},
        "fp64_path": fp64,
        "mixed_over_fp64": round(value / fp64["value"], 3) if fp64 else None,
        "accuracy": acc,
        "work_items": items,
        "host_prep_s": round(host_prep_s, 2),
        "stage_s": round(stage_s, 2),
    }
    ctx.release(h)
    ctx.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
