#!/usr/bin/env python3
"""Throughput of the tonal Farassat 1A hot path on one MI355X, with the broadband, batched and stress
workloads beside it.

Headline workload: the four-blade rotor in forward flight -- 4 blades x 20,000 panels (100
chordwise x 100 spanwise per side), 1024 source steps per revolution, 1024 observers on a 20 m
hemisphere, oM = 250 rad/s, Minf = 0.1 at -5 deg: 8.39e10 source-observer-time evaluations per
pass.  Synthetic, seeded input in the reference's layout, preprocessed on the host before timing;
the element data is resident in HBM when the timed region starts.

A step is one full pass of the mixed FP32/FP64 path: clear partials, eval + binning kernel,
ordered Kahan reduction to pT/pL/pA.  Reported beside it, each in its own JSON key:
  fp64_path  the plain FP64 path on the same case;
  accuracy   both paths against the FP64 CPU oracle on 32 observers spread over the hemisphere:
             rel-L2, OASPL and 1/3-octave band levels (that oracle run, median of three, is also
             the CPU baseline; the float instance of the same restatement is timed beside it);
  bpm        the broadband path (4 x 60 sections, 360 steps, 1024 observers, 34 bands), against
             the FP64 oracle, with the CPU restatement on all threads and on one;
  batched    the first 8 of the 64 seeded uneven cases, 4 per persistent launch, tile shapes per
             case from the autotuner, HIP graph replay, the FP64 path on the same launches, every
             case equal to it run alone and checked against the oracle;
  stress     the coaxial 2 x 4-blade case (3.36e12 evals), against the oracle on 16 observers.
The bench exits with status 3 (after printing its line) when a parity bar is broken: mixed-path
OASPL or a band level more than 0.01 dB from the oracle, FP64 rel-L2 above 1e-6, broadband levels
more than 0.01 dB off.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--oracle-obs N] [--no-fp64] [--no-bpm]
                    [--no-batched] [--no-stress] [--cases N --wave M]
    python bench.py --workload batched|stress|bpm          # one leg alone (batched: --cases of 64)
    python bench.py --summarize-pmc FETCH_DIR WRITE_DIR      # HBM traffic summary into profiles/
    python bench.py [--workload bpm] --summarize-counters DIR...  # instruction mix into profiles/
    python bench.py [--workload bpm] --summarize-stats DIR     # kernel times into profiles/

Roofline of the eval kernel: 100 FLOP per eval (70 FP32 + 30 FP64 in the mixed path) against
the MI355X vector peaks (it is bound by vector-ALU issue); `traffic` is the kernel's measured HBM
bytes per launch, read from the committed rocprofv3 counter summary
(profiles/tonal_eval_hbm_traffic_rotor_4x20000_elements_1024_steps_1024_observers.md) when present.
"""
import argparse
import ctypes as C
import glob
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_EVAL = 100.0          # algorithmic count per (element, step, observer)
FLOP_FP32, FLOP_FP64 = 70.0, 30.0
PEAK_FP32, PEAK_FP64 = 157.3e12, 78.6e12   # MI355X vector peaks, FLOP/s
HBM_PEAK = 8.0e12                           # B/s
# ISO 1/3-octave band centres, 10 Hz - 20 kHz (the bands of the tonal spectra leg)
ISO_BANDS = (10, 12.5, 16, 20, 25, 31.5, 40, 50, 63, 80, 100, 125, 160, 200, 250, 315, 400, 500, 630, 800,
             1000, 1250, 1600, 2000, 2500, 3150, 4000, 5000, 6300, 8000, 10000, 12500, 16000, 20000)
# workloads the committed profiles cover, as they appear in the profiles' file names
TONAL_PROFILE = "rotor_4x20000_elements_1024_steps_1024_observers"
BPM_PROFILE = "4x60_sections_360_steps_1024_observers_34_bands"
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", f"tonal_eval_hbm_traffic_{TONAL_PROFILE}.md")
# the rotor leg alone with the plain FP64 path (the profiled runs leave the other legs out)
TONAL_ONLY = "--no-bpm --no-batched --no-stress"
EVAL_KERNEL = "tonal_eval_kernel"
PLAIN_KERNEL = "tonal_plain_kernel"
CALIB_BYTES = 1 << 30


def mixed_peak():
    # FLOP/s at which 70 FP32 + 30 FP64 FLOP per eval would run at both vector peaks
    return FLOP_PER_EVAL / (FLOP_FP32 / PEAK_FP32 + FLOP_FP64 / PEAK_FP64)


def copy_calibration():
    """A plain device-to-device copy of CALIB_BYTES (reads and writes CALIB_BYTES) through the HIP
    runtime, for scaling the FETCH_SIZE / WRITE_SIZE counters; runs only when profiling."""
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    ok = hip.hipMalloc(C.byref(src), C.c_size_t(CALIB_BYTES)) == 0
    ok = ok and hip.hipMalloc(C.byref(dst), C.c_size_t(CALIB_BYTES)) == 0
    ok = ok and hip.hipMemset(src, 1, C.c_size_t(CALIB_BYTES)) == 0
    ok = ok and hip.hipDeviceSynchronize() == 0
    ok = ok and hip.hipMemcpy(dst, src, C.c_size_t(CALIB_BYTES), 3) == 0  # device to device
    ok = ok and hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    if not ok:
        raise RuntimeError("calibration copy failed")
    print(f"calibration copy of {CALIB_BYTES} bytes done", file=sys.stderr, flush=True)


def read_traffic():
    """Bytes per eval-kernel launch from the committed counter summary, or None."""
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(TRAFFIC_SUMMARY).read())
    return float(m.group(1)) if m else None


def summarize_pmc(fetch_dir, write_dir):
    """Turns two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; KiB per dispatch) of
    `bench.py --steps 1 --warmup 0 --no-fp64 --no-bpm --no-batched --no-stress --oracle-obs 1
    --copy-calibration` into the traffic summary (the mixed eval kernel, and the plain FP64
    kernel), scaling each counter by its reading for the calibration copy."""
    import csv

    def load(d, counter):
        per = {}  # dispatch id -> [kernel name, summed value]
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r.get("Counter_Name") == counter:
                    e = per.setdefault(int(r["Dispatch_Id"]), [r["Kernel_Name"], 0.0])
                    e[1] += float(r["Counter_Value"])
        if not per:
            raise SystemExit(f"no {counter} rows under {d}")
        return [tuple(per[k]) for k in sorted(per)]

    out = {}
    for name, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = load(d, name)
        # the calibration copy is the first copy dispatch of the process (before any staging)
        copy = [v for k, v in rows if "copy" in k.lower()]
        ev = [v for k, v in rows if EVAL_KERNEL in k]
        pl = [v for k, v in rows if PLAIN_KERNEL in k]
        if not copy or not ev:
            raise SystemExit(f"{name}: calibration copy or eval kernel missing in {d}")
        copy_kib = copy[0]
        scale = (CALIB_BYTES / 1024.0) / copy_kib
        out[name] = dict(copy_kib=copy_kib, scale=scale, eval_kib=float(np.mean(ev)),
                         eval_bytes=float(np.mean(ev)) * 1024.0 * scale,
                         plain_kib=float(np.mean(pl)) if pl else None,
                         plain_bytes=float(np.mean(pl)) * 1024.0 * scale if pl else None)
    total = out["FETCH_SIZE"]["eval_bytes"] + out["WRITE_SIZE"]["eval_bytes"]
    lines = [
        "# HBM traffic of the tonal eval kernel: four-blade rotor (4 x 20,000 elements, 1024 steps,",
        "# 1024 observers), mixed path, one launch (and the plain FP64 kernel's, one launch)",
        "",
        "Commands (repository root `R`, output directory `OUT`, on the GPU host, one counter per run):",
        "",
        "```",
        "cd /tmp && export TMPDIR=/tmp",
        "rocprofv3 --pmc FETCH_SIZE --output-format csv -d $OUT/pmcF -o p -- \\",
        f"    python3 $R/bench.py --steps 1 --warmup 0 --no-fp64 {TONAL_ONLY} --oracle-obs 1 --copy-calibration",
        "rocprofv3 --pmc WRITE_SIZE --output-format csv -d $OUT/pmcW -o p -- \\",
        f"    python3 $R/bench.py --steps 1 --warmup 0 --no-fp64 {TONAL_ONLY} --oracle-obs 1 --copy-calibration",
        "python3 $R/bench.py --summarize-pmc $OUT/pmcF $OUT/pmcW",
        "```",
        "",
        f"Calibration: a device copy of {CALIB_BYTES} bytes (reads and writes that many bytes).",
        "",
        "| counter | copy (KiB) | scale (bytes/KiB-count) | eval kernel (KiB) | eval kernel (bytes) |",
        "|---|---|---|---|---|",
    ]
    for name in ("FETCH_SIZE", "WRITE_SIZE"):
        o = out[name]
        lines.append(f"| {name} | {o['copy_kib']:.0f} | {o['scale'] * 1024:.4f} | {o['eval_kib']:.0f} | "
                     f"{o['eval_bytes']:.4g} |")
    lines += ["", f"traffic_bytes_per_launch: {total:.6g}", ""]
    if out["FETCH_SIZE"]["plain_bytes"] is not None and out["WRITE_SIZE"]["plain_bytes"] is not None:
        pf, pw = out["FETCH_SIZE"]["plain_bytes"], out["WRITE_SIZE"]["plain_bytes"]
        lines += [f"Plain FP64 kernel (`{PLAIN_KERNEL}`), same scale: FETCH {pf:.4g} bytes, WRITE {pw:.4g} "
                  f"bytes per launch.", "", f"plain_traffic_bytes_per_launch: {pf + pw:.6g}", ""]
    with open(TRAFFIC_SUMMARY, "w") as f:
        f.write("\n".join(lines))
    print("\n".join(lines))
    return 0


# Counter passes of the instruction-mix profiles (rocprofv3 takes at most 8 SQ counters and 2
# GRBM counters per run).
COUNTER_PASSES = {
    "fp64": ("SQ_INSTS_VALU", "SQ_INSTS_MFMA", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64",
             "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_TRANS_F64", "SQ_INSTS_VALU_CVT", "SQ_INSTS_VALU_INT32"),
    "fp32": ("SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_TRANS_F32",
             "SQ_INSTS_VALU_INT64", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_SMEM"),
    "cycles": ("SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_ANY",
               "SQ_WAIT_INST_ANY", "SQ_VALU_MFMA_BUSY_CYCLES", "SQ_INSTS_VMEM", "GRBM_GUI_ACTIVE", "GRBM_COUNT"),
}
# kernels the instruction-mix profiles report, and the work one launch does (for per-unit rows)
COUNTER_KERNELS = {
    "rotor": (("tonal_eval_kernel<true, true, false, 20>", "mixed eval", 80000 * 1024 * 1024 / 64, "wave-eval"),
              ("tonal_eval_kernel<false, true, false, 20>", "FP64 eval", 80000 * 1024 * 1024 / 64, "wave-eval"),
              ("tonal_plain_kernel", "plain FP64 kernel (the reference's structure)", 80000 * 1024 * 1024 / 64,
               "wave-eval")),
    "bpm": (("bpm_obs_kernel<17>", "broadband observer kernel", 240 * 360 * 1024 * 34 / 64, "wave band-eval"),),
}


def summarize_counters(dirs, workload="rotor"):
    """Per-launch averages of the counters of the rocprofv3 --pmc runs (one COUNTER_PASSES group
    each) under `dirs`, per unit of work, for the kernels of the workload (the four-blade rotor's
    two eval kernels, or the broadband observer kernel), into profiles/."""
    import csv
    per = {}  # kernel -> counter -> list of per-dispatch values
    for d in dirs:
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            disp = {}
            for r in csv.DictReader(open(f)):
                k = r["Kernel_Name"].replace("(anonymous namespace)::", "").replace("acum::", "")
                key = (int(r["Dispatch_Id"]), k, r["Counter_Name"])
                disp[key] = disp.get(key, 0.0) + float(r["Counter_Value"])
            for (_, k, c), v in disp.items():
                per.setdefault(k, {}).setdefault(c, []).append(v)
    if workload == "bpm":
        title = ("broadband observer kernel, 4 x 60 sections, 360 steps, 1024 observers, 34 bands "
                 "(one launch per call)")
        run = "python3 $R/bench.py --workload bpm --steps 1"
        out = os.path.join(ROOT, "profiles", f"bpm_obs_instruction_mix_{BPM_PROFILE}.md")
        summ = "python3 $R/bench.py --workload bpm --summarize-counters $OUT/*"
    else:
        title = ("tonal eval kernels, four-blade rotor (4 x 20,000 elements, 1024 steps, 1024 observers), "
                 "mixed and FP64 paths and the plain FP64 kernel, one launch each")
        run = f"python3 $R/bench.py --steps 1 --warmup 0 {TONAL_ONLY} --oracle-obs 1"
        out = os.path.join(ROOT, "profiles", f"tonal_eval_instruction_mix_{TONAL_PROFILE}.md")
        summ = "python3 $R/bench.py --summarize-counters $OUT/*"
    lines = [f"# Instruction mix and issue counters: {title}", "",
             "Commands (repository root `R`, output directory `OUT`, on the GPU host; one counter group per run):",
             "", "```", "cd /tmp && export TMPDIR=/tmp"]
    for g, cs in COUNTER_PASSES.items():
        lines.append(f"rocprofv3 --pmc {' '.join(cs)} --output-format csv -d $OUT/{g} -o p -- \\")
        lines.append(f"    {run}")
    lines += [summ, "```", "",
              "Values are per launch (averaged over the launches of each run) and per unit of work.",
              "SQ_WAVE_CYCLES, SQ_BUSY_CYCLES, SQ_ACTIVE_INST_* and SQ_WAIT_INST_ANY count in quad-cycles",
              "(units of 4 cycles), summed over waves (WAVE_CYCLES, ACTIVE, WAIT) or over the shader engines",
              "(BUSY); GRBM_GUI_ACTIVE is summed over the 8 XCDs. VALU issue share = ACTIVE_INST_VALU /",
              "WAVE_CYCLES x waves per SIMD: the fraction of a SIMD's cycles in which its VALU (or, two",
              "quad-cycles per instruction, its matrix core) issues.", ""]
    for kname, label, units, unit_name in COUNTER_KERNELS[workload]:
        ks = [k for k in per if kname in k]
        if not ks:
            continue
        cnt = {}
        for k in ks:
            for c, v in per[k].items():
                cnt.setdefault(c, []).extend(v)
        avg = {c: float(np.mean(v)) for c, v in cnt.items()}
        lines += [f"## {label}: `{kname}`", "", f"| counter | per launch | per {unit_name} |", "|---|---|---|"]
        for g in COUNTER_PASSES.values():
            for c in g:
                if c in avg:
                    lines.append(f"| {c} | {avg[c]:.4g} | {avg[c] / units:.3f} |")
        if "SQ_INSTS_VALU" in avg:
            f64 = sum(avg.get(f"SQ_INSTS_VALU_{t}_F64", 0.0) for t in ("ADD", "MUL", "FMA", "TRANS"))
            f32 = sum(avg.get(f"SQ_INSTS_VALU_{t}_F32", 0.0) for t in ("ADD", "MUL", "FMA", "TRANS"))
            lines += ["", f"FP64 arithmetic {f64 / units:.2f}, FP32 arithmetic {f32 / units:.2f}, "
                          f"conversions {avg.get('SQ_INSTS_VALU_CVT', 0) / units:.2f}, "
                          f"INT32 {avg.get('SQ_INSTS_VALU_INT32', 0) / units:.2f}, "
                          f"transcendental F64 {avg.get('SQ_INSTS_VALU_TRANS_F64', 0) / units:.2f} / "
                          f"F32 {avg.get('SQ_INSTS_VALU_TRANS_F32', 0) / units:.2f}, "
                          f"matrix-core {avg.get('SQ_INSTS_MFMA', 0) / units:.2f} instructions per {unit_name}."]
        if "SQ_WAVE_CYCLES" in avg and "SQ_ACTIVE_INST_VALU" in avg:
            lines += [f"VALU issue per wave cycle: {avg['SQ_ACTIVE_INST_VALU'] / avg['SQ_WAVE_CYCLES']:.3f} "
                      f"(times the resident waves per SIMD gives the SIMD's VALU issue share); waits "
                      f"{avg.get('SQ_WAIT_INST_ANY', 0) / avg['SQ_WAVE_CYCLES']:.3f} of wave cycles; "
                      f"matrix-core busy {avg.get('SQ_VALU_MFMA_BUSY_CYCLES', 0) / max(avg.get('GRBM_GUI_ACTIVE', 1), 1):.3g} "
                      f"per GPU cycle."]
        lines.append("")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))
    return 0


def summarize_stats(d, workload="rotor"):
    """Kernel time table of one rocprofv3 --kernel-trace --stats run of
    `bench.py --steps 3 --warmup 1 --no-bpm --no-batched --no-stress --oracle-obs 2` (both tonal
    precisions and the plain FP64 path), or with workload "bpm" of `bench.py --workload bpm
    --steps 3`, "stress" of `bench.py --workload stress --steps 2 --oracle-obs 16`, "batched" of
    `bench.py --workload batched --cases 8 --wave 4 --steps 2`."""
    import csv
    files = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no kernel_stats.csv under {d}")
    rows = list(csv.DictReader(open(files[0])))
    if workload == "bpm":
        head = ["# Kernel statistics: BPM broadband path, 4 x 60 sections, 360 steps, 1024 observers,",
                "# 34 bands (one warm-up and three timed calls)"]
        cmd = ["    python3 $R/bench.py --workload bpm --steps 3",
               "python3 $R/bench.py --workload bpm --summarize-stats $OUT/prof"]
        note_ = ["`bpm_obs_kernel<17>` covers the 34 bands in two passes of 17 within one launch per call.", ""]
        out = os.path.join(ROOT, "profiles", f"bpm_kernel_stats_{BPM_PROFILE}.md")
    elif workload == "stress":
        head = ["# Kernel statistics: stress case, coaxial 2 x 4 blades x 50,000 elements, 2048 steps,",
                "# 4096 observers, mixed path (one planning pass and two timed passes)"]
        cmd = ["    python3 $R/bench.py --workload stress --steps 2 --oracle-obs 16",
               "python3 $R/bench.py --workload stress --summarize-stats $OUT/prof"]
        note_ = ["One `tonal_eval_kernel<true, true, false, 20>` launch per pass covers the whole case "
                 "(3.36e12 evals).", ""]
        out = os.path.join(ROOT, "profiles", "tonal_kernel_stats_stress_coaxial_8x50000_elements_2048_steps_4096_observers.md")
    elif workload == "batched":
        head = ["# Kernel statistics: batched workload, first 8 of the 64 seeded uneven rotor cases, 4 per",
                "# persistent launch, mixed and FP64 passes, tile shapes per case from the autotuner"]
        cmd = ["    python3 $R/bench.py --workload batched --cases 8 --wave 4 --steps 2",
               "python3 $R/bench.py --workload batched --summarize-stats $OUT/prof"]
        note_ = ["The autotuner's candidate passes run on the first 512 observers of a case and are "
                 "counted here too; each timed batch pass is one eval launch per kernel variant in use, "
                 "replayed as a HIP graph.", ""]
        out = os.path.join(ROOT, "profiles", "tonal_kernel_stats_batched_first_8_seeded_cases.md")
    else:
        head = ["# Kernel statistics: tonal path, four-blade rotor (4 x 20,000 elements, 1024 steps,",
                "# 1024 observers), mixed and FP64 passes"]
        cmd = [f"    python3 $R/bench.py --steps 3 --warmup 1 {TONAL_ONLY} --oracle-obs 2",
               "python3 $R/bench.py --summarize-stats $OUT/prof"]
        note_ = ["`tonal_eval_kernel<true, ...>` is the mixed path, `<false, ...>` the FP64 path, "
                 "`tonal_plain_kernel` the plain FP64 path (the reference's kernel structure, one pass).", ""]
        out = os.path.join(ROOT, "profiles", f"tonal_kernel_stats_{TONAL_PROFILE}.md")
    lines = head + ["", "Command (repository root `R`, output directory `OUT`, on the GPU host):", "", "```",
                    "cd /tmp && export TMPDIR=/tmp",
                    "rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/prof -o bench -- \\"] + \
        cmd + ["```", ""] + note_ + ["| kernel | calls | avg ms | min ms | max ms | share |", "|---|---|---|---|---|---|"]
    for r in rows:
        n = r["Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").replace("acum::", "")
        lines.append(f"| `{n}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.3f} | "
                     f"{float(r['MinNs']) / 1e6:.3f} | {float(r['MaxNs']) / 1e6:.3f} | "
                     f"{float(r['Percentage']):.2f}% |")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))
    return 0


def note(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info():
    """CPU model (lscpu), CPUs visible to this process, and the threads the CPU baselines use: all
    of the process's CPUs, or OMP_NUM_THREADS when the host sets a smaller share."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None and os.path.exists("/proc/cpuinfo"):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    cpus = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(cpus, omp) if omp > 0 else cpus
    return dict(cpu_model=model, cpus_visible=cpus, nproc=os.cpu_count(), omp_num_threads=omp or None,
                threads=threads)


def timed_runs(fn, runs):
    """(first result, median seconds, all seconds) of `runs` calls of fn."""
    first, secs = None, []
    for i in range(runs):
        t = time.perf_counter()
        r = fn()
        secs.append(time.perf_counter() - t)
        if i == 0:
            first = r
    return first, float(np.median(secs)), secs


def tonal_accuracy(acum, O, P, g, o, dt):
    """Mixed-path histories g (observer columns of the oracle subset) against the FP64 oracle's o:
    rel-L2 per channel, OASPL and 1/3-octave band levels (bands above the -100 dB floor)."""
    ref_spl, _ = P.band_levels(o[2], dt)
    spl, _ = P.band_levels(g[2], dt)
    d_band, where = P.band_error(spl, ref_spl)
    return dict(observers=int(g[2].shape[1]), rel_l2_pA=O.rel_l2(g[2], o[2]), rel_l2_pT=O.rel_l2(g[0], o[0]),
                rel_l2_pL=O.rel_l2(g[1], o[1]),
                max_doaspl_db=float(np.abs(acum.oaspl(g[2]) - acum.oaspl(o[2].astype(np.float32))).max()),
                max_dband_db=d_band,
                worst_band_hz=None if where is None else float(P.ISO_BANDS[where[1]]),
                live_bands=int((ref_spl > -99.0).sum()))


def leg_bpm(acum, O, ctx, steps, host):
    """The broadband path on the 4 x 60-section workload: GPU kernels and the whole call, accuracy
    against the FP64 oracle on 8 observers, and the CPU restatement timed on all threads and on one
    (the reference's own loop is single-threaded, bb_core.cu:449)."""
    run = acum.Run.synth(acum.synth_params(3))
    c = run.bpm_case()
    ctx.bpm(c)  # warm
    gpu_ms, call_ms = [], []
    for _ in range(max(1, steps)):
        t = time.perf_counter()
        spl, oa = ctx.bpm(c)
        call_ms.append((time.perf_counter() - t) * 1e3)
        tm = ctx.timing()
        gpu_ms.append(tm.prep_ms + tm.eval_ms + tm.reduce_ms)
    nsrc = int(np.ctypeslib.as_array(c.nsect, shape=(c.nSurf_BB,)).sum()) * c.nTime
    evals = nsrc * c.nObs
    band_evals = evals * c.nf2
    g = float(np.median(gpu_ms))
    # algorithmic FLOP (FMA = 2, transcendental = 4): 15 per band-eval (six components, clip),
    # 90 per source-observer pair (retarded position, directivities, 1/r^2), 250 per source-band
    # (spectral shapes)
    flop = 15.0 * band_evals + 90.0 * evals + 250.0 * nsrc * c.nf2
    out = dict(workload=f"BPM broadband: {c.nSurf_BB} x {nsrc // c.nSurf_BB // c.nTime} sections, "
                        f"{c.nTime} steps, {c.nObs} observers, {c.nf2} bands",
               src_obs_time_evals=evals, band_evals=band_evals, gpu_ms=round(g, 3),
               call_ms=round(float(np.median(call_ms)), 2),
               value=round(evals / (g * 1e-3) / 1e9, 3), unit="G src-obs-time evals/s",
               band_evals_per_s=round(band_evals / (g * 1e-3) / 1e9, 2),
               roofline={"bound": "valu", "engines": "VALU (FP32, transcendental)",
                         "achieved": round(flop / (g * 1e-3) / 1e12, 3),
                         "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                         "frac": round(flop / (g * 1e-3) / PEAK_FP32, 4),
                         "flop": flop, "kernels": "bpm_shape + bpm_obs + bpm_final"})
    worst = worst_oa = 0.0
    for obs0 in (0, c.nObs - 4):
        ospl, ooa = O.bpm_f64(c, obs0, 4, host["threads"])
        worst = max(worst, float(np.abs(spl[obs0:obs0 + 4].astype(np.float64) - ospl).max()))
        worst_oa = max(worst_oa, float(np.abs(oa[obs0:obs0 + 4] - ooa).max()))
    out["accuracy"] = dict(observers=8, max_dspl_db=worst, max_doaspl_db=worst_oa)
    # CPU baselines: the FP64 restatement (orc_bpm), median of three runs each
    nall, n1 = 16, 2
    _, s_all, _ = timed_runs(lambda: O.bpm_f64(c, 0, nall, host["threads"]), 3)
    _, s_one, _ = timed_runs(lambda: O.bpm_f64(c, 0, n1, 1), 3)
    out["cpu_baseline"] = dict(value=round(nsrc * nall / s_all / 1e9, 6), unit="G src-obs-time evals/s",
                               cores=host["threads"], kind="port", runs=3,
                               sample=f"FP64 restatement of findBBnoise, first {nall} observers "
                                      f"({nsrc * nall:.3g} evals, median {s_all:.2f} s)")
    out["cpu_baseline_1thread"] = dict(value=round(nsrc * n1 / s_one / 1e9, 6), unit="G src-obs-time evals/s",
                                       cores=1, kind="port", runs=3,
                                       sample=f"same, first {n1} observers ({nsrc * n1:.3g} evals, "
                                              f"median {s_one:.2f} s)")
    out["vs_cpu"] = round(out["value"] / out["cpu_baseline"]["value"], 1)
    return out


def leg_batched(acum, O, P, ctx, host, ncases, per_launch, oracle_obs, steps=2, fp64=True, plain=True):
    """The batched workload's first `ncases` seeded rotor cases (of 64: 2-6 blades, 2,000-40,000
    elements per blade, 256-2048 steps, 16-2048 observers on spheres of 5R-50R, oM 150-400,
    Minf 0-0.2), `per_launch` cases per persistent launch: tile shapes per case from the
    autotuner, the timed passes replayed as a captured HIP graph, the FP64 path on the same
    launches, every case checked bit for bit against it run alone and against the FP64 oracle on
    its first observers; the plain FP64 path timed on a sample of the first case."""
    seed = 31004
    plain_rate, plain_sample = None, None
    tot_evals = 0
    mixed_ms = fp64_ms = eval_ms = 0.0
    host_s = tune_s = 0.0
    tuned = hits = widened = launches = 0
    acc = []
    windows = {}
    cpu_evals, cpu_s, cpu_obs = 0, 0.0, 0
    # Cases go into persistent launches of at most `per_launch`, in order, each launch holding what
    # fits a device-memory budget (the 64 cases need ~446 GB of element data, so they run in
    # groups): per case, records x (72 B uploaded arrays + 64 B positions + 64 B mixed attributes,
    # + 128 B FP64 attributes) and its partial rows (at most a quarter of its records or 8 GiB).
    def dev_bytes(i):
        p = acum.batch_params(seed, i)
        rec = p.nBlades * (2 if p.coaxial else 1) * p.nChord * p.nSpan * 2 * p.nTime
        per = 200 + (128 if fp64 else 0)
        slab = max(8 << 30, rec * 32) + (max(8 << 30, rec * 48) if fp64 else 0)
        return rec * per + slab
    budget = 260e9
    groups, cur, used = [], [], 0.0
    for i in range(ncases):
        b = dev_bytes(i)
        if cur and (len(cur) == per_launch or used + b > budget):
            groups.append(cur)
            cur, used = [], 0.0
        cur.append(i)
        used += b
    if cur:
        groups.append(cur)
    for idx in groups:
        w0 = idx[0]
        t0 = time.time()
        runs = [acum.Run.synth(acum.batch_params(seed, i)).preprocess() for i in idx]
        cases = [r.tonal_case(acum.MIXED) for r in runs]
        host_s += time.time() - t0
        h = ctx.stage(cases)
        t0 = time.time()
        ctx.autotune(h, acum.MIXED)
        tune_s += time.time() - t0
        tt = ctx.timing()
        tuned += tt.tune_cases
        hits += tt.tune_hits
        for i in range(len(cases)):
            w = ctx.case_tuning(h, i, acum.MIXED).window
            windows[w] = windows.get(w, 0) + 1
        tm = ctx.run(h, acum.MIXED)  # plan and warm, with per-launch kernel times
        eval_ms += tm.eval_ms
        launches = max(launches, tm.eval_launches)
        ctx.run(h, acum.MIXED, repeats=1, graph=True)  # capture
        tm = ctx.run(h, acum.MIXED, repeats=steps, graph=True)
        mixed_ms += tm.total_ms
        widened += tm.motion_widened_cases
        ev = sum(sum(c.E[s] for s in range(c.nSurf)) * c.nTime * c.nObs for c in cases)
        tot_evals += ev
        batch_out = [ctx.fetch(h, i, c.nTime, c.nObs) for i, c in enumerate(cases)]
        if fp64:
            ctx.run(h, acum.FP64)
            fp64_ms += ctx.run(h, acum.FP64).total_ms
        ctx.release(h)
        for i, (c, b) in enumerate(zip(cases, batch_out)):
            alone = ctx.tonal(c)
            if not all(np.array_equal(x, y) for x, y in zip(alone, b)):
                raise SystemExit(f"batched case {idx[i]}: histories differ from the case run alone")
            sub = P.obs_subset(c, np.arange(min(oracle_obs, c.nObs)), acum.FP64)
            o = O.tonal_f64(sub, host["threads"])
            acc.append(tonal_accuracy(acum, O, P, [x[:, :sub.nObs] for x in b], o, c.dPsi / c.oM))
            if w0 == 0 and i == 0:
                # CPU baseline: the oracle is parallel over observers, so it is timed on as many
                # observers of the first case as it has threads
                cb = P.obs_subset(c, np.arange(min(host["threads"], c.nObs)), acum.FP64)
                _, cpu_s, _ = timed_runs(lambda: O.tonal_f64(cb, host["threads"]), 3)
                cpu_evals = sum(c.E[s] for s in range(c.nSurf)) * c.nTime * cb.nObs
                cpu_obs = cb.nObs
                if plain:
                    # the plain FP64 path on about 2e10 evals of the first case (its first
                    # observers; a pass of the whole case would take tens of seconds)
                    ne = sum(c.E[s] for s in range(c.nSurf)) * c.nTime
                    nop = int(min(c.nObs, max(64, round(2e10 / ne))))
                    ctx.tonal_plain(P.obs_subset(c, np.arange(nop), acum.FP64))
                    tp = ctx.timing()
                    plain_rate = tp.evals / (tp.eval_ms * 1e-3) / 1e9
                    plain_sample = (f"first {nop} observers of case 0 ({tp.evals:.3g} evals, one pass, "
                                    f"{tp.eval_ms / 1e3:.1f} s)")
        note(f"batched cases {idx[0]}-{idx[-1]}: {ev:.3g} evals, {tm.total_ms:.1f} ms per mixed pass")
        del runs, cases, batch_out
    value = tot_evals / (mixed_ms * 1e-3) / 1e9
    peak = mixed_peak()
    out = dict(workload=f"first {ncases} of the 64 seeded uneven rotor cases, at most {per_launch} per persistent "
                        f"launch ({len(groups)} launches, grouped within a device-memory budget), tile shapes "
                        f"per case from the autotuner, HIP graph replay",
               evals=tot_evals, value=round(value, 3), unit="G src-obs-time evals/s",
               gpu_ms_per_pass=round(mixed_ms, 2),
               roofline={"bound": "valu", "kernel": "tonal_eval_kernel (all launches of a pass)",
                         "achieved": round(FLOP_PER_EVAL * tot_evals / (eval_ms * 1e-3) / 1e12, 3),
                         "peak": round(peak / 1e12, 3), "unit": "TFLOP/s",
                         "frac": round(FLOP_PER_EVAL * tot_evals / (eval_ms * 1e-3) / peak, 4)},
               cases_checked_against_alone=ncases,
               accuracy=dict(observers_per_case=oracle_obs,
                             max_rel_l2_pA=max(a["rel_l2_pA"] for a in acc),
                             max_doaspl_db=max(a["max_doaspl_db"] for a in acc),
                             max_dband_db=max(a["max_dband_db"] for a in acc)),
               autotune=dict(buckets_tuned=tuned, cache_hits=hits, seconds=round(tune_s, 1),
                             windows=windows),
               eval_launches_per_pass=launches, cases_on_wider_motion_bound=widened,
               host_synth_prep_s=round(host_s, 1),
               cpu_baseline=dict(value=round(cpu_evals / cpu_s / 1e9, 4), unit="G src-obs-time evals/s",
                                 cores=host["threads"], kind="port", runs=3,
                                 sample=f"FP64 oracle, first {cpu_obs} observers of case 0 "
                                        f"({cpu_evals:.3g} evals, median of 3 runs {cpu_s:.1f} s)"))
    if fp64:
        out["fp64_value"] = round(tot_evals / (fp64_ms * 1e-3) / 1e9, 3)
        out["mixed_over_fp64"] = round(value / out["fp64_value"], 3)
    if plain:
        out["plain_fp64_value"] = round(plain_rate, 3)
        out["plain_fp64_sample"] = plain_sample
        out["mixed_over_plain_fp64"] = round(value / plain_rate, 2)
    return out


def leg_stress(acum, O, P, ctx, host, steps, oracle_obs):
    """The stress workload: coaxial counter-rotating 2 x 4 blades, 50,000 elements per blade,
    2048 steps, 4096 observers on a 30R sphere, Minf 0.15, impulsive loading (3.36e12 evals),
    mixed path, accuracy against the FP64 oracle on spread observers."""
    t0 = time.time()
    run = acum.Run.synth(acum.synth_params(5))
    synth_s = time.time() - t0
    t0 = time.time()
    run.preprocess_gpu(ctx)
    prep_s = time.time() - t0
    case = run.tonal_case(acum.MIXED)
    E = sum(case.E[s] for s in range(case.nSurf))
    evals = E * case.nTime * case.nObs
    t0 = time.time()
    h = ctx.stage([case])
    stage_s = time.time() - t0
    ctx.run(h, acum.MIXED)
    ms, ems = [], []
    for _ in range(max(1, steps)):
        tm = ctx.run(h, acum.MIXED)
        ms.append(tm.total_ms)
        ems.append(tm.eval_ms)
    widened = tm.motion_widened_cases
    g = ctx.fetch(h, 0, case.nTime, case.nObs)
    ctx.release(h)
    idx = P.spread_observers(case.nObs, oracle_obs)
    sub = P.obs_subset(case, idx, acum.FP64)
    t = time.perf_counter()
    o = O.tonal_f64(sub, host["threads"])
    cpu_s = time.perf_counter() - t
    step_ms, eval_ms = float(np.median(ms)), float(np.median(ems))
    peak = mixed_peak()
    return dict(workload="coaxial 2 x 4 blades x 50,000 elements, 2048 steps, 4096 observers, Minf 0.15, "
                         "impulsive loading", evals=evals,
                value=round(evals / (step_ms * 1e-3) / 1e9, 3), unit="G src-obs-time evals/s",
                ms_per_pass=round(step_ms, 1),
                roofline={"bound": "valu", "kernel": "tonal_eval_kernel<mixed, moving>",
                          "achieved": round(FLOP_PER_EVAL * evals / (eval_ms * 1e-3) / 1e12, 3),
                          "peak": round(peak / 1e12, 3), "unit": "TFLOP/s",
                          "frac": round(FLOP_PER_EVAL * evals / (eval_ms * 1e-3) / peak, 4)},
                on_wider_motion_bound=bool(widened),
                accuracy=tonal_accuracy(acum, O, P, [x[:, idx] for x in g], o, case.dPsi / case.oM),
                cpu_baseline=dict(value=round(E * case.nTime * len(idx) / cpu_s / 1e9, 4),
                                  unit="G src-obs-time evals/s", cores=host["threads"], kind="port", runs=1,
                                  sample=f"FP64 oracle, {len(idx)} spread observers ({cpu_s:.1f} s)"),
                synth_s=round(synth_s, 1), gpu_preprocess_call_s=round(prep_s, 1), stage_s=round(stage_s, 1))


def parity_failures(out):
    """The path's bars that a measured line breaks (the bench then exits with status 3)."""
    bad = []
    acc = out.get("accuracy") or {}
    if acc.get("mixed_vs_oracle_max_doaspl_db", 0) > 0.01:
        bad.append("rotor mixed OASPL")
    if acc.get("mixed_vs_oracle_max_dband_db", 0) > 0.01:
        bad.append("rotor mixed 1/3-octave band")
    if acc.get("fp64_vs_oracle_rel_l2_pA", 0) > 1e-6:
        bad.append("rotor FP64 rel-L2")
    if acc.get("plain_fp64_vs_oracle_rel_l2_pA", 0) > 1e-6:
        bad.append("rotor plain FP64 rel-L2")
    for leg in ("batched", "stress"):
        a = (out.get(leg) or {}).get("accuracy") or {}
        if a.get("max_doaspl_db", 0) > 0.01 or a.get("max_dband_db", 0) > 0.01:
            bad.append(f"{leg} levels")
    b = (out.get("bpm") or {}).get("accuracy") or {}
    if b.get("max_dspl_db", 0) > 0.01 or b.get("max_doaspl_db", 0) > 0.01:
        bad.append("broadband levels")
    return bad


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--oracle-obs", type=int, default=32,
                    help="observers (spread over the layout) of the rotor's CPU baseline and accuracy")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-bpm", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-stress", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="leave out the plain FP64 path")
    ap.add_argument("--graph", action="store_true", help="replay the pass as a captured HIP graph")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain device copy first (for scaling HBM counters)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--summarize-stats", metavar="DIR")
    ap.add_argument("--summarize-counters", nargs="+", metavar="DIR")
    ap.add_argument("--workload", choices=("rotor", "batched", "stress", "bpm"), default="rotor",
                    help="rotor: the four-blade rotor line with the broadband, batched and stress legs "
                         "(default); batched: the uneven cases alone; stress: the coaxial case alone; "
                         "bpm: the broadband leg alone")
    ap.add_argument("--cases", type=int, default=8, help="batched cases (of 64)")
    ap.add_argument("--wave", type=int, default=4, help="batched cases per persistent launch")
    args = ap.parse_args()
    if args.summarize_stats:
        return summarize_stats(args.summarize_stats, args.workload)
    if args.summarize_pmc:
        return summarize_pmc(*args.summarize_pmc)
    if args.summarize_counters:
        return summarize_counters(args.summarize_counters, "bpm" if args.workload == "bpm" else "rotor")
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(observers would be sharded across GPUs with no communication); run with --gpus 1", file=sys.stderr)
        return 2

    import acum3_periodic_amd as acum
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O  # tests/oracle.py: the oracle binding (CPU baseline and accuracy only)
    import parity as P  # tests/parity.py: observer subsets and band-level comparison

    host = host_info()
    if args.copy_calibration:
        copy_calibration()
    if args.workload in ("batched", "stress", "bpm"):
        ctx = acum.Context(0)
        if args.workload == "batched":
            out = leg_batched(acum, O, P, ctx, host, args.cases, args.wave, 4, max(1, args.steps), not args.no_fp64,
                              not args.no_plain)
        elif args.workload == "stress":
            out = leg_stress(acum, O, P, ctx, host, args.steps, args.oracle_obs)
        else:
            out = leg_bpm(acum, O, ctx, args.steps, host)
        out["host"] = host
        ctx.close()
        print(json.dumps(out))
        bad = parity_failures({args.workload: out})
        if bad:
            print("parity bars broken: " + ", ".join(bad), file=sys.stderr)
            return 3
        return 0

    t0 = time.time()
    note("synthesising and preprocessing the four-blade rotor")
    run = acum.Run.synth(acum.synth_params(2))
    synth_s = time.time() - t0
    t0 = time.time()
    run.preprocess()
    host_prep_s = time.time() - t0
    case = run.tonal_case(acum.MIXED)
    E = [case.E[s] for s in range(case.nSurf)]
    evals = sum(E) * case.nTime * case.nObs

    ctx = acum.Context(0)
    prep = None
    if not args.no_bpm:  # the full measurement also times the GPU preprocessing (next row)
        note("GPU preprocessing")
        acum.Run.synth(acum.synth_params(2)).preprocess_gpu(ctx)  # warm-up call, not timed
        run2 = acum.Run.synth(acum.synth_params(2))
        t0 = time.time()
        run2.preprocess_gpu(ctx)
        gpu_call_s = time.time() - t0
        tp = ctx.timing()
        a = acum.case_arrays(case)
        b = acum.case_arrays(run2.tonal_case(acum.MIXED))
        same = all(np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)) for k in ("X", "N", "V", "PR"))
        nrec = sum(E) * case.nTime
        prep = dict(records=nrec, host_fp64_s=round(host_prep_s, 3), host_threads=host["threads"],
                    gpu_kernel_ms=round(tp.eval_ms, 3), gpu_upload_ms=round(tp.prep_ms, 2),
                    gpu_download_ms=round(tp.reduce_ms, 2), gpu_call_s=round(gpu_call_s, 3),
                    kernel_GBps=round(nrec * 92 / (tp.eval_ms * 1e-3) / 1e9, 1),
                    bit_identical_to_host=bool(same))
        del run2
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

    note("timing the mixed path")
    step_s, eval_ms, red_ms, tm = timed(acum.MIXED, args.steps, args.warmup)
    mixed = ctx.fetch(h, 0, case.nTime, case.nObs)
    partials = tm.n_partials
    items = tm.n_items
    tune = ctx.case_tuning(h, 0, acum.MIXED)

    # post-processing of the mixed history on the GPU: 1/3-octave band levels from the resident pA
    note("tonal band spectra")
    iso = np.array(ISO_BANDS, np.float32)
    ctx.fetch_spectra(h, 0, case.nObs, iso)  # warm-up call (first launches load the kernels)
    sp_spl, _ = ctx.fetch_spectra(h, 0, case.nObs, iso)
    sp_ms = ctx.timing().eval_ms

    fp64 = None
    if not args.no_fp64:
        note("timing the FP64 path")
        s64, e64, r64, t64 = timed(acum.FP64, max(1, min(args.steps, 3)), 1)
        plain = ctx.fetch(h, 0, case.nTime, case.nObs)
        t64s = ctx.case_tuning(h, 0, acum.FP64)
        fp64 = dict(value=round(evals / s64 / 1e9, 3), ms_per_step=round(s64 * 1e3, 3),
                    eval_ms=round(e64, 3), reduce_ms=round(r64, 4), partials=t64.n_partials,
                    roofline_frac=round(FLOP_PER_EVAL * evals / (e64 * 1e-3) / PEAK_FP64, 4),
                    tile={"window": t64s.window, "max_spread": t64s.max_spread, "max_tile": t64s.max_tile,
                          "motion_permille": t64s.motion_permille})
    ctx.release(h)
    plain_fp64 = plain_out = None
    if not args.no_plain:
        note("plain FP64 path (the reference's kernel structure in FP64), one pass")
        t0 = time.time()
        plain_out = ctx.tonal_plain(run.tonal_case(acum.FP64))
        call_s = time.time() - t0
        tpl = ctx.timing()
        plain_fp64 = dict(kernel="tonal_plain_kernel", value=round(evals / (tpl.eval_ms * 1e-3) / 1e9, 3),
                          eval_ms=round(tpl.eval_ms, 1), call_s=round(call_s, 2),
                          roofline_frac=round(FLOP_PER_EVAL * evals / (tpl.eval_ms * 1e-3) / PEAK_FP64, 5),
                          structure="one thread per (step, observer), 8x4x8 blocks striding over elements "
                                    "(the reference's launch), three retarded-time solves per eval, FP64 "
                                    "atomic adds into pT/pL/pA",
                          timed="one pass, kernel time by HIP events (element data resident in HBM)")

    # CPU baseline + accuracy: the FP64 oracle on observers spread over the hemisphere (observers
    # are independent, so the rate is exact for the sample), median of three runs; the float
    # instance of the same restatement (the reference's float arithmetic) likewise
    idx = P.spread_observers(case.nObs, args.oracle_obs)
    sub = P.obs_subset(case, idx, acum.FP64)
    nob = len(idx)
    note(f"CPU oracle on {nob} observers, three runs")
    (oT, oL, oA), cpu_s, cpu_runs = timed_runs(lambda: O.tonal_f64(sub, host["threads"]), 3)
    cpu_evals = sum(E) * case.nTime * nob
    note(f"CPU float oracle on {nob} observers, three runs")
    sub32 = P.obs_subset(case, idx, acum.MIXED)
    (fT, fL, fA), cpu32_s, _ = timed_runs(lambda: O.tonal_f32(sub32, host["threads"]), 3)

    dt = case.dPsi / case.oM
    gm = [m[:, idx] for m in mixed]
    a = tonal_accuracy(acum, O, P, gm, (oT, oL, oA), dt)
    ref_spl, _ = P.band_levels(oA, dt)
    acc = dict(observers=nob, mixed_vs_oracle_rel_l2_pA=a["rel_l2_pA"], mixed_vs_oracle_rel_l2_pT=a["rel_l2_pT"],
               mixed_vs_oracle_rel_l2_pL=a["rel_l2_pL"], mixed_vs_oracle_max_doaspl_db=a["max_doaspl_db"],
               mixed_vs_oracle_max_dband_db=a["max_dband_db"], mixed_worst_band_hz=a["worst_band_hz"],
               mixed_gpu_spectra_vs_oracle_max_dband_db=P.band_error(sp_spl[idx], ref_spl)[0],
               live_bands=a["live_bands"],
               float_oracle_vs_oracle_rel_l2_pA=O.rel_l2(fA, oA),
               float_oracle_vs_oracle_max_doaspl_db=float(
                   np.abs(acum.oaspl(fA) - acum.oaspl(oA.astype(np.float32))).max()))
    if fp64:
        acc["fp64_vs_oracle_rel_l2_pA"] = O.rel_l2(plain[2][:, idx], oA)
        acc["fp64_vs_oracle_rel_l2_pT"] = O.rel_l2(plain[0][:, idx], oT)
        d = np.abs(acum.oaspl(mixed[2]) - acum.oaspl(plain[2]))
        acc["mixed_vs_fp64_all_obs_max_doaspl_db"] = float(d.max())
    if plain_out is not None:
        acc["plain_fp64_vs_oracle_rel_l2_pA"] = O.rel_l2(plain_out[2][:, idx], oA)
        acc["plain_fp64_vs_oracle_rel_l2_pT"] = O.rel_l2(plain_out[0][:, idx], oT)
        d = np.abs(acum.oaspl(mixed[2]) - acum.oaspl(plain_out[2]))
        acc["mixed_vs_plain_fp64_all_obs_max_doaspl_db"] = float(d.max())

    o_spl, _ = O.tonal_spectra(mixed[2][:, idx], dt, iso)
    spectra = dict(kernels="spectra_stats + spectra_bins + spectra_bands", ms=round(sp_ms, 3), bands=int(iso.size),
                   max_dspl_db_vs_oracle=float(np.abs(sp_spl[idx] - o_spl).max()))

    bpm = batched = stress = None
    if not args.no_bpm:
        note("broadband path")
        bpm = leg_bpm(acum, O, ctx, args.steps, host)
    if not args.no_batched:
        note(f"batched leg: first {args.cases} cases, {args.wave} per launch")
        batched = leg_batched(acum, O, P, ctx, host, args.cases, args.wave, 4, 2, not args.no_fp64,
                              not args.no_plain)
    if not args.no_stress:
        note("stress leg")
        stress = leg_stress(acum, O, P, ctx, host, 2, 16)

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
                   "parallelism": "1 GPU",
                   "tile": {"window": tune.window, "max_spread": tune.max_spread,
                            "max_tile": tune.max_tile, "motion_permille": tune.motion_permille,
                            "on_wider_motion_bound": bool(tm.motion_widened_cases)}},
        # compute-bound on the vector ALUs: the FP64 observer-time solve and the FP32 integrand
        # (the integrand's four record products go through the matrix cores, 3 of ~73
        # instructions per step); the peak is the FP32/FP64 blend of the 70/30 FLOP split
        "roofline": {"bound": "valu", "engines": "VALU FP64/FP32 (+ MFMA FP32 for the record products)",
                     "achieved": round(achieved / 1e12, 3),
                     "peak": round(peak / 1e12, 3), "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": read_traffic(),
                     "kernel": "tonal_eval_kernel<mixed, moving>", "eval_ms": round(eval_ms, 3),
                     "flop_per_eval": FLOP_PER_EVAL},
        "cpu_baseline": {"value": round(cpu_evals / cpu_s / 1e9, 4), "unit": "G src-obs-time evals/s",
                         "cores": host["threads"], "kind": "port", "runs": 3,
                         "sample": f"FP64 oracle, same workload, {nob} observers spread over the hemisphere "
                                   f"({cpu_evals:.3g} evals, median of 3 runs {cpu_s:.1f} s)",
                         "run_seconds": [round(x, 2) for x in cpu_runs],
                         "float_instance_value": round(cpu_evals / cpu32_s / 1e9, 4)},
        "host": host,
        "binning": {"kernel": "tonal_reduce_kernel<float>", "ms": round(red_ms, 4),
                    "bytes": bin_bytes, "GBps": round(bin_bytes / (red_ms * 1e-3) / 1e9, 1),
                    "frac_hbm": round(bin_bytes / (red_ms * 1e-3) / HBM_PEAK, 4),
                    "partials": partials},
        "fp64_path": fp64,
        "mixed_over_fp64": round(value / fp64["value"], 3) if fp64 else None,
        # the baseline the path's 10x goal names: the reference's kernel structure in FP64
        "plain_fp64_path": plain_fp64,
        "mixed_over_plain_fp64": round(value / plain_fp64["value"], 2) if plain_fp64 else None,
        "accuracy": acc,
        "spectra": spectra,
        "bpm": bpm,
        "batched": batched,
        "stress": stress,
        "preprocess": prep,
        "work_items": items,
        "synth_s": round(synth_s, 2),
        "host_prep_s": round(host_prep_s, 2),
        "stage_s": round(stage_s, 2),
    }
    ctx.close()
    print(json.dumps(out))
    bad = parity_failures(out)
    if bad:
        print("parity bars broken: " + ", ".join(bad), file=sys.stderr)
        return 3
    return 0


if __name__ == "__main__":
    sys.exit(main())
