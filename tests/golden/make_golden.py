"""Writes the golden fixtures of this directory from the reference's own host code, compiled from
its sources into oracle/_ref/libref_host.so (make -C oracle ref).  Run from the repository root:

    python tests/golden/make_golden.py

Fixtures (inputs and the reference build's outputs; no reference source is kept):
  mathfunc.npz    differentiate, differentiate_biased, make_periodic (is = 0, 1), medfilt1,
                  pre_gaussfilt1, gaussfilt1 (mathfunc.cu:132-421) on seeded float series
  preprocess.npz  preprocess_imper (preprocess.cu:228-413) of a seeded synthetic run, lowpass 0/1
  bpm.npz         findBBnoise (bb_core.cu:11-586) SPL/OASPL of seeded synthetic broadband runs
The synthetic runs are regenerated from their seeded parameters by the library's generator, which
the tests re-run; each fixture stores those parameters.
"""
import ctypes as C
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
import acum3_periodic_amd as acum  # noqa: E402
import oracle as O  # noqa: E402

f32p = C.POINTER(C.c_float)

# seeded parameters shared with tests/test_cpu_golden.py
PRE_CASES = [dict(config=1, nChord=3, nSpan=2, nTime=24, lowpass=0),
             dict(config=1, nChord=2, nSpan=2, nTime=20, lowpass=1)]
BPM_CASES = [dict(nSect=6, nTime=10, nObsAz=3, nObsEl=2, trip=1, t_del=2),
             dict(nSect=6, nTime=10, nObsAz=3, nObsEl=2, trip=0, t_del=1),
             dict(nSect=5, nTime=8, nObsAz=2, nObsEl=2, trip=0, t_del=0, nBands=42)]


def fn(name, argtypes):
    f = getattr(O.ref_lib(), name)
    f.restype, f.argtypes = None, argtypes
    return f


def mathfunc():
    out = {}
    p = lambda a: a.ctypes.data_as(f32p)
    for n in (9, 40):
        rng = np.random.default_rng(100 + n)
        u = (rng.standard_normal(n) * 7).astype(np.float32)
        u[2] = u[3]
        dt = np.float32(0.0217)
        out[f"u{n}"] = u
        for key, name in (("diff", "_Z13differentiatePfS_fi"), ("diffb", "_Z20differentiate_biasedPfS_fi")):
            o = np.empty(n, np.float32)
            fn(name, [f32p, f32p, C.c_float, C.c_int])(p(u.copy()), p(o), dt, n)
            out[f"{key}{n}"] = o
        for is_ in (0, 1):
            o = u.copy()
            cub = np.empty(n + 1, np.float32)
            fn("_Z13make_periodicPfiiS_", [f32p, C.c_int, C.c_int, f32p])(p(o), n, is_, p(cub))
            out[f"per{is_}_{n}"] = o
        o = u.copy()
        fn("_Z8medfilt1Pfi", [f32p, C.c_int])(p(o), n)
        out[f"med{n}"] = o
        g = np.empty(13, np.float32)
        fn("_Z14pre_gaussfilt1Pfi", [f32p, C.c_int])(p(g), 13)
        out["gauss"] = g
        if n >= 13:
            o = u.copy()
            fn("_Z10gaussfilt1PfS_ii", [f32p, f32p, C.c_int, C.c_int])(p(o), p(g), n, 13)
            out[f"gf{n}"] = o
    out["dt"] = np.float32(0.0217)
    np.savez_compressed(os.path.join(HERE, "mathfunc.npz"), **out)


def preprocess():
    out = {}
    for i, pc in enumerate(PRE_CASES):
        pc = dict(pc)
        cfg, lowpass = pc.pop("config"), pc.pop("lowpass")
        with tempfile.TemporaryDirectory() as d:
            acum.Run.synth(acum.synth_params(cfg, **pc)).save(d)
            if lowpass:
                txt = open(os.path.join(d, "input")).read().replace("lowpass 0", "lowpass 1")
                open(os.path.join(d, "input"), "w").write(txt)
            o = os.path.join(d, "out")
            subprocess.run([sys.executable, os.path.join(os.path.dirname(HERE), "ref_host.py"), d, o],
                           check=True, capture_output=True)
            for nm in ("X", "N", "V", "PR"):
                out[f"c{i}_{nm}"] = np.load(os.path.join(o, f"pre_{nm}.npy"))
    np.savez_compressed(os.path.join(HERE, "preprocess.npz"), **out)


def bpm():
    out = {}
    for i, bc in enumerate(BPM_CASES):
        c = acum.Run.synth(acum.synth_params(3, **bc)).bpm_case()
        spl, oa = O.ref_bpm(c)
        out[f"c{i}_spl"] = spl
        out[f"c{i}_oaspl"] = oa
    np.savez_compressed(os.path.join(HERE, "bpm.npz"), **out)


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref/libref_host.so is missing: run make -C oracle ref first")
    mathfunc()
    preprocess()
    bpm()
    print("fixtures written to", HERE)
