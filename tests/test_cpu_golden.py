"""The oracle against the committed golden fixtures (tests/golden/, written by make_golden.py from
the reference's own host code compiled from its sources).  These pins hold without the reference
build: mathfunc float bit-exact, preprocess_imper float bit-exact, findBBnoise within 1e-3 dB."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import acum3_periodic_amd as acum
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
from make_golden import PRE_CASES, BPM_CASES  # noqa: E402

f32p = C.POINTER(C.c_float)


def test_mathfunc_fixtures():
    g = np.load(os.path.join(GOLD, "mathfunc.npz"))
    dt = np.float32(g["dt"])
    p = lambda a: a.ctypes.data_as(f32p)
    same = lambda a, b: np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for n in (9, 40):
        u = g[f"u{n}"]
        for key, name in (("diff", "orc_differentiate_f32"), ("diffb", "orc_differentiate_biased_f32")):
            o = np.empty(n, np.float32)
            getattr(O.orc, name)(p(u), p(o), dt, n)
            assert same(o, g[f"{key}{n}"]), (key, n)
        for is_ in (0, 1):
            o = u.copy()
            O.orc.orc_make_periodic_f32(p(o), n, is_)
            assert same(o, g[f"per{is_}_{n}"]), (is_, n)
        o = u.copy()
        O.orc.orc_medfilt1_f32(p(o), n)
        assert same(o, g[f"med{n}"])
        w = np.empty(13, np.float32)
        O.orc.orc_pre_gaussfilt1_f32(p(w), 13)
        assert same(w, g["gauss"])
        if n >= 13:
            o = u.copy()
            O.orc.orc_gaussfilt1_f32(p(o), p(w), n, 13)
            assert same(o, g[f"gf{n}"])


@pytest.mark.parametrize("i", range(len(PRE_CASES)))
def test_preprocess_fixture(i):
    g = np.load(os.path.join(GOLD, "preprocess.npz"))
    pc = dict(PRE_CASES[i])
    cfg, lowpass = pc.pop("config"), pc.pop("lowpass")
    run = acum.Run.synth(acum.synth_params(cfg, **pc))
    c = acum.TonalCase()
    acum.lib.acum_run_tonal_case(run._h, acum.FP64, C.byref(c))
    raw = {k: v.copy() for k, v in acum.case_arrays(c).items()}
    inp = run.inputs()
    E, nT = raw["E"], inp.nTime
    off = np.concatenate([[0], np.cumsum(E)[:-1]]).astype(np.int64) * nT
    i64 = lambda a: np.ascontiguousarray(a, np.int64).ctypes.data_as(C.POINTER(C.c_int64))
    ptr = {w: np.ascontiguousarray(off * w) for w in (2, 4, 6)}
    X, N, PR = raw["X"], raw["N"], raw["PR"]
    V = np.zeros(N.size, np.float32)
    f = np.float32
    O.orc.orc_preprocess_imper_f32(len(E), i64(E), nT, O.p32(X), O.p32(N), O.p32(V), O.p32(PR),
                                   i64(ptr[4]), i64(ptr[6]), i64(ptr[6]), i64(ptr[2]), f(inp.CFDscale),
                                   f(inp.rhoRef), f(inp.Mref), f(inp.a0), f(inp.dTau), inp.periodic, lowpass)
    for nm, arr in (("X", X), ("N", N), ("V", V), ("PR", PR)):
        assert np.array_equal(arr.view(np.uint32), g[f"c{i}_{nm}"].view(np.uint32)), nm


@pytest.mark.parametrize("i", range(len(BPM_CASES)))
def test_bpm_fixture(i):
    g = np.load(os.path.join(GOLD, "bpm.npz"))
    c = acum.Run.synth(acum.synth_params(3, **BPM_CASES[i])).bpm_case()
    spl, oa = O.bpm_f64(c)
    assert np.abs(spl - g[f"c{i}_spl"]).max() < 1e-3
    assert np.abs(oa - g[f"c{i}_oaspl"]).max() < 1e-3
