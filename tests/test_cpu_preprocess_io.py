"""Host side without a GPU: preprocessing numerics and the file contract.

Pins of the oracle's preprocessing restatement (oracle/orc_preprocess.cpp):
  * closed forms (periodic central difference of a sine, make_periodic invariants, Gaussian
    weights, the median tie rule);
  * bit-for-bit agreement of the float instance with the reference's own mathfunc.cu /
    preprocess.cu, compiled from its sources into oracle/_ref (skipped when that build is absent).
The product's FP64 preprocessing (acum_run_preprocess) must equal the oracle's FP64 instance bit
for bit, and the product's readers and writers must agree with the reference's io.cu.
"""
import ctypes as C
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import acum3_periodic_amd as acum
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference sources)")
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


# ---- closed forms ------------------------------------------------------------------------
def test_central_difference_of_sine():
    n, w, dt = 64, 3.0, 0.013
    t = np.arange(n) * dt
    # an integer number of periods over the series so the wrap is smooth
    w = 2 * np.pi * 4 / (n * dt)
    u = np.sin(w * t)
    ud = np.empty(n)
    O.orc.orc_differentiate_f64(O.p64(u), O.p64(ud), dt, n)
    expect = w * np.cos(w * t) * np.sin(w * dt) / (w * dt)
    np.testing.assert_allclose(ud, expect, rtol=0, atol=1e-12 * w)


def test_make_periodic_invariants():
    n = 50
    rng = np.random.default_rng(3)
    f = np.cumsum(rng.standard_normal(n)) + 5.0
    g = f.copy()
    O.orc.orc_make_periodic_f64(O.p64(g), n, 0)
    scale = np.abs(f).max()
    # the last value is unchanged
    assert abs(g[-1] - f[-1]) < 1e-12 * scale
    # the change is a cubic c(i) in the extended index i = 1..n (i = 0 is the point 2 f0 - f1
    # prepended before the series) with c(0) = f[n-1] - (2 f0 - f1), so the prepended point
    # moves onto the last value; c(n) = 0, c'(n) = 0, and c'(0) = s2 - s1 with
    # s1 = f1 - f0, s2 = f[n-1] - f[n-2]
    i = np.arange(1, n + 1)
    coef = np.polyfit(i, g - f, 3)
    assert np.abs(np.polyval(coef, i) - (g - f)).max() < 1e-9 * scale
    s1, s2 = f[1] - f[0], f[-1] - f[-2]
    d = np.polyder(coef)
    assert abs(np.polyval(coef, 0) - (f[-1] - (2 * f[0] - f[1]))) < 1e-8 * scale
    assert abs(np.polyval(coef, n)) < 1e-9 * scale
    assert abs(np.polyval(d, n)) < 1e-9 * scale
    assert abs(np.polyval(d, 0) - (s2 - s1)) < 1e-8 * scale
    # a smooth periodic input changes by O(h^2)
    for m in (64, 128):
        x = np.sin(2 * np.pi * np.arange(m) / m)
        y = x.copy()
        O.orc.orc_make_periodic_f64(O.p64(y), m, 0)
        assert np.abs(y - x).max() < 40.0 / m ** 2


def test_gaussian_weights():
    g = np.empty(13)
    O.orc.orc_pre_gaussfilt1_f64(O.p64(g), 13)
    sd = 2.0
    x = np.arange(13) - 6
    w = np.exp(-0.5 * x * x / sd / sd)
    np.testing.assert_allclose(g, w / w.sum(), rtol=1e-14)
    assert abs(g.sum() - 1) < 1e-15


def test_median_tie_rule():
    """3-point periodic median with strict tests: a tie returns the right neighbour
    (mathfunc.cu:244-255)."""
    f = np.array([1.0, 5.0, 2.0, 2.0, 7.0, 7.0, 3.0, 4.0])
    g = f.copy()
    O.orc.orc_medfilt1_f64(O.p64(g), len(f))
    # i=0: (4,1,5) -> 4;  i=1: (1,5,2) -> 2;  i=2: (5,2,2) tie -> c=2;  i=3: (2,2,7) tie -> 7
    # i=4: (2,7,7) tie -> 7;  i=5: (7,7,3) tie -> 3;  i=6: (7,3,4) -> 4;  i=7: (3,4,1) -> 3
    assert g.tolist() == [4.0, 2.0, 2.0, 7.0, 7.0, 3.0, 4.0, 3.0]


# ---- float instance vs the reference's own mathfunc.cu --------------------------------------
def _ref_fn(name, restype, argtypes):
    f = getattr(O.ref_lib(), name)
    f.restype, f.argtypes = restype, argtypes
    return f


@needs_ref
@pytest.mark.parametrize("n", [8, 37, 360])
def test_mathfunc_float_bit_exact(n):
    rng = np.random.default_rng(n)
    u = (rng.standard_normal(n) * 10).astype(np.float32)
    u[3] = u[4]  # a tie for the median
    dt = np.float32(0.0123)
    p = lambda a: a.ctypes.data_as(f32p)
    for orc_name, ref_name in (("orc_differentiate_f32", "_Z13differentiatePfS_fi"),
                               ("orc_differentiate_biased_f32", "_Z20differentiate_biasedPfS_fi")):
        a = np.empty(n, np.float32)
        b = np.empty(n, np.float32)
        getattr(O.orc, orc_name)(p(u), p(a), dt, n)
        _ref_fn(ref_name, None, [f32p, f32p, C.c_float, C.c_int])(p(u.copy()), p(b), dt, n)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), orc_name
    for is_ in (0, 1):
        a = u.copy()
        b = u.copy()
        cub = np.empty(n + 1, np.float32)
        O.orc.orc_make_periodic_f32(p(a), n, is_)
        _ref_fn("_Z13make_periodicPfiiS_", None, [f32p, C.c_int, C.c_int, f32p])(p(b), n, is_, p(cub))
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"make_periodic is={is_}"
    a = u.copy()
    b = u.copy()
    O.orc.orc_medfilt1_f32(p(a), n)
    _ref_fn("_Z8medfilt1Pfi", None, [f32p, C.c_int])(p(b), n)
    assert np.array_equal(a, b)
    ga = np.empty(13, np.float32)
    gb = np.empty(13, np.float32)
    O.orc.orc_pre_gaussfilt1_f32(p(ga), 13)
    _ref_fn("_Z14pre_gaussfilt1Pfi", None, [f32p, C.c_int])(p(gb), 13)
    assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
    if n >= 13:
        a = u.copy()
        b = u.copy()
        O.orc.orc_gaussfilt1_f32(p(a), p(ga), n, 13)
        _ref_fn("_Z10gaussfilt1PfS_ii", None, [f32p, f32p, C.c_int, C.c_int])(p(b), p(gb), n, 13)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---- whole run directories ----------------------------------------------------------------
def _small_run(tmp_path, bb=False, lowpass=0, periodic=1, name="case"):
    if bb:
        p = acum.synth_params(3, nSect=10, nTime=24, nObsAz=3, nObsEl=2, nBands=42)
    else:
        p = acum.synth_params(1, nChord=5, nSpan=4, nTime=36)
    d = tmp_path / name
    acum.Run.synth(p).save(str(d))
    if lowpass or not periodic:
        txt = (d / "input").read_text()
        txt = txt.replace("lowpass 0", f"lowpass {lowpass}").replace("periodic 1", f"periodic {periodic}")
        (d / "input").write_text(txt)
    return d


def _run_ref(case_dir, out_dir, bb=False, write=False):
    cmd = [sys.executable, os.path.join(HERE, "ref_host.py"), str(case_dir), str(out_dir)]
    if bb:
        cmd.append("--bb")
    if write:
        cmd.append("--write")
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    return lambda n: np.load(os.path.join(out_dir, n + ".npy"))


def _raw_arrays(run):
    """The run's AoS arrays before preprocessing, read through a case view."""
    c = acum.TonalCase()
    acum.lib.acum_run_tonal_case(run._h, acum.FP64, C.byref(c))  # fills pointers even if not preprocessed
    return {k: v.copy() for k, v in acum.case_arrays(c).items()}


@needs_ref
@pytest.mark.parametrize("lowpass,periodic", [(0, 1), (1, 1), (0, 0), (1, 0)])
def test_preprocess_matches_reference(tmp_path, lowpass, periodic):
    d = _small_run(tmp_path, lowpass=lowpass, periodic=periodic)
    ref = _run_ref(d, tmp_path / "ref")
    run = acum.Run.load(str(d))
    raw = _raw_arrays(run)
    # readers (io.cu:106-135, 206-333): identical floats
    assert np.array_equal(raw["E"], ref("E"))
    assert np.array_equal(raw["rObs"].reshape(-1), ref("rObs"))
    assert np.array_equal(raw["X"], ref("raw_X"))
    assert np.array_equal(raw["N"][np.arange(raw["N"].size) % 6 < 3], ref("raw_N")[np.arange(raw["N"].size) % 6 < 3])
    assert np.array_equal(raw["PR"][0::2], ref("raw_PR")[0::2])
    # derived constants (io.cu:70-83) in float as the reference keeps them
    names = list(ref("scalar_names"))
    sc = dict(zip(names, ref("scalars")))
    inp = run.inputs()
    assert np.float32(inp.dTau) == np.float32(sc["dTau"])
    assert np.float32(inp.dPsi) == np.float32(sc["dPsi"])
    assert np.allclose([inp.vInf[i] for i in range(3)], ref("vInf"), rtol=1e-6, atol=1e-6)

    # the float instance of the oracle reproduces preprocess_imper bit for bit
    E = raw["E"]
    nT = int(sc["nTime"])
    off = np.concatenate([[0], np.cumsum(E)[:-1]]).astype(np.int64) * nT
    X, N, PR = raw["X"].copy(), raw["N"].copy(), raw["PR"].copy()
    V = np.zeros(raw["N"].size, np.float32)
    i64 = lambda a: np.ascontiguousarray(a, np.int64).ctypes.data_as(C.POINTER(C.c_int64))
    ptr = {w: np.ascontiguousarray(off * w) for w in (2, 4, 6)}
    args = (len(E), i64(E), nT, O.p32(X), O.p32(N), O.p32(V), O.p32(PR), i64(ptr[4]), i64(ptr[6]),
            i64(ptr[6]), i64(ptr[2]))
    O.orc.orc_preprocess_imper_f32(*args, np.float32(sc["CFDscale"]), np.float32(sc["rhoRef"]),
                                   np.float32(sc["Mref"]), np.float32(sc["a0"]), np.float32(sc["dTau"]),
                                   periodic, lowpass)
    for nm, arr in (("X", X), ("N", N), ("V", V), ("PR", PR)):
        assert np.array_equal(arr.view(np.uint32), ref("pre_" + nm).view(np.uint32)), nm

    # the product's FP64 preprocessing equals the oracle's FP64 instance bit for bit
    X, N, PR = raw["X"].copy(), raw["N"].copy(), raw["PR"].copy()
    V = np.zeros(raw["N"].size, np.float32)
    args = (len(E), i64(E), nT, O.p32(X), O.p32(N), O.p32(V), O.p32(PR), i64(ptr[4]), i64(ptr[6]),
            i64(ptr[6]), i64(ptr[2]))
    O.orc.orc_preprocess_imper_f64(*args, inp.CFDscale, inp.rhoRef, inp.Mref, inp.a0, inp.dTau,
                                   periodic, lowpass)
    run.preprocess()
    got = acum.case_arrays(run.tonal_case(acum.FP64))
    for nm, arr in (("X", X), ("N", N), ("V", V), ("PR", PR)):
        assert np.array_equal(got[nm].view(np.uint32), arr.view(np.uint32)), nm
        # and stays close to the reference's float pipeline
        r = ref("pre_" + nm)
        assert O.rel_l2(got[nm], r) < 1e-4, nm


def test_product_preprocess_equals_oracle_fp64(tmp_path):
    """Same check as above without the reference build (synthetic run in memory)."""
    run = acum.Run.synth(acum.synth_params(1, nChord=4, nSpan=3, nTime=40))
    raw = _raw_arrays(run)
    inp = run.inputs()
    E, nT = raw["E"], inp.nTime
    off = np.concatenate([[0], np.cumsum(E)[:-1]]).astype(np.int64) * nT
    i64 = lambda a: np.ascontiguousarray(a, np.int64).ctypes.data_as(C.POINTER(C.c_int64))
    ptr = {w: np.ascontiguousarray(off * w) for w in (2, 4, 6)}
    X, N, PR = raw["X"].copy(), raw["N"].copy(), raw["PR"].copy()
    V = np.zeros(N.size, np.float32)
    O.orc.orc_preprocess_imper_f64(len(E), i64(E), nT, O.p32(X), O.p32(N), O.p32(V), O.p32(PR),
                                   i64(ptr[4]), i64(ptr[6]), i64(ptr[6]), i64(ptr[2]), inp.CFDscale,
                                   inp.rhoRef, inp.Mref, inp.a0, inp.dTau, inp.periodic, inp.lowpass)
    run.preprocess()
    got = acum.case_arrays(run.tonal_case(acum.FP64))
    for nm, arr in (("X", X), ("N", N), ("V", V), ("PR", PR)):
        assert np.array_equal(got[nm].view(np.uint32), arr.view(np.uint32)), nm


@needs_ref
def test_writers_byte_identical(tmp_path):
    d = _small_run(tmp_path, bb=True)
    ref_out = tmp_path / "ref"
    _run_ref(d, ref_out, bb=True, write=True)
    sys.path.insert(0, HERE)
    from ref_host import seeded_outputs
    run = acum.Run.load(str(d))
    inp = run.inputs()
    c = run.bpm_case()
    (pT, pL, pA), oaspl, spl = seeded_outputs(inp.nTime, c.nObs, c.nf2)
    mine = tmp_path / "mine"
    mine.mkdir()
    run.write_tonal(str(mine), pT, pL, pA)
    run.write_bb(str(mine), oaspl, spl)
    for f in ("pT.out", "pL.out", "pA.out", "bb_oaspl.out", "bb_spl.dat"):
        assert (mine / f).read_bytes() == (ref_out / f).read_bytes(), f


@needs_ref
def test_broadband_reader_matches_reference(tmp_path):
    d = _small_run(tmp_path, bb=True)
    ref = _run_ref(d, tmp_path / "ref", bb=True)
    c = acum.Run.load(str(d)).bpm_case()
    nS = c.nSurf_BB
    g = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy()
    nsect = g(c.nsect, nS)
    assert np.array_equal(nsect, ref("bb_nsect"))
    tot = int(nsect.sum())
    assert np.array_equal(g(c.xyzcR, 4 * tot), ref("bb_xyzcR"))
    assert np.array_equal(g(c.bbdata, 7 * tot * c.nTime), ref("bb_bbdata"))  # cn, cc divided by (r/R)^2
    for nm in ("span", "omega", "psi_offset", "ccw", "trip", "nsect_t", "nmach_t", "naoa_t", "nre_t"):
        assert np.array_equal(g(getattr(c, nm), nS), ref("bb_" + nm)), nm
    for nm in ("rotate", "translate"):
        assert np.array_equal(g(getattr(c, nm), 3 * nS), ref("bb_" + nm)), nm
    nst, nm_, na, nr = (ref("bb_nsect_t"), ref("bb_nmach_t"), ref("bb_naoa_t"), ref("bb_nre_t"))
    assert np.array_equal(g(c.sect_t, int(nst.sum())), ref("bb_sect_t"))
    assert np.array_equal(g(c.mach_t, int(nm_.sum())), ref("bb_mach_t"))
    assert np.array_equal(g(c.aoa_t, int(na.sum())), ref("bb_aoa_t"))
    assert np.array_equal(g(c.re_t, int(nr.sum())), ref("bb_re_t"))
    assert np.array_equal(g(c.bbdata_t, int((nst * nm_ * na * nr * 7).sum())), ref("bb_bbdata_t"))
    assert np.array_equal(g(c.ff, c.nf2), ref("bb_ff"))
    assert [nS, c.t_del, c.nf2] == list(ref("bb_ints"))


def test_surface_reader_matches_strtof(tmp_path):
    """The parallel surface reader (exact fast path + strtof fallback) gives strtof's float for
    every token spelling, including values on and next to float rounding midpoints."""
    run = acum.Run.synth(acum.synth_params(1, nChord=3, nSpan=2, nTime=32))
    d = tmp_path / "r"
    run.save(str(d))
    libc = C.CDLL(None)
    libc.strtof.restype = C.c_float
    libc.strtof.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]
    rng = np.random.default_rng(17)
    E = run.tonal_case_raw_sizes()[0][0]
    nT = 32
    n = nT * E * 8
    vals = rng.standard_normal(n) * 10.0 ** rng.integers(-12, 12, n)
    # exact float midpoints and their neighbours, printed with 17 significant digits
    f = rng.standard_normal(n // 4).astype(np.float32)
    nxt = np.nextafter(f, np.float32(np.inf))
    mid = (f.astype(np.float64) + nxt.astype(np.float64)) / 2
    vals[: n // 4] = mid
    vals[n // 4: n // 2] = np.nextafter(mid, np.inf)
    fmts = ["%.9e", "%.17g", "%.6f", "%g", "%+.8E", "%.3e", "%.25f", "%d"]
    toks = []
    for i, v in enumerate(vals):
        fm = fmts[i % len(fmts)]
        toks.append(fm % (int(v) if fm == "%d" else v))
    toks[5] = "1e-40"      # subnormal
    toks[13] = "-0.0"
    toks[21] = "0.000000000000000000000000001234567"
    toks[29] = "3.14159265358979323846264338327950288419716939937510"
    lines = [f"{E} {nT}"] + [" ".join(toks[i:i + 8]) for i in range(0, n, 8)]
    (d / "surfaces" / "surface_1.dat").write_text("\n".join(lines) + "\n")
    got = acum.Run.load(str(d)).surface_raw(0).reshape(-1)
    want = np.array([libc.strtof(t.encode(), None) for t in toks], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_tonal_writer_matches_printf(tmp_path):
    """pT/pL/pA.out rows are " %14.8E" of the float values as io.cu:143-192 prints them,
    checked against C's snprintf over random floats across the whole range and special values."""
    run = acum.Run.synth(acum.synth_params(1, nChord=2, nSpan=2, nTime=24, nObs=37))
    inp = run.inputs()
    c = run.preprocess().tonal_case(acum.FP64)
    nT, nObs = c.nTime, c.nObs
    rng = np.random.default_rng(3)
    bits = rng.integers(0, 2 ** 32, nT * nObs, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32).copy()
    vals[:8] = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -3.4028235e38, 1.0]
    pT = vals.copy()
    pL = (vals * 0.5).astype(np.float32)
    pA = np.arange(nT * nObs, dtype=np.float32) * np.float32(1.37e-3)
    run.write_tonal(str(tmp_path), pT, pL, pA)
    libc = C.CDLL(None)
    buf = C.create_string_buffer(64)

    def pf(v):
        libc.snprintf(buf, 64, b" %14.8E", C.c_double(float(v)))
        return buf.value.decode()

    dpsi = np.float32(inp.dPsi)
    for name, arr in (("pT.out", pT), ("pL.out", pL), ("pA.out", pA)):
        lines = [f"# ntime nobs", f"{nT} {nObs}"]
        for i in range(nT):
            row = pf(np.float32(i) * dpsi) + "".join(pf(arr[i * nObs + j]) for j in range(nObs))
            lines.append(row)
        assert (tmp_path / name).read_text() == "\n".join(lines) + "\n", name


def test_surface_cache_round_trip_and_invalidation(tmp_path):
    """A run loaded through the binary surface caches equals the run parsed from the text; a
    cache whose text file has changed, or a damaged cache, is ignored."""
    import os
    import time
    d = tmp_path / "r"
    p = acum.synth_params(1, nChord=3, nSpan=2, nTime=12)
    acum.Run.synth(p).save(str(d))
    text = acum.Run.load(str(d))
    text.save_cache(str(d))
    cache = d / "surfaces" / "surface_1.acumc"
    assert cache.exists()
    # parse from the cache: same arrays; make the text unreadable as numbers to prove it is not parsed
    dat = d / "surfaces" / "surface_1.dat"
    st = os.stat(dat)
    raw = dat.read_bytes()
    body = raw.split(b"\n", 1)
    dat.write_bytes(body[0] + b"\n" + bytes(b if b in b" \n" else ord("x") for b in body[1]))
    os.utime(dat, ns=(st.st_atime_ns, st.st_mtime_ns))
    via = acum.Run.load(str(d))
    for i in range(text.inputs().nSurf):
        assert np.array_equal(text.surface_raw(i), via.surface_raw(i))
    # the text changes (new time): the cache no longer matches and the (garbled) text is parsed
    time.sleep(0.01)
    os.utime(dat, None)
    with pytest.raises(acum.AcumError):
        acum.Run.load(str(d))
    # restore the text: no cache matches it until one is written again
    dat.write_bytes(raw)
    again = acum.Run.load(str(d))
    assert np.array_equal(text.surface_raw(0), again.surface_raw(0))
    again.save_cache(str(d))
    # a damaged cache is ignored
    b = bytearray(cache.read_bytes())
    b[0] ^= 0xFF
    cache.write_bytes(bytes(b))
    assert np.array_equal(text.surface_raw(0), acum.Run.load(str(d)).surface_raw(0))
    # a preprocessed run cannot be cached
    with pytest.raises(acum.AcumError):
        acum.Run.load(str(d)).preprocess().save_cache(str(d))
