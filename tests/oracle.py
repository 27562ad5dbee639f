"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the host library compiled from
the reference's own sources (oracle/_ref/libref_host.so).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

import acum3_periodic_amd as acum

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# ACUM_ORACLE_PATH selects another build of the oracle (its sanitizer build)
ORACLE_PATH = os.environ.get("ACUM_ORACLE_PATH") or os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libref_host.so")

_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


def _load_oracle():
    lib = C.CDLL(ORACLE_PATH)
    tc = C.POINTER(acum.TonalCase)
    bc = C.POINTER(acum.BpmCase)
    sig = {
        "orc_tonal_f64": (C.c_int, [tc, _f64p, _f64p, _f64p, C.c_int]),
        "orc_tonal_f32": (C.c_int, [tc, _f32p, _f32p, _f32p, C.c_int]),
        "orc_tonal_bins": (C.c_int, [tc, C.c_int64, _i32p, _i32p, _i32p, _i32p]),
        "orc_obs_tau": (C.c_double, [tc, _f32p, _f32p, C.c_int64]),
        "orc_retarded_r": (C.c_double, [tc, _f32p, _f32p]),
        "orc_differentiate_f32": (None, [_f32p, _f32p, C.c_float, C.c_int]),
        "orc_differentiate_biased_f32": (None, [_f32p, _f32p, C.c_float, C.c_int]),
        "orc_make_periodic_f32": (None, [_f32p, C.c_int, C.c_int]),
        "orc_medfilt1_f32": (None, [_f32p, C.c_int]),
        "orc_pre_gaussfilt1_f32": (None, [_f32p, C.c_int]),
        "orc_gaussfilt1_f32": (None, [_f32p, _f32p, C.c_int, C.c_int]),
        "orc_differentiate_f64": (None, [_f64p, _f64p, C.c_double, C.c_int]),
        "orc_differentiate_biased_f64": (None, [_f64p, _f64p, C.c_double, C.c_int]),
        "orc_make_periodic_f64": (None, [_f64p, C.c_int, C.c_int]),
        "orc_medfilt1_f64": (None, [_f64p, C.c_int]),
        "orc_pre_gaussfilt1_f64": (None, [_f64p, C.c_int]),
        "orc_gaussfilt1_f64": (None, [_f64p, _f64p, C.c_int, C.c_int]),
        "orc_preprocess_imper_f32": (None, [C.c_int, _i64p, C.c_int, _f32p, _f32p, _f32p, _f32p, _i64p,
                                            _i64p, _i64p, _i64p, C.c_float, C.c_float, C.c_float,
                                            C.c_float, C.c_float, C.c_int, C.c_int]),
        "orc_preprocess_imper_f64": (None, [C.c_int, _i64p, C.c_int, _f32p, _f32p, _f32p, _f32p, _i64p,
                                            _i64p, _i64p, _i64p, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_double, C.c_int, C.c_int]),
        "orc_a_min": (C.c_double, [C.c_double]),
        "orc_a_max": (C.c_double, [C.c_double]),
        "orc_b_min": (C.c_double, [C.c_double]),
        "orc_b_max": (C.c_double, [C.c_double]),
        "orc_g5comp": (C.c_double, [C.c_double, C.c_double]),
        "orc_bpm_f64": (C.c_int, [bc, C.c_int64, C.c_int64, _f64p, _f64p, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


orc = _load_oracle()


def ref_available():
    return os.path.exists(REF_PATH)


def p64(a):
    return a.ctypes.data_as(_f64p)


def p32(a):
    return a.ctypes.data_as(_f32p)


def tonal_f64(case, nthreads=0):
    n = case.nTime * case.nObs
    out = [np.empty(n, np.float64) for _ in range(3)]
    rc = orc.orc_tonal_f64(C.byref(case), *[p64(a) for a in out], nthreads)
    assert rc == 0, f"oracle tonal failed: {rc}"
    return [a.reshape(case.nTime, case.nObs) for a in out]


def tonal_f32(case, nthreads=0):
    n = case.nTime * case.nObs
    out = [np.empty(n, np.float32) for _ in range(3)]
    rc = orc.orc_tonal_f32(C.byref(case), *[p32(a) for a in out], nthreads)
    assert rc == 0, f"oracle tonal failed: {rc}"
    return [a.reshape(case.nTime, case.nObs) for a in out]


def tonal_bins(case, iobs):
    nE = sum(case.E[s] for s in range(case.nSurf))
    n = nE * case.nTime
    out = [np.empty(n, np.int32) for _ in range(4)]
    rc = orc.orc_tonal_bins(C.byref(case), iobs, *[a.ctypes.data_as(_i32p) for a in out])
    assert rc == 0, f"oracle bins failed: {rc}"
    return out


def bpm_f64(case, obs0=0, nobs=0, nthreads=0):
    n = nobs if nobs > 0 else case.nObs
    spl = np.empty(n * case.nf2 * 7, np.float64)
    oa = np.empty(n, np.float64)
    rc = orc.orc_bpm_f64(C.byref(case), obs0, nobs, p64(spl), p64(oa), nthreads)
    assert rc == 0, f"oracle bpm failed: {rc}"
    return spl.reshape(n, case.nf2, 7), oa


# ---- tonal band spectra (numpy restatement of the definition in include/acum.h) -------------
SIXTH_DOWN = 0.8908987181403393  # 2^(-1/6), the same literal as the library's
SIXTH_UP = 1.122462048309373     # 2^(1/6)
PREF2 = 2e-5 * 2e-5


def band_bins(nT, dt, fc):
    """[lo_b, hi_b) DFT-bin range of each band: bin k (k / (nT dt) Hz) lies in band b when
    edge_b <= k df < edge_{b+1}; inner edges are geometric means of adjacent centres."""
    fc = np.asarray(fc, np.float32).astype(np.float64)
    nbins = nT // 2
    df = 1.0 / (float(nT) * dt)
    edges = [fc[0] * SIXTH_DOWN] + [float(np.sqrt(fc[b - 1] * fc[b])) for b in range(1, fc.size)] + \
            [fc[-1] * SIXTH_UP]
    k = np.arange(1, nbins + 1, dtype=np.float64) * df

    def first(f):
        idx = np.nonzero(k >= f)[0]
        return int(idx[0]) + 1 if idx.size else nbins + 1
    lo = np.array([first(e) for e in edges[:-1]])
    hi = np.array([first(e) for e in edges[1:]])
    return lo, hi


def bin_energy(pA):
    """One-sided mean-square share of DFT bins k = 1 .. nT/2 of each column (mean removed)."""
    x = np.asarray(pA, np.float64)
    nT = x.shape[0]
    x = x - x.mean(axis=0)
    X = np.fft.rfft(x, axis=0)[1:nT // 2 + 1]
    w = np.full(X.shape[0], 2.0)
    if nT % 2 == 0:
        w[-1] = 1.0
    return w[:, None] * np.abs(X) ** 2 / float(nT) ** 2


def tonal_spectra(pA, dt, fc):
    """(spl [nObs, nBands], oaspl [nObs]) in dB re 20 uPa, FP64."""
    x = np.asarray(pA, np.float64)
    E = bin_energy(x)
    lo, hi = band_bins(x.shape[0], dt, fc)
    eb = np.stack([E[l - 1:h - 1].sum(axis=0) for l, h in zip(lo, hi)], axis=1)
    spl = 10.0 * np.log10(np.maximum(eb, 1e-10 * PREF2) / PREF2)
    var = ((x - x.mean(axis=0)) ** 2).mean(axis=0)
    return spl, 10.0 * np.log10(var / PREF2)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---- the reference's own host code (oracle/_ref), for pinning the oracle --------------------
_ref = None


def ref_lib():
    """libref_host.so, built by oracle/Makefile from the reference sources (None if absent)."""
    global _ref
    if _ref is None and ref_available():
        _ref = C.CDLL(REF_PATH)
    return _ref


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def ref_bpm(case, obs0=0, nobs=0):
    """findBBnoise (bb_core.cu:11) of the reference build on a broadband case (float)."""
    lib = ref_lib()
    nS, nT, nf2 = case.nSurf_BB, case.nTime, case.nf2
    nobs = nobs if nobs > 0 else case.nObs
    get = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy()
    nsect = get(case.nsect, nS)
    tot = int(nsect.sum())
    nst = get(case.nsect_t, nS)
    nm = get(case.nmach_t, nS)
    na = get(case.naoa_t, nS)
    nr = get(case.nre_t, nS)
    arrs = dict(
        span=_f32(get(case.span, nS)), omega=_f32(get(case.omega, nS)), ccw=_i32(get(case.ccw, nS)),
        psi=_f32(get(case.psi_offset, nS)), trip=_i32(get(case.trip, nS)),
        rot=_f32(get(case.rotate, 3 * nS)), tra=_f32(get(case.translate, 3 * nS)),
        xs=_i32(get(case.xsPtr, nS)), ds=_i32(get(case.dsPtr, nS)), nsect=_i32(nsect),
        xyzc=_f32(get(case.xyzcR, 4 * tot)), bb=_f32(get(case.bbdata, 7 * tot * nT)),
        nst=_i32(nst), nm=_i32(nm), na=_i32(na), nr=_i32(nr),
        aP=_i32(get(case.aPtr, nS)), mP=_i32(get(case.mPtr, nS)), rP=_i32(get(case.rPtr, nS)),
        sP=_i32(get(case.sPtr, nS)), dP=_i32(get(case.dPtr, nS)),
        st=_f32(get(case.sect_t, int(nst.sum()))), mt=_f32(get(case.mach_t, int(nm.sum()))),
        at=_f32(get(case.aoa_t, int(na.sum()))), rt=_f32(get(case.re_t, int(nr.sum()))),
        bt=_f32(get(case.bbdata_t, int((nst * nm * na * nr * 7).sum()))),
        vinf=_f32([case.vInf[0], case.vInf[1], case.vInf[2]]),
        robs=_f32(get(case.rObs, 3 * case.nObs)[3 * obs0:3 * (obs0 + nobs)]),
        tobs=_f32(np.zeros(nT)), ff=_f32(get(case.ff, nf2)))
    spl = np.full(nobs * nf2 * 7, -100.0, np.float32)
    oa = np.full(nobs, -100.0, np.float32)
    fn = lib._Z11findBBnoiseiiifPfS_PiS_S0_S_S_S0_S0_S0_S_S_S0_S0_S0_S0_S0_S0_S0_S0_S0_S_S_S_S_S_iS_S_S_iS_fS_S_
    fn.restype = None
    ip = lambda a: a.ctypes.data_as(_i32p)
    fp = lambda a: a.ctypes.data_as(_f32p)
    a = arrs
    fn(C.c_int(nS), C.c_int(nobs), C.c_int(nT), C.c_float(case.dTau), fp(a["span"]), fp(a["omega"]),
       ip(a["ccw"]), fp(a["psi"]), ip(a["trip"]), fp(a["rot"]), fp(a["tra"]), ip(a["xs"]), ip(a["ds"]),
       ip(a["nsect"]), fp(a["xyzc"]), fp(a["bb"]), ip(a["nst"]), ip(a["nm"]), ip(a["na"]), ip(a["nr"]),
       ip(a["aP"]), ip(a["mP"]), ip(a["rP"]), ip(a["sP"]), ip(a["dP"]), fp(a["st"]), fp(a["mt"]),
       fp(a["at"]), fp(a["rt"]), fp(a["bt"]), C.c_int(case.t_del), fp(a["vinf"]), fp(a["robs"]),
       fp(a["tobs"]), C.c_int(nf2), fp(a["ff"]), C.c_float(case.a0), fp(spl), fp(oa))
    return spl.reshape(nobs, nf2, 7), oa
