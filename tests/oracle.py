"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the host library compiled from
the reference's own sources (oracle/_ref/libref_host.so).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

import acum3_periodic_amd as acum

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
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
        "orc_obs_time": (C.c_double, [tc, _f32p, _f32p, C.c_int64]),
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


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
