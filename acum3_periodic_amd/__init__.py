"""Python binding of the MI355X acoustic hot path (include/acum.h) over ctypes.

The library ``lib/libacum_hip.so`` (HIP kernels for gfx950 + the host file contract) is loaded
from this package directory; importing fails loudly when it has not been built.  Structures mirror
the C-ABI field for field.  The tonal path replaces the reference's device launch of
findPressureTimeHistory (main.cu:156-247, core.cu:18); the broadband path replaces its host call
of findBBnoise (main.cu:150, bb_core.cu:11).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ACUM_LIB_PATH selects another build of the same library (the sanitizer build of the host code)
LIB_PATH = os.environ.get("ACUM_LIB_PATH") or os.path.join(_HERE, "lib", "libacum_hip.so")
DRIVER_PATH = os.path.join(_HERE, "bin", "acum")

ACUM_OK, ACUM_ERR_ARG, ACUM_ERR_IO, ACUM_ERR_DEVICE, ACUM_ERR_NOMEM, ACUM_ERR_NUMERIC, ACUM_ERR_PLAN = (
    0, -1, -2, -3, -4, -5, -6)
FP64, MIXED = 0, 1

_f32p = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class TonalCase(C.Structure):
    _fields_ = [("nSurf", C.c_int64), ("nTime", C.c_int64), ("nObs", C.c_int64), ("E", _i64p),
                ("X", _f32p), ("N", _f32p), ("V", _f32p), ("PR", _f32p),
                ("XPtr", _i64p), ("NPtr", _i64p), ("VPtr", _i64p), ("PRPtr", _i64p),
                ("rObs", _f32p), ("dPsi", C.c_double), ("dTau", C.c_double), ("a0", C.c_double),
                ("rhoRef", C.c_double), ("oM", C.c_double), ("vInf", C.c_double * 3),
                ("precision", C.c_int32), ("reserved", C.c_int32)]


class BpmCase(C.Structure):
    _fields_ = [("nSurf_BB", C.c_int32), ("nTime", C.c_int32), ("nObs", C.c_int32), ("nf2", C.c_int32),
                ("t_del", C.c_int32), ("reserved", C.c_int32), ("dTau", C.c_double), ("a0", C.c_double),
                ("vInf", C.c_double * 3), ("rObs", _f32p), ("ff", _f32p), ("span", _f32p),
                ("omega", _f32p), ("psi_offset", _f32p), ("ccw", _i32p), ("trip", _i32p),
                ("rotate", _f32p), ("translate", _f32p), ("nsect", _i32p), ("xsPtr", _i64p),
                ("dsPtr", _i64p), ("xyzcR", _f32p), ("bbdata", _f32p), ("nsect_t", _i32p),
                ("nmach_t", _i32p), ("naoa_t", _i32p), ("nre_t", _i32p), ("aPtr", _i64p),
                ("mPtr", _i64p), ("rPtr", _i64p), ("sPtr", _i64p), ("dPtr", _i64p),
                ("sect_t", _f32p), ("mach_t", _f32p), ("aoa_t", _f32p), ("re_t", _f32p),
                ("bbdata_t", _f32p)]


class Timing(C.Structure):
    _fields_ = [("prep_ms", C.c_double), ("eval_ms", C.c_double), ("reduce_ms", C.c_double),
                ("total_ms", C.c_double), ("evals", C.c_int64), ("partial_bytes", C.c_int64),
                ("n_partials", C.c_int32), ("n_items", C.c_int32), ("launches", C.c_int32),
                ("motion_widened_cases", C.c_int32), ("eval_launches", C.c_int32),
                ("tune_cases", C.c_int32), ("tune_hits", C.c_int32), ("tune_passes", C.c_int32)]


class Tune(C.Structure):
    _fields_ = [("window", C.c_int32), ("max_spread", C.c_int32), ("max_tile", C.c_int32),
                ("groups", C.c_int32), ("max_chunk", C.c_int32), ("waves", C.c_int32),
                ("motion_permille", C.c_int32)]


class Inputs(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("a0", "rhoRef", "pRef", "oM", "dPsi_deg", "Minf", "Mref",
                                           "xy_angle", "xz_angle", "CFDscale", "OBSscale")] + \
               [(n, C.c_int32) for n in ("nSurf", "nTime", "periodic", "lowpass", "impermeable",
                                          "BB_noise")] + \
               [("dTau", C.c_double), ("dPsi", C.c_double), ("vInf", C.c_double * 3)]


class SynthParams(C.Structure):
    _fields_ = [("seed", C.c_uint64)] + \
               [(n, C.c_int32) for n in ("nBlades", "nChord", "nSpan", "nTime", "coaxial", "obs_layout",
                                          "nObsAz", "nObsEl", "nObs", "cp_model", "bb", "nSect",
                                          "t_del", "trip", "nBands")] + \
               [(n, C.c_double) for n in ("R", "chord", "rootCut", "oM", "a0", "rhoRef", "pRef", "Minf",
                                           "xy_angle", "xz_angle", "obsRadius", "obsElMin", "obsElMax",
                                           "cpNoise", "coaxSep", "twistRoot", "twistTip")]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "acum_device_count": (C.c_int, []),
        "acum_ctx_create": (C.c_int, [C.POINTER(P), C.c_int]),
        "acum_ctx_destroy": (None, [P]),
        "acum_last_error": (C.c_char_p, [P]),
        "acum_tonal": (C.c_int, [P, C.POINTER(TonalCase), _f32p, _f32p, _f32p]),
        "acum_tonal_plain": (C.c_int, [P, C.POINTER(TonalCase), _f32p, _f32p, _f32p]),
        "acum_tonal_batch": (C.c_int, [P, C.POINTER(TonalCase), C.c_int32, C.POINTER(_f32p),
                                       C.POINTER(_f32p), C.POINTER(_f32p)]),
        "acum_tonal_stage": (C.c_int, [P, C.POINTER(TonalCase), C.c_int32, _i32p]),
        "acum_tonal_run": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
        "acum_tonal_fetch": (C.c_int, [P, C.c_int32, C.c_int32, _f32p, _f32p, _f32p]),
        "acum_tonal_release": (C.c_int, [P, C.c_int32]),
        "acum_tonal_bins": (C.c_int, [P, C.POINTER(TonalCase), C.c_int64, _i32p, _i32p, _i32p, _i32p]),
        "acum_get_timing": (C.c_int, [P, C.POINTER(Timing)]),
        "acum_get_tuning": (C.c_int, [P, C.c_int32, C.POINTER(Tune)]),
        "acum_set_tuning": (C.c_int, [P, C.c_int32, C.POINTER(Tune)]),
        "acum_autotune": (C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(Tune), C.POINTER(C.c_double)]),
        "acum_bpm": (C.c_int, [P, C.POINTER(BpmCase), _f32p, _f32p]),
        "acum_host_error": (C.c_char_p, []),
        "acum_run_load": (C.c_int, [C.c_char_p, C.POINTER(P)]),
        "acum_run_free": (None, [P]),
        "acum_run_preprocess": (C.c_int, [P]),
        "acum_preprocess_gpu": (C.c_int, [P, P]),
        "acum_run_inputs": (C.c_int, [P, C.POINTER(Inputs)]),
        "acum_run_tonal_case": (C.c_int, [P, C.c_int32, C.POINTER(TonalCase)]),
        "acum_run_bpm_case": (C.c_int, [P, C.POINTER(BpmCase)]),
        "acum_run_write_tonal": (C.c_int, [P, C.c_char_p, _f32p, _f32p, _f32p]),
        "acum_run_write_bb": (C.c_int, [P, C.c_char_p, _f32p, _f32p]),
        "acum_run_save": (C.c_int, [P, C.c_char_p]),
        "acum_run_surface_raw": (C.c_int, [P, C.c_int32, _f32p]),
        "acum_synth_defaults": (C.c_int, [C.c_int32, C.POINTER(SynthParams)]),
        "acum_synth_batch_case": (C.c_int, [C.c_uint64, C.c_int32, C.POINTER(SynthParams)]),
        "acum_synth": (C.c_int, [C.POINTER(SynthParams), C.POINTER(P)]),
        "acum_oaspl": (C.c_int, [C.c_int64, C.c_int64, _f32p, C.POINTER(C.c_double)]),
        "acum_tonal_spectra": (C.c_int, [P, C.c_int64, C.c_int64, _f32p, C.c_double, C.c_int32, _f32p, _f32p,
                                         _f32p]),
        "acum_tonal_fetch_spectra": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, _f32p, _f32p, _f32p]),
        "acum_run_stage": (C.c_int, [P, P, _i32p]),
        "acum_run_save_cache": (C.c_int, [P, C.c_char_p]),
        "acum_tonal_spectra_noalloc": (C.c_int, [P, C.c_int64, C.c_int64, P, C.c_double, C.c_int32, _f32p, P, P]),
        "acum_get_case_tuning": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(Tune)]),
        "acum_tune_cache_save": (C.c_int, [P, C.c_char_p]),
        "acum_tune_cache_load": (C.c_int, [P, C.c_char_p]),
        "acum_tune_cache_clear": (C.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()

# Names every C-ABI entry point declared in include/acum.h.
EXPORTS = ("acum_device_count", "acum_ctx_create", "acum_ctx_destroy", "acum_last_error", "acum_tonal",
           "acum_tonal_plain",
           "acum_tonal_batch", "acum_tonal_stage", "acum_tonal_run", "acum_tonal_fetch",
           "acum_tonal_release", "acum_tonal_bins", "acum_get_timing", "acum_get_tuning",
           "acum_set_tuning", "acum_autotune", "acum_bpm", "acum_host_error", "acum_run_load",
           "acum_run_free", "acum_run_preprocess", "acum_preprocess_gpu", "acum_run_inputs", "acum_run_tonal_case",
           "acum_run_bpm_case", "acum_run_write_tonal", "acum_run_write_bb", "acum_run_save",
           "acum_run_surface_raw", "acum_synth_defaults", "acum_synth_batch_case", "acum_synth",
           "acum_oaspl", "acum_tonal_spectra", "acum_tonal_fetch_spectra",
           "acum_run_stage", "acum_run_save_cache", "acum_tonal_spectra_noalloc", "acum_get_case_tuning",
           "acum_tune_cache_save", "acum_tune_cache_load", "acum_tune_cache_clear")


class AcumError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"acum error {code}: {msg}")
        self.code = code


def _host_check(rc):
    if rc != ACUM_OK:
        raise AcumError(rc, lib.acum_host_error().decode())


def fptr(a):
    return a.ctypes.data_as(_f32p)


def synth_params(config, **over):
    p = SynthParams()
    _host_check(lib.acum_synth_defaults(config, C.byref(p)))
    for k, v in over.items():
        setattr(p, k, v)
    return p


def batch_params(seed, icase, **over):
    p = SynthParams()
    _host_check(lib.acum_synth_batch_case(seed, icase, C.byref(p)))
    for k, v in over.items():
        setattr(p, k, v)
    return p


class Run:
    """One run directory's worth of inputs (the reference's globals, variables.cu:1-64)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def load(cls, path):
        h = C.c_void_p()
        _host_check(lib.acum_run_load(str(path).encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def synth(cls, params):
        h = C.c_void_p()
        _host_check(lib.acum_synth(C.byref(params), C.byref(h)))
        return cls(h.value)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.acum_run_free(self._h)
            self._h = C.c_void_p()

    def preprocess(self):
        _host_check(lib.acum_run_preprocess(self._h))
        return self

    def preprocess_gpu(self, ctx):
        """The same preprocessing on the context's GPU (bit-identical to preprocess())."""
        ctx._check(lib.acum_preprocess_gpu(ctx._h, self._h))
        return self

    def inputs(self):
        i = Inputs()
        _host_check(lib.acum_run_inputs(self._h, C.byref(i)))
        return i

    def save(self, path):
        _host_check(lib.acum_run_save(self._h, str(path).encode()))

    def save_cache(self, path):
        """Binary caches of the parsed surface files of run directory `path` (as read)."""
        _host_check(lib.acum_run_save_cache(self._h, str(path).encode()))

    def tonal_case(self, precision=MIXED):
        c = TonalCase()
        _host_check(lib.acum_run_tonal_case(self._h, precision, C.byref(c)))
        c._keep = self
        return c

    def bpm_case(self):
        c = BpmCase()
        _host_check(lib.acum_run_bpm_case(self._h, C.byref(c)))
        c._keep = self
        return c

    def surface_raw(self, isurf):
        c = self.tonal_case_raw_sizes()
        E = c[0][isurf]
        out = np.empty((c[1], E, 8), np.float32)
        _host_check(lib.acum_run_surface_raw(self._h, isurf, fptr(out)))
        return out

    def tonal_case_raw_sizes(self):
        i = self.inputs()
        # E per surface is exposed through a preprocessed case view only; the raw view reads
        # it from the tonal case struct without requiring preprocessing
        c = TonalCase()
        lib.acum_run_tonal_case(self._h, MIXED, C.byref(c))
        return [c.E[s] for s in range(i.nSurf)], i.nTime

    def write_tonal(self, path, pT, pL, pA):
        _host_check(lib.acum_run_write_tonal(self._h, str(path).encode(), fptr(pT), fptr(pL), fptr(pA)))

    def write_bb(self, path, oaspl, spl):
        _host_check(lib.acum_run_write_bb(self._h, str(path).encode(), fptr(oaspl), fptr(spl)))


def case_arrays(c):
    """numpy views of a tonal case's arrays (reference AoS layout)."""
    nS, nT = c.nSurf, c.nTime
    E = np.array([c.E[s] for s in range(nS)], np.int64)
    q = int(E.sum()) * nT
    v = lambda p, n: np.ctypeslib.as_array(p, shape=(n,))
    return dict(E=E, X=v(c.X, 4 * q), N=v(c.N, 6 * q), V=v(c.V, 6 * q), PR=v(c.PR, 2 * q),
                rObs=v(c.rObs, 3 * c.nObs).reshape(-1, 3))


def oaspl(pA):
    """Tonal overall level per observer (dB re 20 uPa) of a [nTime, nObs] pressure history."""
    pA = np.ascontiguousarray(pA, np.float32)
    out = np.empty(pA.shape[1], np.float64)
    _host_check(lib.acum_oaspl(pA.shape[0], pA.shape[1], fptr(pA), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def device_count():
    return lib.acum_device_count()


class Context:
    """One GPU context (acum_ctx); calls are synchronous on its stream."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib.acum_ctx_create(C.byref(h), device)
        if rc != ACUM_OK:
            raise AcumError(rc, f"cannot create a context on device {device}")
        self._h = h

    def close(self):
        if self._h and self._h.value:
            lib.acum_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != ACUM_OK:
            raise AcumError(rc, lib.acum_last_error(self._h).decode())

    def tonal(self, case):
        n = case.nTime * case.nObs
        out = [np.empty(n, np.float32) for _ in range(3)]
        self._check(lib.acum_tonal(self._h, C.byref(case), *[fptr(a) for a in out]))
        return [a.reshape(case.nTime, case.nObs) for a in out]

    def tonal_plain(self, case):
        """The plain FP64 path (the reference's kernel structure in FP64, the measurement
        baseline); the case's precision field is not used."""
        n = case.nTime * case.nObs
        out = [np.empty(n, np.float32) for _ in range(3)]
        self._check(lib.acum_tonal_plain(self._h, C.byref(case), *[fptr(a) for a in out]))
        return [a.reshape(case.nTime, case.nObs) for a in out]

    def tonal_batch(self, cases):
        arr = (TonalCase * len(cases))(*cases)
        outs = [[np.empty(c.nTime * c.nObs, np.float32) for _ in range(3)] for c in cases]
        ptr = [(_f32p * len(cases))(*[fptr(o[i]) for o in outs]) for i in range(3)]
        self._check(lib.acum_tonal_batch(self._h, arr, len(cases), *ptr))
        return [[a.reshape(c.nTime, c.nObs) for a in o] for o, c in zip(outs, cases)]

    def stage(self, cases):
        arr = (TonalCase * len(cases))(*cases)
        h = C.c_int32()
        self._check(lib.acum_tonal_stage(self._h, arr, len(cases), C.byref(h)))
        return h.value

    def run(self, handle, precision, repeats=1, graph=False):
        self._check(lib.acum_tonal_run(self._h, handle, precision, repeats, 1 if graph else 0))
        return self.timing()

    def fetch(self, handle, icase, nTime, nObs):
        out = [np.empty(nTime * nObs, np.float32) for _ in range(3)]
        self._check(lib.acum_tonal_fetch(self._h, handle, icase, *[fptr(a) for a in out]))
        return [a.reshape(nTime, nObs) for a in out]

    def release(self, handle):
        self._check(lib.acum_tonal_release(self._h, handle))

    def stage_run(self, run):
        """Stages a run's tonal case; a run not yet preprocessed is preprocessed on the GPU in
        place (its element data stay on the device).  Returns the handle."""
        h = C.c_int32()
        self._check(lib.acum_run_stage(self._h, run._h, C.byref(h)))
        return h.value

    def spectra(self, pA, dt, fc):
        """Band levels [nObs, nBands] and overall levels [nObs] (dB re 20 uPa) of a [nTime, nObs]
        pressure history with observer-time step dt (s), bands centred at fc (Hz)."""
        pA = np.ascontiguousarray(pA, np.float32)
        fc = np.ascontiguousarray(fc, np.float32)
        nT, nO = pA.shape
        spl = np.empty((nO, fc.size), np.float32)
        oa = np.empty(nO, np.float32)
        self._check(lib.acum_tonal_spectra(self._h, nT, nO, fptr(pA), float(dt), fc.size, fptr(fc), fptr(spl),
                                           fptr(oa)))
        return spl, oa

    def fetch_spectra(self, handle, icase, nObs, fc):
        """spectra() of case icase of a staged batch, from its device-resident pA."""
        fc = np.ascontiguousarray(fc, np.float32)
        spl = np.empty((nObs, fc.size), np.float32)
        oa = np.empty(nObs, np.float32)
        self._check(lib.acum_tonal_fetch_spectra(self._h, handle, icase, fc.size, fptr(fc), fptr(spl), fptr(oa)))
        return spl, oa

    def spectra_noalloc(self, d_pA, nTime, nObs, dt, fc, d_spl, d_oaspl):
        """acum_tonal_spectra_noalloc: d_pA [nTime, nObs], d_spl [nObs, nBands] and d_oaspl [nObs]
        are device addresses (ints); the call allocates nothing."""
        fc = np.ascontiguousarray(fc, np.float32)
        self._check(lib.acum_tonal_spectra_noalloc(self._h, nTime, nObs, C.c_void_p(d_pA), float(dt), fc.size,
                                                   fptr(fc), C.c_void_p(d_spl), C.c_void_p(d_oaspl)))

    def bins(self, case, iobs):
        nE = sum(case.E[s] for s in range(case.nSurf))
        n = nE * case.nTime
        out = [np.empty(n, np.int32) for _ in range(4)]
        self._check(lib.acum_tonal_bins(self._h, C.byref(case), iobs,
                                        *[a.ctypes.data_as(_i32p) for a in out]))
        return out

    def timing(self):
        t = Timing()
        self._check(lib.acum_get_timing(self._h, C.byref(t)))
        return t

    def get_tuning(self, precision):
        t = Tune()
        self._check(lib.acum_get_tuning(self._h, precision, C.byref(t)))
        return t

    def set_tuning(self, precision, **kw):
        t = self.get_tuning(precision)
        for k, v in kw.items():
            setattr(t, k, v)
        self._check(lib.acum_set_tuning(self._h, precision, C.byref(t)))

    def autotune(self, handle, precision):
        """Tunes the buckets of the staged cases not yet in the context's cache; returns case 0's
        shape and its time per pass on the tuning subset."""
        t = Tune()
        ms = C.c_double()
        self._check(lib.acum_autotune(self._h, handle, precision, C.byref(t), C.byref(ms)))
        return t, ms.value

    def case_tuning(self, handle, icase, precision):
        t = Tune()
        self._check(lib.acum_get_case_tuning(self._h, handle, icase, precision, C.byref(t)))
        return t

    def tune_cache_save(self, path):
        self._check(lib.acum_tune_cache_save(self._h, str(path).encode()))

    def tune_cache_load(self, path):
        self._check(lib.acum_tune_cache_load(self._h, str(path).encode()))

    def tune_cache_clear(self):
        self._check(lib.acum_tune_cache_clear(self._h))

    def bpm(self, case):
        spl = np.full(case.nObs * case.nf2 * 7, -100.0, np.float32)
        oa = np.full(case.nObs, -100.0, np.float32)
        self._check(lib.acum_bpm(self._h, C.byref(case), fptr(spl), fptr(oa)))
        return spl.reshape(case.nObs, case.nf2, 7), oa
