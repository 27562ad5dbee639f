"""Helpers of the GPU parity tests: observer subsets of a tonal case (the oracle runs on a few
observers of a full-size case; observers are independent, so a subset's histories are exactly
those columns of the full run) and the band-level comparison of a history against the oracle's."""
import atexit
import ctypes as C
import json
import os

import numpy as np

import acum3_periodic_amd as acum
import oracle as O

# ISO 1/3-octave band centres, 10 Hz - 20 kHz
ISO_BANDS = np.array((10, 12.5, 16, 20, 25, 31.5, 40, 50, 63, 80, 100, 125, 160, 200, 250, 315, 400,
                      500, 630, 800, 1000, 1250, 1600, 2000, 2500, 3150, 4000, 5000, 6300, 8000, 10000,
                      12500, 16000, 20000), np.float32)
FLOOR_DB = -100.0


def spread_observers(nObs, n):
    """n observer indices spread evenly over 0 .. nObs-1 (first and last included)."""
    n = min(n, nObs)
    return np.unique(np.round(np.linspace(0, nObs - 1, n)).astype(np.int64))


def obs_subset(case, idx, precision=None):
    """A TonalCase over the same element data with only the observers `idx` (in that order)."""
    rObs = np.ctypeslib.as_array(case.rObs, shape=(3 * case.nObs,)).reshape(-1, 3)
    sub_obs = np.ascontiguousarray(rObs[np.asarray(idx)], np.float32).reshape(-1)
    sub = acum.TonalCase()
    C.pointer(sub)[0] = case
    sub.nObs = len(idx)
    sub.rObs = sub_obs.ctypes.data_as(C.POINTER(C.c_float))
    if precision is not None:
        sub.precision = precision
    sub._keep = (case, sub_obs)
    return sub


def band_levels(pA, dt, fc=ISO_BANDS):
    """(band SPL [nObs, nBands], OASPL [nObs]) of a [nT, nObs] history, FP64 numpy restatement."""
    return O.tonal_spectra(np.asarray(pA, np.float64), dt, fc)


def band_error(spl, ref_spl, floor=FLOOR_DB + 1.0):
    """Largest |spl - ref| over the bands whose reference level is above the -100 dB floor, and the
    (observer, band) where it occurs; bands at the floor hold no energy in either history."""
    spl = np.asarray(spl, np.float64)
    ref = np.asarray(ref_spl, np.float64)
    live = ref > floor
    if not live.any():
        return 0.0, None
    d = np.where(live, np.abs(spl - ref), 0.0)
    i = np.unravel_index(int(np.argmax(d)), d.shape)
    return float(d[i]), (int(i[0]), int(i[1]))


# Measured errors of the parity tests, written as JSON to the file named by ACUM_PARITY_REPORT
# (when set) at exit, so a run can show how far below its bars each check sits.
_measured = {}


def report(test, **values):
    _measured.setdefault(test, {}).update({k: (float(v) if isinstance(v, (float, np.floating)) else v)
                                           for k, v in values.items()})
    path = os.environ.get("ACUM_PARITY_REPORT")
    if path:
        print(f"[parity] {test}: {json.dumps(_measured[test])}", flush=True)


@atexit.register
def _write_report():
    path = os.environ.get("ACUM_PARITY_REPORT")
    if path and _measured:
        with open(path, "w") as f:
            json.dump(_measured, f, indent=1)
