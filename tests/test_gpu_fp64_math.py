"""The FP64 helpers of the tonal kernels on the device: sqrt_pos must be correctly rounded (the
canonical observer time, and so every bin index, depends on it) and rcp_nr within 1 ulp."""
import ctypes as C
import os

import numpy as np
import pytest

import acum3_periodic_amd as acum

pytestmark = pytest.mark.gpu
PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "probe", "libprobe.so")


def test_sqrt_correctly_rounded_and_rcp():
    if acum.device_count() < 1:
        pytest.fail("no GPU visible")
    lib = C.CDLL(PROBE)
    p = C.POINTER(C.c_double)
    lib.probe_fp64.argtypes = [p, p, p, C.c_int64]
    lib.probe_fp64.restype = C.c_int
    rng = np.random.default_rng(2024)
    n = 1 << 22
    x = 10.0 ** rng.uniform(-12, 12, n)
    x[:4] = [1.0, 4.0, 2.0, 1e12]
    x[4:4 + 1000] = np.nextafter(np.arange(1, 1001, dtype=np.float64) ** 2, 0)  # just below squares
    s = np.empty(n)
    r = np.empty(n)
    assert lib.probe_fp64(x.ctypes.data_as(p), s.ctypes.data_as(p), r.ctypes.data_as(p), n) == 0
    bad = np.flatnonzero(s != np.sqrt(x))
    assert bad.size == 0, f"{bad.size} sqrt results not correctly rounded, e.g. x={x[bad[0]]!r}"
    ulp = np.abs(r - 1.0 / x) / np.spacing(1.0 / x)
    assert ulp.max() <= 1.0
