"""Device helpers of the tonal kernels: sqrt_pos must be correctly rounded (the canonical observer
time, and so every bin index, depends on it), rcp_nr within 1 ulp, and the XCC_ID register the
work queues are picked by must spread the workgroups over the XCDs."""
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


def test_xcc_ids_spread_over_the_xcds():
    """The eval kernel starts each wave on the work queue of its XCD (hardware register XCC_ID);
    on an MI355X the workgroups of a launch see all eight XCDs."""
    lib = C.CDLL(PROBE)
    lib.probe_xcc.argtypes = [C.POINTER(C.c_int), C.c_int]
    lib.probe_xcc.restype = C.c_int
    n = 4096
    ids = np.zeros(n, np.int32)
    assert lib.probe_xcc(ids.ctypes.data_as(C.POINTER(C.c_int)), n) == 0
    counts = np.bincount(ids, minlength=16)
    print("XCC_ID counts:", counts.tolist())
    assert ids.min() >= 0 and ids.max() < 8
    assert (counts[:8] > 0).sum() == 8


def test_mfma_4x4x1_lane_layout():
    """v_mfma_f32_4x4x1f32 as the tonal kernel uses it: 16 blocks of four lanes; lane l = 4b + j
    supplies row j of block b's A and column j of its B, and receives column j of block b's
    result in its four accumulators (one per row)."""
    lib = C.CDLL(PROBE)
    fp = C.POINTER(C.c_float)
    lib.probe_mfma4x4.argtypes = [fp, fp, fp]
    lib.probe_mfma4x4.restype = C.c_int
    rng = np.random.default_rng(5)
    a = rng.integers(-8, 9, size=(3, 64)).astype(np.float32)   # small integers: exact products
    b = rng.integers(-8, 9, size=(3, 64)).astype(np.float32)
    out = np.zeros((64, 4), np.float32)
    assert lib.probe_mfma4x4(a.ctypes.data_as(fp), b.ctypes.data_as(fp), out.ctypes.data_as(fp)) == 0
    want = np.zeros((64, 4), np.float32)
    for l in range(64):
        blk, j = l // 4, l % 4
        for v in range(4):
            want[l, v] = sum(a[k, 4 * blk + v] * b[k, 4 * blk + j] for k in range(3))
    assert np.array_equal(out, want), f"layout differs:\n{out[:8]}\n{want[:8]}"


def test_rec_products64_row_broadcast():
    """rec_products64 as the FP64 tonal kernel uses it: lane l holds field l % 16 of its 16-lane
    row's record, and each lane gets its row's record dotted into its own r: rm = d - v.r,
    nVd = aM1 - vd.r, dL = l.r, nLd = aLM - ld.r and cT (fields 12, 13, 14)."""
    lib = C.CDLL(PROBE)
    dp = C.POINTER(C.c_double)
    lib.probe_rec64.argtypes = [dp, dp, dp, dp]
    lib.probe_rec64.restype = C.c_int
    rng = np.random.default_rng(7)
    f = rng.integers(-9, 10, size=64).astype(np.float64)  # small integers: every sum exact
    r = rng.integers(-9, 10, size=(3, 64)).astype(np.float64)
    d = rng.integers(-99, 100, size=64).astype(np.float64)
    out = np.zeros((64, 5))
    assert lib.probe_rec64(f.ctypes.data_as(dp), r.ctypes.data_as(dp), d.ctypes.data_as(dp),
                           out.ctypes.data_as(dp)) == 0
    want = np.zeros((64, 5))
    for l in range(64):
        rec = f[16 * (l // 16):16 * (l // 16) + 16]
        dot = lambda j: sum(rec[3 * j + c] * r[c, l] for c in range(3))
        want[l] = (d[l] - dot(0), rec[12] - dot(1), dot(2), rec[13] - dot(3), rec[14])
    assert np.array_equal(out, want), f"broadcast differs:\n{out[:4]}\n{want[:4]}"
