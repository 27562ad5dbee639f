"""The allocation-free spectra entry point (acum_tonal_spectra_noalloc), through the C-ABI, against
the allocating one (acum_tonal_spectra): two calls in a row, the second smaller, must give the same
band levels and overall levels bit for bit (the same kernels in the same order, on the library's
fixed device scratch instead of its workspace).  Sizes beyond the scratch, or more bands than the
kernel argument holds, are rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

import acum3_periodic_amd as acum
from parity import ISO_BANDS

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
_hip_lib = None


def _hip():
    """The HIP runtime, loaded on first use (collection on a host without ROCm does not need it)."""
    global _hip_lib
    if _hip_lib is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip_lib = h
    return _hip_lib


class DeviceArray:
    """Device memory owned by the test (the entry point under test allocates none)."""

    def __init__(self, shape, dtype=np.float32):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        assert _hip().hipMalloc(C.byref(p), self.nbytes) == 0, "hipMalloc failed"
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        assert _hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D) == 0

    def download(self):
        out = np.empty(self.shape, self.dtype)
        assert _hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            _hip().hipFree(self.ptr)
            self.ptr = None


@pytest.fixture(scope="module")
def ctx():
    if acum.device_count() < 1:
        pytest.fail("no GPU visible")
    c = acum.Context(0)
    yield c
    c.close()


def history(nT, nObs, seed):
    """Tones over broadband noise, one column per observer."""
    rng = np.random.default_rng(seed)
    t = np.arange(nT)[:, None]
    f = rng.integers(1, nT // 2, size=(1, nObs))
    a = rng.uniform(0.1, 10.0, size=(1, nObs))
    x = a * np.sin(2 * np.pi * f * t / nT + rng.uniform(0, 6.28, size=(1, nObs)))
    return (x + 0.05 * rng.standard_normal((nT, nObs)) + 3.0).astype(np.float32)


def noalloc(ctx, pA, dt, fc):
    nT, nObs = pA.shape
    d_pA = DeviceArray((nT, nObs))
    d_spl = DeviceArray((nObs, len(fc)))
    d_oa = DeviceArray((nObs,))
    try:
        d_pA.upload(pA)
        ctx.spectra_noalloc(d_pA.ptr, nT, nObs, dt, fc, d_spl.ptr, d_oa.ptr)
        return d_spl.download(), d_oa.download()
    finally:
        for d in (d_pA, d_spl, d_oa):
            d.free()


def test_noalloc_equals_the_allocating_call_twice(ctx):
    calls = [  # (steps, observers, dt, bands): the first nearly fills the 64 MiB scratch
        (2048, 8000, 2.45e-5, ISO_BANDS),
        (360, 70, 1.3e-4, ISO_BANDS[8:30]),
    ]
    for i, (nT, nObs, dt, fc) in enumerate(calls):
        assert (nT // 2 + 1) * nObs * 8 <= 64 << 20
        pA = history(nT, nObs, 100 + i)
        ref_spl, ref_oa = ctx.spectra(pA, dt, fc)
        spl, oa = noalloc(ctx, pA, dt, fc)
        assert np.array_equal(spl, ref_spl), f"call {i}: band levels differ"
        assert np.array_equal(oa, ref_oa), f"call {i}: overall levels differ"


def test_noalloc_rejects_what_its_scratch_cannot_hold(ctx):
    # device buffers of the requested sizes, so nothing could be read out of bounds even if the
    # size check were missing; the call must fail before any launch
    for nT, nObs, nb in ((4096, 4200, 34), (256, 64, 129)):
        fc = np.geomspace(10.0, 20000.0, nb).astype(np.float32)
        d_pA, d_spl, d_oa = DeviceArray((nT, nObs)), DeviceArray((nObs, nb)), DeviceArray((nObs,))
        try:
            with pytest.raises(acum.AcumError) as e:
                ctx.spectra_noalloc(d_pA.ptr, nT, nObs, 1e-4, fc, d_spl.ptr, d_oa.ptr)
            assert e.value.code == acum.ACUM_ERR_ARG
        finally:
            for d in (d_pA, d_spl, d_oa):
                d.free()
    # the context still works afterwards
    pA = history(64, 5, 7)
    spl, oa = noalloc(ctx, pA, 1e-3, ISO_BANDS[10:20])
    ref_spl, ref_oa = ctx.spectra(pA, 1e-3, ISO_BANDS[10:20])
    assert np.array_equal(spl, ref_spl) and np.array_equal(oa, ref_oa)
