"""GPU band spectra of tonal pressure histories against the FP64 numpy oracle, through the C-ABI
(acum_tonal_spectra, acum_tonal_fetch_spectra)."""
import numpy as np
import pytest

import acum3_periodic_amd as acum
import oracle as O

pytestmark = pytest.mark.gpu

TOL_DB = 1e-3   # FP64 Goertzel-Reinsch on the GPU vs FP64 FFT on the host; outputs are float dB
DEEP_DB = 120.0  # bands this far below the overall level are held to DEEP_TOL_DB: at -180 dB and
DEEP_TOL_DB = 0.05  # below, the O(N eps) rounding of either DFT is a visible share of the band


def close(spl, ospl, oa):
    deep = ospl < (oa[:, None] - DEEP_DB)
    err = np.abs(spl - ospl)
    return err[~deep].max(initial=0.0) <= TOL_DB and err[deep].max(initial=0.0) <= DEEP_TOL_DB

ISO = np.array([10, 12.5, 16, 20, 25, 31.5, 40, 50, 63, 80, 100, 125, 160, 200, 250, 315, 400, 500, 630,
                800, 1000, 1250, 1600, 2000, 2500, 3150, 4000, 5000, 6300, 8000, 10000, 12500, 16000,
                20000], np.float32)


@pytest.fixture(scope="module")
def ctx():
    if acum.device_count() < 1:
        pytest.fail("no GPU visible")
    c = acum.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("nT,nObs", [(36, 1), (37, 70), (64, 130), (1024, 65)])
def test_random_histories_match_oracle(ctx, nT, nObs):
    rng = np.random.default_rng(nT + nObs)
    p = (rng.standard_normal((nT, nObs)) * 10.0 + 3.0).astype(np.float32)
    dt = 1.0 / (40.0 * nT)          # bins every 40 Hz up to 20 nT Hz
    spl, oa = ctx.spectra(p, dt, ISO)
    ospl, ooa = O.tonal_spectra(p, dt, ISO)
    assert close(spl, ospl, ooa)
    assert np.abs(oa - ooa).max() <= TOL_DB
    # the overall level agrees with the host routine
    np.testing.assert_allclose(oa, acum.oaspl(p), atol=1e-4)


def test_low_bin_tone_and_floor(ctx):
    """A tone in bin 1 of a long history (Goertzel's least favourable case) and bands below the
    first bin, which hold no bin and read the -100 dB floor."""
    nT = 2048
    n = np.arange(nT)
    p = (2.0 * np.cos(2 * np.pi * n / nT + 0.1) + 0.01 * np.cos(2 * np.pi * 300 * n / nT)).astype(np.float32)[:, None]
    dt = 1.0 / (100.0 * nT)         # bin 1 at 100 Hz
    spl, oa = ctx.spectra(p, dt, ISO)
    ospl, ooa = O.tonal_spectra(p, dt, ISO)
    assert close(spl, ospl, ooa)
    assert np.abs(oa - ooa).max() <= TOL_DB
    assert np.all(spl[0, :9] == np.float32(-100.0))


def test_staged_case_spectra(ctx):
    run = acum.Run.synth(acum.synth_params(2, nChord=5, nSpan=4, nTime=64, nObsAz=6, nObsEl=3)).preprocess()
    c = run.tonal_case(acum.MIXED)
    h = ctx.stage([c])
    ctx.run(h, acum.MIXED)
    pA = ctx.fetch(h, 0, c.nTime, c.nObs)[2]
    dt = c.dPsi / c.oM
    a = ctx.fetch_spectra(h, 0, c.nObs, ISO)
    b = ctx.spectra(pA, dt, ISO)
    ctx.release(h)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    ospl, ooa = O.tonal_spectra(pA, dt, ISO)
    assert close(a[0], ospl, ooa)
    assert np.abs(a[1] - ooa).max() <= TOL_DB


def test_spectra_argument_errors(ctx):
    p = np.ones((16, 2), np.float32)
    with pytest.raises(acum.AcumError) as e:
        ctx.spectra(p, 1e-3, ISO[::-1].copy())
    assert e.value.code == acum.ACUM_ERR_ARG
    with pytest.raises(acum.AcumError) as e:
        ctx.spectra(p, -1.0, ISO)
    assert e.value.code == acum.ACUM_ERR_ARG
    with pytest.raises(acum.AcumError) as e:
        ctx.spectra(p[:1], 1e-3, ISO)
    assert e.value.code == acum.ACUM_ERR_ARG
    ctx.spectra(p, 1e-3, ISO)
