"""Pins of the band-spectrum oracle (tests/oracle.py: band_bins, bin_energy, tonal_spectra) by
closed forms: a pure tone puts A^2/2 in its band, a Nyquist tone A^2, Parseval over bands that
cover every bin, and the energies of a brute-force DFT on short histories."""
import numpy as np
import pytest

import oracle as O

ISO = np.array([10, 12.5, 16, 20, 25, 31.5, 40, 50, 63, 80, 100, 125, 160, 200, 250, 315, 400, 500, 630,
                800, 1000, 1250, 1600, 2000, 2500, 3150, 4000, 5000, 6300, 8000, 10000, 12500, 16000,
                20000], np.float32)


@pytest.mark.parametrize("nT,k0", [(64, 5), (65, 17), (1024, 100)])
def test_pure_tone_in_one_band(nT, k0):
    dt = 1.0 / 40960.0
    A = 3.0
    n = np.arange(nT)
    p = A * np.cos(2 * np.pi * k0 * n / nT + 0.3) + 7.0
    spl, oa = O.tonal_spectra(p[:, None], dt, ISO)
    e = 10 ** (spl[0] / 10) * O.PREF2
    lo, hi = O.band_bins(nT, dt, ISO)
    b = [i for i in range(ISO.size) if lo[i] <= k0 < hi[i]][0]
    np.testing.assert_allclose(e[b], A * A / 2, rtol=1e-12)
    others = np.delete(e, b)
    assert np.all(others < 1e-20 + 1e-10 * O.PREF2 * 1.0001)
    np.testing.assert_allclose(oa[0], 10 * np.log10(A * A / 2 / O.PREF2), atol=1e-10)


def test_nyquist_tone_counts_once():
    nT, dt, A = 32, 1.0 / 640.0, 2.0
    p = A * (-1.0) ** np.arange(nT)
    E = O.bin_energy(p[:, None])
    np.testing.assert_allclose(E[-1, 0], A * A, rtol=1e-13)
    assert np.abs(E[:-1]).max() < 1e-24


def test_parseval_over_covering_bands():
    rng = np.random.default_rng(3)
    nT, dt = 360, 1.0 / 3600.0            # bins at 10 Hz .. 1800 Hz
    fc = np.array([8.0 * 2 ** (i / 3) for i in range(24)], np.float32)   # 8 Hz .. ~1.6 kHz
    lo, hi = O.band_bins(nT, dt, fc)
    assert lo[0] == 1 and hi[-1] == nT // 2 + 1 and np.all(lo[1:] == hi[:-1])
    p = rng.standard_normal((nT, 5))
    spl, oa = O.tonal_spectra(p, dt, fc)
    tot = (10 ** (spl / 10)).sum(axis=1)
    np.testing.assert_allclose(10 * np.log10(tot), oa, atol=1e-9)


@pytest.mark.parametrize("nT", [7, 8])
def test_energy_matches_brute_force_dft(nT):
    rng = np.random.default_rng(nT)
    p = rng.standard_normal((nT, 3))
    x = p - p.mean(axis=0)
    n = np.arange(nT)
    E = O.bin_energy(p)
    for k in range(1, nT // 2 + 1):
        X = (x * np.exp(-2j * np.pi * k * n / nT)[:, None]).sum(axis=0)
        w = 1.0 if 2 * k == nT else 2.0
        np.testing.assert_allclose(E[k - 1], w * np.abs(X) ** 2 / nT ** 2, rtol=1e-12)


def test_band_edges_are_geometric_means():
    lo, hi = O.band_bins(100, 1.0 / 1000.0, np.array([20.0, 40.0, 80.0], np.float32))
    # bins every 10 Hz; edges 20/2^(1/6)=17.8, sqrt(800)=28.3, sqrt(3200)=56.6, 80*2^(1/6)=89.8
    assert list(lo) == [2, 3, 6] and list(hi) == [3, 6, 9]
