"""Pins of the broadband oracle (oracle/orc_bpm.cpp, FP64) without a GPU.

  * closed forms of the BPM shape functions (NASA RP-1218): A_min(0) = A_max(0) =
    sqrt(67.552) - 8.219, B_min(0) = B_max(0) = sqrt(16.888) - 4.109, continuity at the published
    breakpoints; g5comp continuity;
  * the whole findBBnoise (bb_core.cu:11-586) against the reference's own bb_core.cu/bb_utils.cu
    compiled from source into oracle/_ref (float arithmetic): 1/3-octave SPL and OASPL within
    1e-3 dB on tripped/untripped, empirical/table/file boundary layers and both band tables.
"""
import math

import numpy as np
import pytest

import acum3_periodic_amd as acum
import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference sources)")


def test_shape_function_zeros():
    a0 = math.sqrt(67.552) - 8.219
    b0 = math.sqrt(16.888) - 4.109
    assert abs(O.orc.orc_a_min(0.0) - a0) < 1e-12
    assert abs(O.orc.orc_a_max(0.0) - a0) < 1e-12
    assert abs(O.orc.orc_b_min(0.0) - b0) < 1e-12
    assert abs(O.orc.orc_b_max(0.0) - b0) < 1e-12


@pytest.mark.parametrize("fn,breaks", [("orc_a_min", (0.204, 0.244)), ("orc_a_max", (0.13, 0.321)),
                                       ("orc_b_min", (0.13, 0.145)), ("orc_b_max", (0.10, 0.187))])
def test_shape_function_continuity(fn, breaks):
    """Piecewise forms meet at the published breakpoints to within the rounding of the
    published coefficients (about 0.02 dB)."""
    f = getattr(O.orc, fn)
    for b in breaks:
        lo, hi = f(b * (1 - 1e-9)), f(b * (1 + 1e-9))
        assert abs(lo - hi) < 0.05, (fn, b, lo, hi)


def test_g5comp_continuity():
    for hdav in (0.2, 0.5, 1.0, 2.0):
        prev = None
        for eta in np.linspace(-0.5, 0.5, 201):
            v = O.orc.orc_g5comp(hdav, float(eta))
            if prev is not None:
                assert abs(v - prev) < 3.0
            prev = v


VARIANTS = [
    dict(trip=1, t_del=2),
    dict(trip=0, t_del=2),
    dict(trip=0, t_del=1),
    dict(trip=0, t_del=0, nBands=42),
    dict(trip=1, t_del=0),
]


@needs_ref
@pytest.mark.parametrize("over", VARIANTS, ids=lambda o: "-".join(f"{k}{v}" for k, v in o.items()))
def test_bpm_oracle_matches_reference_build(over):
    p = dict(nSect=12, nTime=20, nObsAz=4, nObsEl=3)
    p.update(over)
    c = acum.Run.synth(acum.synth_params(3, **p)).bpm_case()
    spl, oa = O.bpm_f64(c)
    rspl, roa = O.ref_bpm(c)
    d = np.abs(spl - rspl.astype(np.float64))
    assert d.max() < 1e-3, f"SPL differs by {d.max():.2e} dB at {np.unravel_index(d.argmax(), d.shape)}"
    assert np.abs(oa - roa).max() < 1e-3


@needs_ref
def test_bpm_oracle_matches_reference_rotated_surfaces():
    c = acum.Run.synth(acum.synth_params(3, nSect=8, nTime=16, nObsAz=3, nObsEl=2, trip=0)).bpm_case()
    rot = np.ctypeslib.as_array(c.rotate, shape=(3 * c.nSurf_BB,))
    rot[0::3] = 5.0
    rot[1::3] = -3.0
    tra = np.ctypeslib.as_array(c.translate, shape=(3 * c.nSurf_BB,))
    tra[2::3] = 0.3
    spl, oa = O.bpm_f64(c)
    rspl, roa = O.ref_bpm(c)
    assert np.abs(spl - rspl).max() < 1e-3
    assert np.abs(oa - roa).max() < 1e-3


def test_bpm_oracle_level_identities():
    """Energy bookkeeping of the aggregation (bb_core.cu:565-583): the total of each band is the
    energy sum of the six components over the time-average, and OASPL is the energy sum of the
    band totals; cells with no energy sit at the 1e-10 floor."""
    c = acum.Run.synth(acum.synth_params(3, nSect=6, nTime=12, nObsAz=2, nObsEl=2, trip=0)).bpm_case()
    spl, oa = O.bpm_f64(c)
    e_tot = 10 ** (spl[:, :, 0] / 10)
    oa_expect = 10 * np.log10(1e-10 + e_tot.sum(axis=1))
    np.testing.assert_allclose(oa, oa_expect, atol=1e-9)
    assert np.all(spl[:, :, 0] >= spl[:, :, 1:].max(axis=2) - 1e-9)
