"""GPU parity of every benchmark configuration at its real shape, against the FP64 CPU oracle,
through the C-ABI.  The GPU runs each case whole; the oracle runs a few observers of it (observers
are independent, so those columns of the GPU histories are exactly comparable).

  * four-blade rotor in forward flight (4 x 20,000 elements, 1024 steps, 1024 observers):
    FP64 path within 1e-6 rel-L2 per channel, mixed path within the mixed bar, OASPL and every
    1/3-octave band within 0.01 dB, bin indices bit-exact for two observers;
  * broadband (4 x 60 sections, 360 steps, 1024 observers, 34 bands), tripped and untripped:
    every component and band within the broadband bar on 8 observers;
  * batched: the first four seeded cases of the uneven 64-case workload at their drawn sizes,
    tile shapes from the autotuner, one persistent launch replayed as a HIP graph: each case
    bit-identical to it run alone and within the mixed bars of the oracle on 8 observers;
  * stress (coaxial 2 x 4 blades x 50,000 elements, 2048 steps, 4096 observers, impulsive
    loading): mixed path within 1e-5 rel-L2 and 0.01 dB (OASPL and bands) on 16 observers.
"""
import numpy as np
import pytest

import acum3_periodic_amd as acum
import oracle as O
from parity import ISO_BANDS, band_error, band_levels, obs_subset, report, spread_observers

pytestmark = pytest.mark.gpu

REL_L2_FP64 = 1e-6    # FP64 GPU vs FP64 oracle, each of pT, pL, pA
REL_L2_MIXED = 1e-5   # mixed GPU vs FP64 oracle, pA
LEVEL_DB = 0.01       # OASPL and 1/3-octave band levels of the mixed path vs the FP64 oracle
BPM_DB = 1e-3         # broadband band levels and OASPL vs the FP64 oracle


@pytest.fixture(scope="module")
def ctx():
    if acum.device_count() < 1:
        pytest.fail("no GPU visible")
    c = acum.Context(0)
    yield c
    c.close()


def check_mixed(name, g, o, dt, gpu_spl=None):
    """Mixed-path histories g (the observer columns of the oracle's subset) against the FP64
    oracle's o: rel-L2 of pA, OASPL and band levels.  gpu_spl: the GPU spectra of the same
    columns, compared as well when given."""
    rel = O.rel_l2(g[2], o[2])
    rel_T = O.rel_l2(g[0], o[0])
    rel_L = O.rel_l2(g[1], o[1])
    d_oa = float(np.abs(acum.oaspl(g[2]) - acum.oaspl(o[2].astype(np.float32))).max())
    ref_spl, _ = band_levels(o[2], dt)
    spl, _ = band_levels(g[2], dt)
    d_band, where = band_error(spl, ref_spl)
    vals = dict(rel_l2_pA=rel, rel_l2_pT=rel_T, rel_l2_pL=rel_L, max_doaspl_db=d_oa,
                max_dband_db=d_band, worst_band=None if where is None else float(ISO_BANDS[where[1]]),
                live_bands=int((ref_spl > -99.0).sum()))
    if gpu_spl is not None:
        d_gpu, _ = band_error(gpu_spl, ref_spl)
        vals["max_dband_db_gpu_spectra"] = d_gpu
    report(name, **vals)
    assert rel <= REL_L2_MIXED, f"{name}: pA rel L2 {rel:.3e}"
    assert d_oa <= LEVEL_DB, f"{name}: OASPL differs by {d_oa:.3e} dB"
    assert d_band <= LEVEL_DB, f"{name}: band {where} differs by {d_band:.3e} dB"
    if gpu_spl is not None:
        assert vals["max_dband_db_gpu_spectra"] <= LEVEL_DB


def test_rotor_full_size(ctx):
    run = acum.Run.synth(acum.synth_params(2)).preprocess()
    c = run.tonal_case(acum.MIXED)
    assert (c.nSurf, c.nTime, c.nObs, c.E[0]) == (4, 1024, 1024, 20000)
    h = ctx.stage([c])
    try:
        ctx.run(h, acum.MIXED)
        mixed = ctx.fetch(h, 0, c.nTime, c.nObs)
        gspl, _ = ctx.fetch_spectra(h, 0, c.nObs, ISO_BANDS)
        ctx.run(h, acum.FP64)
        plain = ctx.fetch(h, 0, c.nTime, c.nObs)
    finally:
        ctx.release(h)
    idx = spread_observers(c.nObs, 16)
    o = O.tonal_f64(obs_subset(c, idx, acum.FP64))
    errs = [O.rel_l2(g[:, idx], r) for g, r in zip(plain, o)]
    report("rotor_fp64", observers=len(idx), rel_l2_pT=errs[0], rel_l2_pL=errs[1], rel_l2_pA=errs[2])
    for e, what in zip(errs, ("pT", "pL", "pA")):
        assert e <= REL_L2_FP64, f"FP64 {what}: rel L2 {e:.3e}"
    dt = c.dPsi / c.oM
    check_mixed("rotor_mixed", [m[:, idx] for m in mixed], o, dt, gspl[idx])
    # bin indices of every source for two observers
    cf = run.tonal_case(acum.FP64)
    for iobs in (0, c.nObs - 1):
        for a, b, what in zip(ctx.bins(cf, iobs), O.tonal_bins(cf, iobs), ("ipts", "npi", "ndown", "nup")):
            assert np.array_equal(a, b), f"{what} differs at observer {iobs}"


@pytest.mark.parametrize("trip", [1, 0])
def test_broadband_full_size(ctx, trip):
    run = acum.Run.synth(acum.synth_params(3, trip=trip))
    c = run.bpm_case()
    nsect = np.ctypeslib.as_array(c.nsect, shape=(c.nSurf_BB,))
    assert (c.nSurf_BB, c.nTime, c.nObs, c.nf2) == (4, 360, 1024, 34) and (nsect == 60).all()
    spl, oa = ctx.bpm(c)
    worst = worst_oa = 0.0
    for obs0 in (0, c.nObs - 4):  # 8 observers: the first four and the last four
        ospl, ooa = O.bpm_f64(c, obs0, 4)
        worst = max(worst, float(np.abs(spl[obs0:obs0 + 4].astype(np.float64) - ospl).max()))
        worst_oa = max(worst_oa, float(np.abs(oa[obs0:obs0 + 4] - ooa).max()))
    report(f"broadband_trip{trip}", observers=8, max_dspl_db=worst, max_doaspl_db=worst_oa)
    assert worst <= BPM_DB and worst_oa <= BPM_DB


def test_batched_cases_at_drawn_sizes(ctx):
    seed, n = 31004, 4
    runs = [acum.Run.synth(acum.batch_params(seed, i)).preprocess() for i in range(n)]
    cases = [r.tonal_case(acum.MIXED) for r in runs]
    h = ctx.stage(cases)
    try:
        ctx.autotune(h, acum.MIXED)
        ctx.run(h, acum.MIXED, repeats=1, graph=True)  # plan, capture, first replay
        ctx.run(h, acum.MIXED, repeats=2, graph=True)
        batch = [ctx.fetch(h, i, c.nTime, c.nObs) for i, c in enumerate(cases)]
    finally:
        ctx.release(h)
    for i, (c, b) in enumerate(zip(cases, batch)):
        alone = ctx.tonal(c)
        assert all(np.array_equal(x, y) for x, y in zip(alone, b)), f"case {i}: batch differs from alone"
        idx = np.arange(min(8, c.nObs))
        o = O.tonal_f64(obs_subset(c, idx, acum.FP64))
        check_mixed(f"batched_case{i}", [x[:, idx] for x in b], o, c.dPsi / c.oM)


def test_stress_full_size(ctx):
    run = acum.Run.synth(acum.synth_params(5))
    run.preprocess_gpu(ctx)
    c = run.tonal_case(acum.MIXED)
    nE = sum(c.E[s] for s in range(c.nSurf))
    assert (c.nSurf, nE, c.nTime, c.nObs) == (8, 400000, 2048, 4096)
    h = ctx.stage([c])
    try:
        ctx.run(h, acum.MIXED)
        g = ctx.fetch(h, 0, c.nTime, c.nObs)
        gspl, _ = ctx.fetch_spectra(h, 0, c.nObs, ISO_BANDS)
    finally:
        ctx.release(h)
    idx = spread_observers(c.nObs, 16)
    o = O.tonal_f64(obs_subset(c, idx, acum.FP64))
    check_mixed("stress_mixed", [x[:, idx] for x in g], o, c.dPsi / c.oM, gspl[idx])
