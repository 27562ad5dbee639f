"""GPU parity of the tonal Farassat 1A path against the FP64 CPU oracle, through the C-ABI.

Bars (from the path's contract): observer-time bin indices bit-exact; FP64 pressure histories
within 1e-6 relative L2; the mixed FP32/FP64 path within 0.01 dB OASPL of the FP64 oracle.
"""
import numpy as np
import pytest

import acum3_periodic_amd as acum
import oracle as O
from parity import report
from tonal_cases import make_case, rotor_surface

pytestmark = pytest.mark.gpu

REL_L2_FP64 = 1e-6       # FP64 GPU vs FP64 oracle, each of pT, pL, pA
OASPL_DB = 0.01          # mixed GPU vs FP64 oracle
REL_L2_MIXED = 1e-5      # mixed GPU vs FP64 oracle on pA (FP32 integrands, FP32 tile partials)


def small_case(config=1, precision=acum.FP64, **over):
    base = dict(nChord=4, nSpan=3, nTime=36, nObs=8)
    if config == 2:
        base = dict(nChord=5, nSpan=4, nTime=64, nObsAz=6, nObsEl=3)
    base.update(over)
    run = acum.Run.synth(acum.synth_params(config, **base)).preprocess()
    return run, run.tonal_case(precision)


@pytest.fixture(scope="module")
def ctx():
    if acum.device_count() < 1:
        pytest.fail("no GPU visible")
    c = acum.Context(0)
    yield c
    c.close()


CASES = [
    ("hover", 1, {}),
    ("hover_odd_obs", 1, dict(nObs=70, nTime=40)),
    ("flight", 2, {}),
    ("flight_coarse_time", 2, dict(nTime=24)),
    ("hover_period_below_window", 1, dict(nTime=16)),  # 16 steps < the 20-bin window: bins repeat
]


@pytest.mark.parametrize("name,config,over", CASES)
def test_bins_bit_exact(ctx, name, config, over):
    run, c = small_case(config, acum.FP64, **over)
    for iobs in (0, c.nObs - 1):
        g = ctx.bins(c, iobs)
        o = O.tonal_bins(c, iobs)
        for a, b, what in zip(g, o, ("ipts", "npi", "ndown", "nup")):
            assert np.array_equal(a, b), f"{name}: {what} differs at observer {iobs}"
    # the mixed path uses the same FP64 observer time from the same float positions
    cm = run.tonal_case(acum.MIXED)
    g = ctx.bins(cm, 0)
    o = O.tonal_bins(cm, 0)
    assert all(np.array_equal(a, b) for a, b in zip(g, o))


@pytest.mark.parametrize("name,config,over", CASES)
def test_fp64_matches_oracle(ctx, name, config, over):
    run, c = small_case(config, acum.FP64, **over)
    gT, gL, gA = ctx.tonal(c)
    oT, oL, oA = O.tonal_f64(c)
    errs = [O.rel_l2(g, o) for g, o in ((gT, oT), (gL, oL), (gA, oA))]
    report(f"fp64_{name}", rel_l2_pT=errs[0], rel_l2_pL=errs[1], rel_l2_pA=errs[2])
    for err, what in zip(errs, ("pT", "pL", "pA")):
        assert err <= REL_L2_FP64, f"{name} {what}: rel L2 {err:.3e}"


@pytest.mark.parametrize("name,config,over", CASES)
def test_mixed_matches_oracle(ctx, name, config, over):
    run, c = small_case(config, acum.MIXED, **over)
    gT, gL, gA = ctx.tonal(c)
    oT, oL, oA = O.tonal_f64(c)
    err = O.rel_l2(gA, oA)
    d = np.abs(acum.oaspl(gA) - acum.oaspl(oA.astype(np.float32)))
    report(f"mixed_{name}", rel_l2_pA=err, rel_l2_pT=O.rel_l2(gT, oT), rel_l2_pL=O.rel_l2(gL, oL),
           max_doaspl_db=float(d.max()))
    assert err <= REL_L2_MIXED, f"{name}: pA rel L2 {err:.3e}"
    assert d.max() <= OASPL_DB, f"{name}: OASPL differs by {d.max():.4f} dB"


def test_deterministic_and_graph(ctx):
    run, c = small_case(2, acum.MIXED)
    h = ctx.stage([c])
    ctx.run(h, acum.MIXED)
    a = ctx.fetch(h, 0, c.nTime, c.nObs)
    ctx.run(h, acum.MIXED)
    b = ctx.fetch(h, 0, c.nTime, c.nObs)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    for rep in range(3):  # replays of one captured graph, with eager work in between
        ctx.run(h, acum.MIXED, repeats=2, graph=True)
        g = ctx.fetch(h, 0, c.nTime, c.nObs)
        for x, y in zip(a, g):
            assert np.array_equal(x, y), f"graph replay {rep} differs"
        ctx.run(h, acum.FP64)
    ctx.release(h)


def test_batch_equals_single(ctx):
    runs = [small_case(1, acum.MIXED, nTime=32 + 4 * i, nObs=5 + 30 * i, nSpan=2 + i) for i in range(3)]
    cases = [c for _, c in runs]
    batch = ctx.tonal_batch(cases)
    for c, b in zip(cases, batch):
        s = ctx.tonal(c)
        for x, y in zip(s, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("prec,window,spread,tile", [
    (acum.MIXED, 16, 2, 32), (acum.MIXED, 24, 4, 256), (acum.MIXED, 48, 6, 64),
    (acum.MIXED, 64, 8, 256), (acum.FP64, 16, 1, 16), (acum.FP64, 24, 3, 256),
    (acum.FP64, 32, 6, 128)])
def test_tile_shapes_agree(ctx, prec, window, spread, tile):
    """Window size, tile spread and tile length change only the partial-sum grouping."""
    run, c = small_case(2, prec)
    ref = ctx.tonal(c)[2]
    keep = ctx.get_tuning(prec)
    ctx.set_tuning(prec, window=window, max_spread=spread, max_tile=tile)
    try:
        got = ctx.tonal(c)[2]
    finally:
        ctx.set_tuning(prec, window=keep.window, max_spread=keep.max_spread, max_tile=keep.max_tile)
    assert O.rel_l2(got, ref) < (2e-6 if prec == acum.MIXED else 1e-12)


def test_batch_mixes_still_and_moving_cases(ctx):
    """A hover case (still air) and a forward-flight case run through the two kernel variants in
    one batch and give the same results as alone."""
    cases = [small_case(1, acum.MIXED)[1], small_case(2, acum.MIXED)[1], small_case(1, acum.MIXED, nTime=20)[1]]
    assert cases[1].vInf[0] != 0.0 and cases[0].vInf[0] == 0.0
    batch = ctx.tonal_batch(cases)
    for c, b in zip(cases, batch):
        s = ctx.tonal(c)
        for x, y in zip(s, b):
            assert np.array_equal(x, y)


def test_errors_are_returned(ctx):
    run, c = small_case(1, acum.FP64)
    c.vInf[0] = 400.0  # supersonic free stream
    with pytest.raises(acum.AcumError) as e:
        ctx.tonal(c)
    assert e.value.code == acum.ACUM_ERR_NUMERIC
    run, c = small_case(1, acum.FP64)
    c.precision = 7
    with pytest.raises(acum.AcumError) as e:
        ctx.tonal(c)
    assert e.value.code == acum.ACUM_ERR_ARG
    # a non-finite load or position anywhere in the element data is reported, in both precisions
    for field, idx in (("PR", 17), ("X", 4 * 5 + 1)):
        for prec in (acum.FP64, acum.MIXED):
            run, c = small_case(1, prec)
            getattr(c, field)[idx] = float("nan")
            with pytest.raises(acum.AcumError) as e:
                ctx.tonal(c)
            assert e.value.code == acum.ACUM_ERR_NUMERIC
    # empty and degenerate shapes are argument errors, caught before any launch
    for field, value in (("nObs", 0), ("nTime", 3), ("nSurf", 0)):
        run, c = small_case(1, acum.MIXED)
        setattr(c, field, value)
        with pytest.raises(acum.AcumError) as e:
            ctx.tonal(c)
        assert e.value.code == acum.ACUM_ERR_ARG, field
    run, c = small_case(1, acum.MIXED)
    c.E[0] = 0  # a surface with no elements
    with pytest.raises(acum.AcumError) as e:
        ctx.tonal(c)
    assert e.value.code == acum.ACUM_ERR_ARG
    run, c = small_case(1, acum.FP64)
    ctx.tonal(c)  # the context keeps working after those errors


def test_c1_full_size(ctx):
    """The reference's CPU-runnable configuration at its full size (2 surfaces x 1000 elements,
    360 steps, 36 observers on an arc, hover): both precisions against the FP64 oracle, and the
    bin indices of every source for the first and last observer."""
    run = acum.Run.synth(acum.synth_params(1)).preprocess()
    for prec, bar in ((acum.FP64, REL_L2_FP64), (acum.MIXED, REL_L2_MIXED)):
        c = run.tonal_case(prec)
        assert c.nTime == 360 and c.nObs == 36 and c.E[0] == 1000
        g = ctx.tonal(c)
        o = O.tonal_f64(c)
        for x, y in zip(g, o):
            assert O.rel_l2(x, y) <= bar
        d = np.abs(acum.oaspl(g[2]) - acum.oaspl(o[2].astype(np.float32)))
        assert d.max() <= OASPL_DB
    c = run.tonal_case(acum.FP64)
    for iobs in (0, c.nObs - 1):
        assert all(np.array_equal(a, b) for a, b in zip(ctx.bins(c, iobs), O.tonal_bins(c, iobs)))


@pytest.mark.parametrize("prec", [acum.MIXED, acum.FP64])
def test_window_overflow_widens_the_motion_bound(ctx, prec):
    """Planned with a far too small bound on observer-time change per step, the passes overflow
    their windows; the kernel's window check catches it and the call plans the case again with a
    wider bound (0.6 of the worst case, then 0.8, then the full one) and runs it again.  The
    results are bit-identical to planning with that bound from the start (a plan that kept the
    small bound would group the sums differently) and agree with the oracle."""
    keep = ctx.get_tuning(prec)
    try:
        run, c = small_case(2, prec)
        refs = []
        for pm in (600, 800, 1000):
            ctx.set_tuning(prec, motion_permille=pm)
            refs.append(ctx.tonal(c))
        ctx.set_tuning(prec, motion_permille=100)
        got = ctx.tonal(c)
        assert ctx.timing().motion_widened_cases == 1
    finally:
        ctx.set_tuning(prec, window=keep.window, max_spread=keep.max_spread, max_tile=keep.max_tile,
                       motion_permille=keep.motion_permille)
    assert any(all(np.array_equal(x, y) for x, y in zip(ref, got)) for ref in refs)
    o = O.tonal_f64(c)
    assert O.rel_l2(got[2], o[2]) <= (REL_L2_MIXED if prec == acum.MIXED else REL_L2_FP64)


def test_autotune_picks_a_candidate_and_leaves_the_stage_consistent(ctx):
    """acum_autotune times the candidate tile shapes on a staged case and keeps the fastest; a
    later run of that stage equals a fresh context's run with the chosen shape set explicitly."""
    ctx.tune_cache_clear()
    try:
        run, c = small_case(2, acum.MIXED)
        h = ctx.stage([c])
        best, ms = ctx.autotune(h, acum.MIXED)
        assert best.window in (16, 20, 24) and best.max_spread in (2, 3, 4) and ms > 0
        assert best.max_tile in (64, 256) and best.motion_permille in (500, 600, 800, 1000)
        ctx.run(h, acum.MIXED)
        tuned = ctx.fetch(h, 0, c.nTime, c.nObs)
        ctx.release(h)
        other = acum.Context(0)
        other.set_tuning(acum.MIXED, window=best.window, max_spread=best.max_spread,
                         max_tile=best.max_tile, motion_permille=best.motion_permille)
        again = other.tonal(c)
        other.close()
    finally:
        ctx.tune_cache_clear()
    for x, y in zip(tuned, again):
        assert np.array_equal(x, y)


def test_autotune_cache_is_reused_and_persists(ctx, tmp_path):
    """A tuned shape is cached by bucket: a later autotune of a stage whose cases fall in tuned
    buckets times nothing and plans them with the cached shapes; saved to a file and loaded into
    a fresh context, the cache gives the same shapes and bit-identical results without tuning."""
    ctx.tune_cache_clear()
    try:
        run, c = small_case(2, acum.MIXED)
        run2, c2 = small_case(2, acum.MIXED, seed=31077)  # same sizes, other data: same bucket
        h = ctx.stage([c])
        best, _ = ctx.autotune(h, acum.MIXED)
        t = ctx.timing()
        assert (t.tune_cases, t.tune_hits) == (1, 0) and t.tune_passes > 0
        ctx.run(h, acum.MIXED)
        first = ctx.fetch(h, 0, c.nTime, c.nObs)
        ctx.release(h)
        h2 = ctx.stage([c2, c])
        ctx.autotune(h2, acum.MIXED)
        t = ctx.timing()
        assert (t.tune_cases, t.tune_hits, t.tune_passes) == (0, 2, 0)
        for i in range(2):
            got = ctx.case_tuning(h2, i, acum.MIXED)
            assert (got.window, got.max_spread, got.max_tile, got.motion_permille) == \
                (best.window, best.max_spread, best.max_tile, best.motion_permille)
        ctx.release(h2)
        path = tmp_path / "tune.txt"
        ctx.tune_cache_save(path)
        fresh = acum.Context(0)
        try:
            fresh.tune_cache_load(path)
            h3 = fresh.stage([c])
            fresh.autotune(h3, acum.MIXED)
            t = fresh.timing()
            assert (t.tune_cases, t.tune_hits, t.tune_passes) == (0, 1, 0)
            fresh.run(h3, acum.MIXED)
            again = fresh.fetch(h3, 0, c.nTime, c.nObs)
            fresh.release(h3)
        finally:
            fresh.close()
        assert all(np.array_equal(x, y) for x, y in zip(first, again))
    finally:
        ctx.tune_cache_clear()


def test_batch_of_cases_with_different_windows(ctx, tmp_path):
    """Cases whose buckets carry different windows run in one pass as several eval launches (one
    per kernel variant) and give the same histories as each case alone."""
    cases = [small_case(2, acum.MIXED)[1], small_case(2, acum.MIXED, nTime=128, nObsAz=12)[1],
             small_case(1, acum.MIXED)[1]]
    ctx.tune_cache_clear()
    try:
        # buckets of the three cases from a stage's own view of them, shapes written as a file
        h = ctx.stage(cases)
        ctx.autotune(h, acum.MIXED)
        ctx.tune_cache_save(tmp_path / "tuned.txt")
        ctx.release(h)
        lines = [l.split() for l in open(tmp_path / "tuned.txt") if not l.startswith("#")]
        assert len(lines) == 3
        for l, w in zip(lines, (16, 24, 20)):  # force three different windows
            l[6] = str(w)
        with open(tmp_path / "forced.txt", "w") as f:
            f.write("\n".join(" ".join(l) for l in lines) + "\n")
        ctx.tune_cache_clear()
        ctx.tune_cache_load(tmp_path / "forced.txt")
        windows = set()
        h = ctx.stage(cases)
        ctx.run(h, acum.MIXED)
        assert ctx.timing().eval_launches == 3
        for i in range(3):
            windows.add(ctx.case_tuning(h, i, acum.MIXED).window)
        batch = [ctx.fetch(h, i, c.nTime, c.nObs) for i, c in enumerate(cases)]
        ctx.release(h)
        assert windows == {16, 20, 24}
        for c, b in zip(cases, batch):
            assert all(np.array_equal(x, y) for x, y in zip(ctx.tonal(c), b))
    finally:
        ctx.tune_cache_clear()


@pytest.mark.parametrize("prec", [acum.FP64, acum.MIXED])
@pytest.mark.parametrize("E,nT,nobs,moving", [(1, 4, 1, False), (1, 4, 1, True), (3, 5, 65, True),
                                              (2, 7, 63, False)])
def test_smallest_and_ragged_shapes(ctx, prec, E, nT, nobs, moving):
    """The smallest case the library accepts (one element, four steps, one observer) and ragged
    ones (65 and 63 observers: a wave with one or 63 live lanes; odd step counts, periods shorter
    than the window) against the FP64 oracle, in both precisions."""
    omega = 150.0
    dPsi = 2.0 * np.pi / nT
    dTau = dPsi / omega
    rng = np.random.default_rng(100 * E + nT)
    P = rng.uniform(500.0, 2000.0, size=(E, nT))
    surf = rotor_surface(nT, np.linspace(0.4, 0.9, E), np.zeros(E), omega, dTau, P)
    i = np.arange(nobs) + 0.5
    th, ph = np.arccos(1.0 - 2.0 * i / nobs), np.pi * (1.0 + 5.0 ** 0.5) * i
    rad = 4.0 + 4.0 * (i / nobs)
    obs = np.stack([rad * np.sin(th) * np.cos(ph), rad * np.sin(th) * np.sin(ph), rad * np.cos(th)], 1)
    vinf = (20.0, -3.0, -2.0) if moving else (0.0, 0.0, 0.0)
    c = make_case([surf], obs.astype(np.float32), nT, dPsi, dTau, oM=omega, vinf=vinf, precision=prec)
    g = ctx.tonal(c)
    o = O.tonal_f64(c)
    chans = ("pT", "pL", "pA") if prec == acum.FP64 else ("pA",)
    bar = REL_L2_FP64 if prec == acum.FP64 else REL_L2_MIXED
    for ch in chans:
        k = ("pT", "pL", "pA").index(ch)
        err = O.rel_l2(g[k], o[k])
        assert err <= bar, f"E={E} nT={nT} nObs={nobs} {ch}: rel L2 {err:.3e}"
