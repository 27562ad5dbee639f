"""The plain FP64 path (acum_tonal_plain: the reference's kernel structure in FP64, the baseline
the restructured paths are measured against) against the FP64 CPU oracle, through the C-ABI.

Its bins come from a floor of ((k dTau + r / a0) oM) / dPsi and its sums from FP64 atomics in
whatever order the hardware serves them, so it equals the oracle to FP64 rounding, not bit for
bit.  Its outputs are float (as the reference's), whose rounding alone is about 3e-8 relative L2
against the FP64 oracle: each output within 1e-7 relative L2 (the path's bar is 1e-6), and a
tent's bins
exactly where the oracle puts them (impulse cases, steps 0, mid and nT-1, still air and moving
medium, sides longer than two bins).
"""
import numpy as np
import pytest

import acum3_periodic_amd as acum
import oracle as O
from parity import report
from test_gpu_placement import impulse_case, long_tent_case, tent_bins, _code
from test_gpu_tonal import CASES, small_case

pytestmark = pytest.mark.gpu

REL_L2_PLAIN = 1e-7


@pytest.fixture(scope="module")
def ctx():
    if acum.device_count() < 1:
        pytest.fail("no GPU visible")
    c = acum.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name,config,over", CASES)
def test_plain_matches_oracle(ctx, name, config, over):
    run, c = small_case(config, acum.FP64, **over)
    g = ctx.tonal_plain(c)
    t = ctx.timing()
    assert t.evals == sum(c.E[s] for s in range(c.nSurf)) * c.nTime * c.nObs
    o = O.tonal_f64(c)
    errs = [O.rel_l2(a, b) for a, b in zip(g, o)]
    report(f"plain_{name}", rel_l2_pT=errs[0], rel_l2_pL=errs[1], rel_l2_pA=errs[2])
    for err, what in zip(errs, ("pT", "pL", "pA")):
        assert err <= REL_L2_PLAIN, f"{name} {what}: rel L2 {err:.3e}"


@pytest.mark.parametrize("nT,k0,kind", [(48, 0, "still"), (48, 31, "moving"), (48, 47, "moving"),
                                        (48, 0, "long"), (48, 30, "long")])
def test_plain_impulse_lands_on_its_tent(ctx, nT, k0, kind):
    if kind == "long":
        j0, c = 4, long_tent_case(nT, k0, True)
    else:
        j0, c = 2, impulse_case(nT, k0, 2, kind == "moving")
    g = ctx.tonal_plain(c)
    o = O.tonal_f64(c)
    for iobs in range(c.nObs):
        bins, _, _ = tent_bins(c, iobs, j0, k0)
        got = np.nonzero(g[2][:, iobs])[0].tolist()
        assert got == bins, f"{kind} step {k0}: plain bins {got} != tent {bins} (observer {iobs})"
        want = o[2][:, iobs]
        assert np.abs(g[2][:, iobs] - want).max() <= 1e-6 * np.abs(want).max()


def test_plain_errors(ctx):
    # a supersonic tip (1 - Mr changes sign) and a source on an observer are ACUM_ERR_NUMERIC
    obs = np.array([[6.0, 0.0, 0.1], [0.0, -7.0, 0.2]], np.float32)
    assert _code(ctx.tonal_plain, impulse_case(64, 5, 2, False, omega=500.0, obs=obs)) == acum.ACUM_ERR_NUMERIC
    c0 = impulse_case(36, 5, 2, True)
    X = np.ctypeslib.as_array(c0.X, shape=(c0.nTime * 5 * 4,)).reshape(5, c0.nTime, 4)
    on = np.array([[3.1, 1.7, -2.3], X[3, 11, :3]], np.float32)
    assert _code(ctx.tonal_plain, impulse_case(36, 5, 2, True, obs=on)) == acum.ACUM_ERR_NUMERIC
    bad = np.array([[3.1, 1.7, -2.3], [float("nan"), 0.0, 1.0]], np.float32)
    assert _code(ctx.tonal_plain, impulse_case(36, 5, 2, True, obs=bad)) == acum.ACUM_ERR_ARG
    assert _code(ctx.tonal_plain, impulse_case(36, 5, 2, True)) == acum.ACUM_OK
