"""Placement of the eval kernel's tent contributions, checked with impulses, and the error codes of
invalid evaluations, both against the oracle, through the C-ABI.

Impulse cases: a rotating blade whose normals are normal to its velocity (so the thickness term
and the L.M term vanish exactly), with zero pressure everywhere except one (element, step).  The
loading term is then non-zero for that one evaluation only, so the FP64 history of an observer is
non-zero exactly on that evaluation's tent (core.cu:166-231): the rising bins ic, ic-1, ...
(integers in (tau_prev, tau]) and the falling bins ic+1, ... (integers in (tau, tau_next)), wrapped
at the period.  The bins are the oracle's (orc_tonal_bins); the test covers still air and a
moving medium, steps 0 and nT-1 (where the tent wraps around the period) and steps that span
three bins (a fast source receding against a moving medium: sides longer than the kernel's
four-slot footprint, which go through its tail).
"""
import math

import numpy as np
import pytest

import acum3_periodic_amd as acum
import oracle as O
from tonal_cases import Surface, make_case

pytestmark = pytest.mark.gpu

# The FP64 path sums in FP64 and rounds the sums to the float outputs (pT/pL/pA are float, as
# the reference's): each output must be the oracle's FP64 value rounded to float, to within one
# float ulp (the two FP64 sums differ far below float rounding, so they round alike except where
# they straddle a rounding boundary).
ULPS = 1


@pytest.fixture(scope="module")
def ctx():
    if acum.device_count() < 1:
        pytest.fail("no GPU visible")
    c = acum.Context(0)
    yield c
    c.close()


def impulse_case(nT, k0, j0, moving, E=5, omega=200.0, phase=0.3, vinf=None, precision=acum.FP64,
                 obs=None, P0=2.0e3):
    """E points of one blade rotating about z at omega (radius 0.35 .. 0.95 m, one revolution over
    nT steps), normal (0, 0, 1), velocity in the xy plane; P = P0 at (j0, k0) only."""
    dPsi = 2.0 * math.pi / nT
    dTau = dPsi / omega
    radii = np.linspace(0.35, 0.95, E)
    X = np.zeros((E, nT, 4))
    N = np.zeros((E, nT, 6))
    V = np.zeros((E, nT, 6))
    PR = np.zeros((E, nT, 2))
    for j in range(E):
        for k in range(nT):
            psi = phase + 0.17 * j + omega * dTau * k
            cs, sn = math.cos(psi), math.sin(psi)
            x, y = radii[j] * cs, radii[j] * sn
            X[j, k] = (x, y, 0.02 * j, 1.5e-3 * (1 + 0.2 * j))
            V[j, k, :3] = (-omega * y, omega * x, 0.0)
            V[j, k, 3:] = (-omega * omega * x, -omega * omega * y, 0.0)
            N[j, k, :3] = (0.0, 0.0, 1.0)
    PR[j0, k0, 0] = P0
    if obs is None:
        obs = np.array([[3.1, 1.7, -2.3], [-4.2, 0.6, 1.9], [0.4, -5.3, -3.7]], np.float32)
    if vinf is None:
        vinf = (25.0, -4.0, -3.0) if moving else (0.0, 0.0, 0.0)
    return make_case([Surface(X, N, V, PR)], obs, nT, dPsi, dTau, oM=omega, vinf=vinf,
                     precision=precision)


def long_tent_case(nT, k0, moving, precision=acum.FP64):
    """Observer time advancing by more than two bins per step, which a subsonic rotor never does
    (dtau/dk = 1 - Mr stays below 2 where it stays above 0): E points on a sawtooth path along x
    that recede from the observers (all at x < 0) by 0.36 m per step for 12 steps, 1.6 to 1.9 bins
    of delay per step, and come back slowly over the other steps.  Their declared velocities are
    zero, so only positions (hence observer times) move; the recession is centred on step k0,
    so its evaluation has tent sides of three bins."""
    omega = 200.0
    dPsi = 2.0 * math.pi / nT
    dTau = dPsi / omega
    E, ramp, d_fast = 5, 12, 0.36
    d_slow = d_fast * ramp / (nT - ramp)
    X = np.zeros((E, nT, 4))
    N = np.zeros((E, nT, 6))
    V = np.zeros((E, nT, 6))
    PR = np.zeros((E, nT, 2))
    for j in range(E):
        for k in range(nT):
            m = (k - (k0 - ramp // 2)) % nT
            x = d_fast * m if m < ramp else d_fast * ramp - d_slow * (m - ramp)
            X[j, k] = (x + 0.05 * j, 0.3 * j - 0.6, 0.02 * j, 1.5e-3 * (1 + 0.2 * j))
            N[j, k, :3] = (0.0, 0.0, 1.0)
    PR[4, k0, 0] = 2.0e3
    obs = np.array([[-8.0, 0.5, -1.0], [-6.5, 1.5, 1.2], [-9.0, -0.8, 0.4]], np.float32)
    vinf = (40.0, 5.0, 0.0) if moving else (0.0, 0.0, 0.0)  # sound to the observers runs upstream
    return make_case([Surface(X, N, V, PR)], obs, nT, dPsi, dTau, oM=omega, vinf=vinf,
                     precision=precision)


def tent_bins(case, iobs, e, k):
    """Physical bins of the tent of evaluation (element e, step k) for observer iobs (oracle)."""
    ipts, npi, nd, nu = O.tonal_bins(case, iobs)
    i = e * case.nTime + k
    ic = int(npi[i]) * case.nTime + int(ipts[i])
    bins = [ic - j for j in range(int(nd[i]))] + [ic + 1 + j for j in range(int(nu[i]))]
    return sorted({b % case.nTime for b in bins}), int(nd[i]), int(nu[i])


IMPULSES = [  # (name, nT, step, kind): "still", "moving", or "long"/"long_still" (long_tent_case)
    ("still_mid", 48, 20, "still"),
    ("still_step0", 48, 0, "still"),
    ("still_last_step", 48, 47, "still"),
    ("moving_mid", 48, 31, "moving"),
    ("moving_step0", 48, 0, "moving"),
    ("moving_last_step", 48, 47, "moving"),
    ("long_tent_still_mid", 48, 17, "long_still"),
    ("long_tent_step0", 48, 0, "long"),
    ("long_tent_last_step", 48, 47, "long"),
    ("long_tent_mid", 48, 30, "long"),
]


@pytest.mark.parametrize("name,nT,k0,kind", IMPULSES)
def test_impulse_lands_on_its_tent(ctx, name, nT, k0, kind):
    if kind.startswith("long"):
        j0, c = 4, long_tent_case(nT, k0, kind == "long")
    else:
        j0, c = 2, impulse_case(nT, k0, 2, kind == "moving")
    g = ctx.tonal(c)
    o = O.tonal_f64(c)
    long_side = False
    for iobs in range(c.nObs):
        bins, nd, nu = tent_bins(c, iobs, j0, k0)
        long_side |= max(nd, nu) > 2
        for gx, ox, what in ((g[1], o[1], "pL"), (g[2], o[2], "pA")):
            got = np.nonzero(gx[:, iobs])[0].tolist()
            want = np.nonzero(ox[:, iobs])[0].tolist()
            assert want == bins, f"{name}: oracle {what} bins {want} != tent {bins}"
            assert got == bins, f"{name}: GPU {what} bins {got} != tent {bins} (observer {iobs})"
            want32 = ox[:, iobs].astype(np.float32)
            ulps = np.abs(gx[:, iobs] - want32) / np.spacing(np.abs(want32))
            assert ulps.max() <= ULPS, f"{name}: {what} differs by {ulps.max():.1f} float ulps"
        assert not np.any(g[0][:, iobs]), f"{name}: thickness must vanish"
    assert long_side == kind.startswith("long"), f"{name}: tent sides {nd}, {nu}"


# ---- error codes: the library and the oracle agree ----------------------------------------
def _code(fn, *a):
    try:
        fn(*a)
        return acum.ACUM_OK
    except acum.AcumError as e:
        return e.code


def _oracle_code(c):
    import ctypes as C
    n = c.nTime * c.nObs
    out = [np.empty(n, np.float64) for _ in range(3)]
    return O.orc.orc_tonal_f64(C.byref(c), *[O.p64(a) for a in out], 0)


@pytest.mark.parametrize("moving", [False, True])
@pytest.mark.parametrize("prec", [acum.FP64, acum.MIXED])
def test_source_on_observer_is_an_error(ctx, moving, prec):
    c0 = impulse_case(36, 5, 2, moving)
    X = np.ctypeslib.as_array(c0.X, shape=(c0.nTime * 5 * 4,)).reshape(5, c0.nTime, 4)
    obs = np.array([[3.1, 1.7, -2.3], X[3, 11, :3]], np.float32)  # observer 1 sits on element 3 at step 11
    c = impulse_case(36, 5, 2, moving, obs=obs, precision=prec)
    assert _oracle_code(c) == acum.ACUM_ERR_NUMERIC
    assert _code(ctx.tonal, c) == acum.ACUM_ERR_NUMERIC
    assert _code(ctx.tonal, impulse_case(36, 5, 2, moving, precision=prec)) == acum.ACUM_OK


@pytest.mark.parametrize("prec", [acum.FP64, acum.MIXED])
def test_non_finite_observer_is_rejected(ctx, prec):
    for bad in (float("nan"), float("inf")):
        obs = np.array([[3.1, 1.7, -2.3], [bad, 0.0, 1.0]], np.float32)
        c = impulse_case(36, 5, 2, True, obs=obs, precision=prec)
        assert _oracle_code(c) == acum.ACUM_ERR_ARG
        assert _code(ctx.tonal, c) == acum.ACUM_ERR_ARG


@pytest.mark.parametrize("prec", [acum.FP64, acum.MIXED])
def test_supersonic_tip_is_an_error(ctx, prec):
    """A blade tip at Mach 1.4: observers in the rotor plane see Mr > 1, where 1 - Mr changes sign
    and observer time runs backwards; both the library and the oracle report it."""
    obs = np.array([[6.0, 0.0, 0.1], [0.0, -7.0, 0.2]], np.float32)
    c = impulse_case(64, 5, 2, False, omega=500.0, obs=obs, precision=prec)
    assert _oracle_code(c) == acum.ACUM_ERR_NUMERIC
    assert _code(ctx.tonal, c) == acum.ACUM_ERR_NUMERIC
