"""Pins of the tonal oracle (oracle/orc_tonal.cpp) without a GPU: closed forms, invariants and a
brute-force comparison with a literal restatement of the reference's per-eval steps.

  * stationary compact dipole: the discrete closed form of the binned history (core.cu:68-231);
  * thickness vanishes for a stationary element, loading is linear in P, pA = pT + pL;
  * acous_dist (core.cu:261-279): residual and smallest positive root of the convected quadratic;
  * bins: the bin-unit tent selects the same (ipts, npi, sweep lengths) as findpts/bsearch and the
    radian-domain sweeps (core.cu:156-231), exactly;
  * hover rotor: rotating the observer by dPsi shifts the history by one bin; two identical
    blades 180 degrees apart give the one-blade history plus its half-period shift.
"""
import ctypes as C
import math

import numpy as np
import pytest

import acum3_periodic_amd as acum
import oracle as O
from tonal_cases import Surface, make_case, rotor_surface, literal_reference


def _consts(nT, oM=200.0):
    dPsi = 2.0 * math.pi / nT
    dTau = dPsi / oM   # one revolution: dTau = 2 pi dPsi_deg / oM / 360 (io.cu:74-75)
    return dPsi, dTau


def test_stationary_dipole_closed_form():
    """A fixed element with pressure history P_k: the thickness term is zero and the loading
    term is the compact dipole f_k = ds cos(theta) (Pdot_k/(4 pi a0 r) + P_k/(4 pi r^2)); every
    step lands exactly delta = r oM/(a0 dPsi) bins later, so the tent gives linear interpolation
    out[b] = (1 - phi) f[b - m] + phi f[b - m - 1] with delta = m + phi."""
    nT, a0, oM = 24, 340.0, 200.0
    dPsi, dTau = _consts(nT, oM)
    k = np.arange(nT)
    P = 50.0 * np.sin(2 * np.pi * 3 * k / nT) + 20.0 * np.cos(2 * np.pi * 5 * k / nT) + 7.0
    Pd = 300.0 * np.cos(2 * np.pi * 2 * k / nT) - 40.0
    n = np.array([0.6, 0.0, 0.8], np.float32)
    X = np.zeros((1, nT, 4), np.float32)
    X[0, :, 3] = 2.5e-3
    N = np.zeros((1, nT, 6), np.float32)
    N[0, :, :3] = n
    V = np.zeros((1, nT, 6), np.float32)
    PR = np.stack([P, Pd], axis=-1)[None].astype(np.float32)
    obs = np.array([[3.0, 1.0, 4.0]], np.float32)
    case = make_case([Surface(X, N, V, PR)], obs, nT, dPsi, dTau, a0=a0, oM=oM)
    pT, pL, pA = O.tonal_f64(case)

    o = obs[0].astype(np.float64)
    r = float(np.linalg.norm(o))
    cos_t = float(n.astype(np.float64) @ (o / r))
    ds = float(X[0, 0, 3])
    P32, Pd32 = PR[0, :, 0].astype(np.float64), PR[0, :, 1].astype(np.float64)
    f = ds * cos_t * (Pd32 / (4 * np.pi * a0 * r) + P32 / (4 * np.pi * r * r))
    delta = r * (oM / a0) / dPsi
    m = math.floor(delta)
    phi = delta - m
    b = np.arange(nT)
    expect = (1 - phi) * f[(b - m) % nT] + phi * f[(b - m - 1) % nT]
    assert np.all(pT == 0.0)
    np.testing.assert_allclose(pL[:, 0], expect, rtol=1e-11, atol=1e-14 * np.abs(f).max())
    np.testing.assert_allclose(pA, pT + pL, rtol=1e-14, atol=0)


def _moving_case(rng, nT=12, vinf=(20.0, -5.0, 3.0), nobs=3, nsurf=2, E=3):
    dPsi, dTau = _consts(nT)
    surfs = []
    for s in range(nsurf):
        radii = 0.5 + rng.random(E)
        chord = 0.05 * rng.standard_normal(E)
        P = 100.0 * (1 + 0.3 * rng.standard_normal((E, nT)))
        surfs.append(rotor_surface(nT, radii, chord, 200.0, dTau, P, phase=math.pi * s, z=0.01 * s,
                                   cone=0.05))
    obs = rng.standard_normal((nobs, 3)) * 3.0
    obs[:, 2] -= 4.0
    return make_case(surfs, obs, nT, dPsi, dTau, vinf=vinf)


@pytest.mark.parametrize("seed,vinf", [(1, (0.0, 0.0, 0.0)), (2, (20.0, -5.0, 3.0)), (3, (-40.0, 0.0, -10.0))])
def test_brute_force_matches_literal_reference_steps(seed, vinf):
    """orc_tonal_f64 and orc_tonal_bins against the literal radian-domain algorithm."""
    rng = np.random.default_rng(seed)
    case = _moving_case(rng, vinf=vinf)
    lT, lL, lA, lbins = literal_reference(case)
    oT, oL, oA = O.tonal_f64(case)
    for o, l in ((oT, lT), (oL, lL), (oA, lA)):
        assert O.rel_l2(o, l) <= 1e-11
    ipts, npi, nd, nu = O.tonal_bins(case, 0)
    assert np.array_equal(ipts, lbins[:, 0])
    assert np.array_equal(-npi, lbins[:, 1])  # findpts shifts by -(whole periods)
    assert np.array_equal(nd, lbins[:, 2])
    assert np.array_equal(nu, lbins[:, 3])


def test_loading_linear_thickness_independent_of_pressure():
    rng = np.random.default_rng(7)
    case = _moving_case(rng, vinf=(10.0, 0.0, 0.0))
    t1, l1, a1 = O.tonal_f64(case)
    case._keep["PR"] *= 2.0  # exact in binary floating point
    t2, l2, a2 = O.tonal_f64(case)
    assert np.array_equal(t1, t2)
    assert np.array_equal(l2, 2.0 * l1)
    np.testing.assert_allclose(a1, t1 + l1, rtol=1e-12, atol=1e-12 * np.abs(a1).max())


@pytest.mark.parametrize("vinf", [(0.0, 0.0, 0.0), (30.0, 10.0, -5.0), (-60.0, 0.0, 20.0)])
def test_retarded_distance(vinf):
    """r = a0 t with t the smallest positive root of |d - vinf t| = a0 t (core.cu:261-279)."""
    rng = np.random.default_rng(11)
    nT = 8
    dPsi, dTau = _consts(nT)
    case = make_case([rotor_surface(nT, [1.0], [0.0], 200.0, dTau, [1.0])], [[1, 2, 3]], nT, dPsi,
                     dTau, vinf=vinf)
    a0 = case.a0
    v = np.array(vinf)
    for _ in range(50):
        src = (rng.standard_normal(3) * 2).astype(np.float32)
        obs = (rng.standard_normal(3) * 20).astype(np.float32)
        r = O.orc.orc_retarded_r(C.byref(case), O.p32(src), O.p32(obs))
        d = obs.astype(np.float64) - src.astype(np.float64)
        t = r / a0
        assert t > 0
        np.testing.assert_allclose(np.linalg.norm(d - v * t), a0 * t, rtol=1e-13)
        roots = np.roots([v @ v - a0 * a0, -2 * (d @ v), d @ d])
        pos = np.sort(roots[np.isreal(roots) & (roots.real > 0)].real)
        np.testing.assert_allclose(t, pos[0], rtol=1e-12)
        if not any(vinf):
            # still air: r = |d|, to the rounding of the canonical form (|d|^2 expanded about the
            # sources' centre and scaled into bins and back)
            assert abs(r / math.sqrt(float(d @ d)) - 1) < 1e-14


def _hover_blade(nT, dTau, radii, chord, P, phase=0.0):
    return rotor_surface(nT, radii, chord, 200.0, dTau, P, phase=phase)


def test_hover_observer_rotation_shifts_one_bin():
    """Minf = 0, psi-independent loading: observer rotated by +dPsi about the axis sees the same
    history one bin later (up to float rounding of the rotated coordinates)."""
    nT = 32
    dPsi, dTau = _consts(nT)
    rng = np.random.default_rng(5)
    E = 6
    blade = _hover_blade(nT, dTau, 0.3 + 0.7 * rng.random(E), 0.05 * rng.standard_normal(E),
                         100.0 * (1 + rng.random(E)))
    o = np.array([6.0, 2.0, -3.0])
    c, s = math.cos(dPsi), math.sin(dPsi)
    o2 = np.array([c * o[0] - s * o[1], s * o[0] + c * o[1], o[2]])
    case = make_case([blade], np.stack([o, o2]), nT, dPsi, dTau)
    pT, pL, pA = O.tonal_f64(case)
    for p in (pT, pL, pA):
        assert O.rel_l2(p[:, 1], np.roll(p[:, 0], 1)) < 2e-5


def test_two_blades_half_period_apart():
    """Blade 2 = blade 1 rotated by 180 degrees (exact negation of x, y): total history = one
    blade's history + its nT/2-bin shift."""
    nT = 24
    dPsi, dTau = _consts(nT)
    rng = np.random.default_rng(9)
    E = 4
    half = _hover_blade(nT, dTau, 0.3 + 0.7 * rng.random(E), 0.05 * rng.standard_normal(E),
                        100.0 * (1 + rng.random(E)))
    h = nT // 2
    b1 = Surface(half.X.copy(), half.N.copy(), half.V.copy(), half.PR.copy())
    for arr, comps in ((b1.X, (0, 1)), (b1.N, (0, 1, 3, 4)), (b1.V, (0, 1, 3, 4))):
        for ci in comps:
            arr[:, h:, ci] = -arr[:, :h, ci]   # positions at psi + pi are exact negations
    b2 = Surface(np.roll(b1.X, -h, axis=1), np.roll(b1.N, -h, axis=1), np.roll(b1.V, -h, axis=1),
                 np.roll(b1.PR, -h, axis=1))
    obs = [[7.0, -2.0, -3.0], [0.0, 5.0, -5.0]]
    one = O.tonal_f64(make_case([b1], obs, nT, dPsi, dTau))
    two = O.tonal_f64(make_case([b1, b2], obs, nT, dPsi, dTau))
    for p1, p2 in zip(one, two):
        np.testing.assert_allclose(p2, p1 + np.roll(p1, h, axis=0), rtol=1e-11,
                                   atol=1e-12 * np.abs(p1).max())


def test_float_instance_tracks_fp64():
    rng = np.random.default_rng(4)
    case = _moving_case(rng, nT=16, vinf=(25.0, 0.0, -2.0), nobs=4)
    d = O.tonal_f64(case)
    f = O.tonal_f32(case)
    for a, b in zip(f, d):
        assert O.rel_l2(a, b) < 1e-5
