"""The tonal oracle's Farassat 1A thickness and loading terms (oracle/orc_tonal.cpp, restating
core.cu:93-154) against an independent formulation: Farassat's Formulation 1, in which the
observer-time derivative is taken numerically outside the retarded-time sums,

    4 pi p_T(t) = d/dt  sum_e [rho0 Vn dS / (r (1 - Mr))]_ret
    4 pi p_L(t) = (1/a0) d/dt sum_e [L_r dS / (r (1 - Mr))]_ret + sum_e [L_r dS / (r^2 (1 - Mr))]_ret

with L_r = P n.r_hat.  The retarded-time sums use the same tent (core.cu:166-231, a partition of
unity in observer time, so binning a source-time quantity interpolates it at the retarded time)
and the derivative is a periodic central difference in observer time.  For a blade in rigid
rotation in still air V.n is constant along each element's path, so the dotted normal velocity
the reference forces to zero (core.cu:117-118) is zero and the two formulations describe the same
field: they must agree to discretisation error, which falls with the time step (between first
and second order: the central difference of a linearly interpolated sum is first order in the
interpolation error; measured 1.1e-3 -> 3.5e-4 for thickness and 1.1e-3 -> 4.2e-4 for loading
from 360 to 720 steps).  A dropped or mis-signed 1A term leaves an error that does not fall with
the step: planted in the oracle, dropping the dMr thickness term gives 4.9, Mr + M^2 for Mr - M^2
in the thickness term 0.43, and L_r + L_M for L_r - L_M in the loading term 0.13, at both steps.
"""
import math

import numpy as np
import pytest

import oracle as O
from tonal_cases import Surface, make_case

A0, RHO, OM = 340.0, 1.225, 200.0


def rigid_blade(nT, P_of):
    """Six points of a blade rotating about z at OM (radius 0.3 .. 1.0 m, chordwise offsets,
    z = 0.04 r), normals tilted out of the rotor plane, analytic V = Om x X, dV/dt, dn/dt and
    pressure P_of(j, phase) with its analytic time derivative."""
    dPsi = 2.0 * math.pi / nT
    dTau = dPsi / OM
    E = 6
    radii = np.linspace(0.3, 1.0, E)
    chord = 0.03 * np.sin(np.arange(E) + 0.5)
    X = np.zeros((E, nT, 4))
    N = np.zeros((E, nT, 6))
    V = np.zeros((E, nT, 6))
    PR = np.zeros((E, nT, 2))
    for j in range(E):
        n0 = np.array([0.3 * math.cos(j + 0.2), 0.45 * math.sin(1.3 * j + 0.4), 1.0])
        n0 /= np.linalg.norm(n0)
        for k in range(nT):
            psi = OM * dTau * k + 0.7 * j
            cs, sn = math.cos(psi), math.sin(psi)
            x0, y0 = radii[j], chord[j]
            x, y = cs * x0 - sn * y0, sn * x0 + cs * y0
            X[j, k] = (x, y, 0.04 * radii[j], 2.0e-3 * (1.0 + 0.1 * j))
            V[j, k, :3] = (-OM * y, OM * x, 0.0)
            V[j, k, 3:] = (-OM * OM * x, -OM * OM * y, 0.0)
            nx, ny = cs * n0[0] - sn * n0[1], sn * n0[0] + cs * n0[1]
            N[j, k] = (nx, ny, n0[2], -OM * ny, OM * nx, 0.0)
            PR[j, k] = P_of(j, psi)
    return [Surface(X, N, V, PR)], dPsi, dTau


def loading(j, psi):
    """P and dP/dt of point j at blade azimuth psi (rotor speed OM)."""
    a, b, m = 900.0 + 150.0 * j, 250.0, 3
    return a + b * math.sin(m * psi + 0.3 * j), b * m * OM * math.cos(m * psi + 0.3 * j)


OBS = np.array([[6.0, 1.5, -2.0], [-3.0, -7.5, 1.0], [2.0, 4.0, -8.0]], np.float32)


def formulation1(surfs, nT, dPsi, dTau):
    """Thickness and loading by Formulation 1, in FP64 numpy, from the same float inputs."""
    s = surfs[0]
    X, N, V, PR = (a.astype(np.float32).astype(np.float64) for a in (s.X, s.N, s.V, s.PR))
    E = X.shape[0]
    dt = dTau                               # observer-time bin = dPsi / oM = one source step
    pT = np.zeros((nT, len(OBS)))
    pL = np.zeros((nT, len(OBS)))
    for io, ob in enumerate(OBS.astype(np.float64)):
        Q = np.zeros(nT)
        F = np.zeros(nT)
        G = np.zeros(nT)
        for j in range(E):
            d = ob[None, :] - X[j, :, :3]
            r = np.linalg.norm(d, axis=1)
            rh = d / r[:, None]
            Mr = np.einsum("ki,ki->k", V[j, :, :3], rh) / A0
            Vn = np.einsum("ki,ki->k", V[j, :, :3], N[j, :, :3])
            Lr = PR[j, :, 0] * np.einsum("ki,ki->k", N[j, :, :3], rh)
            ds = X[j, :, 3]
            q = RHO * Vn * ds / (4 * math.pi * r * (1 - Mr))
            f = Lr * ds / (4 * math.pi * A0 * r * (1 - Mr))
            g = Lr * ds / (4 * math.pi * r * r * (1 - Mr))
            tau = np.arange(nT) + r / (A0 * dt)  # observer time in bins
            for k in range(nT):
                tp = tau[k - 1] - (nT if k == 0 else 0)
                tc = tau[k]
                tn = tau[(k + 1) % nT] + (nT if k == nT - 1 else 0)
                b = math.floor(tc)
                while b > tp:                 # rising side (tau_prev, tau]
                    w = (b - tp) / (tc - tp)
                    Q[b % nT] += q[k] * w
                    F[b % nT] += f[k] * w
                    G[b % nT] += g[k] * w
                    b -= 1
                b = math.floor(tc) + 1
                while b < tn:                 # falling side (tau, tau_next)
                    w = (tn - b) / (tn - tc)
                    Q[b % nT] += q[k] * w
                    F[b % nT] += f[k] * w
                    G[b % nT] += g[k] * w
                    b += 1
        ddt = lambda y: (np.roll(y, -1) - np.roll(y, 1)) / (2 * dt)
        pT[:, io] = ddt(Q)
        pL[:, io] = ddt(F) + G
    return pT, pL


def errors(nT):
    surfs, dPsi, dTau = rigid_blade(nT, loading)
    c = make_case(surfs, OBS, nT, dPsi, dTau, a0=A0, rhoRef=RHO, oM=OM)
    oT, oL, _ = O.tonal_f64(c)
    fT, fL = formulation1(surfs, nT, dPsi, dTau)
    return O.rel_l2(fT, oT), O.rel_l2(fL, oL)


def test_formulation1_agrees_with_1a_as_the_step_falls():
    eT1, eL1 = errors(360)
    eT2, eL2 = errors(720)
    # agreement to discretisation error that falls at least twice per halving of the step
    assert eT2 < 1e-3 and eL2 < 1e-3, (eT2, eL2)
    assert eT1 / eT2 > 2.0 and eL1 / eL2 > 2.0, (eT1, eT2, eL1, eL2)
