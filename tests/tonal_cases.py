"""Small hand-built tonal cases (reference AoS layout, io.cu:243-261) for the oracle and GPU
parity tests, and a literal restatement of the reference's per-eval steps for brute force.

Layout per surface s, element j, step k (io.cu:280-284):
  X[XPtr[s] + (j nT + k) 4 + (x, y, z, ds)], N[.. 6 + (n, ndot)], V[.. 6 + (v, vdot)],
  PR[.. 2 + (P, Pdot)].
"""
import ctypes as C
import math

import numpy as np

import acum3_periodic_amd as acum


class Surface:
    """Arrays of one surface, shaped [E][nT][w]."""

    def __init__(self, X, N, V, PR):
        self.X = np.asarray(X, np.float32)
        self.N = np.asarray(N, np.float32)
        self.V = np.asarray(V, np.float32)
        self.PR = np.asarray(PR, np.float32)
        assert self.X.shape[2] == 4 and self.N.shape[2] == 6 and self.V.shape[2] == 6
        assert self.PR.shape[2] == 2


def make_case(surfaces, obs, nT, dPsi, dTau, a0=340.0, rhoRef=1.225, oM=200.0, vinf=(0.0, 0.0, 0.0),
              precision=acum.FP64):
    """TonalCase over numpy-owned arrays (kept alive on the returned object)."""
    E = np.array([s.X.shape[0] for s in surfaces], np.int64)
    for s in surfaces:
        assert s.X.shape[1] == nT
    X = np.concatenate([s.X.reshape(-1) for s in surfaces]).astype(np.float32)
    N = np.concatenate([s.N.reshape(-1) for s in surfaces]).astype(np.float32)
    V = np.concatenate([s.V.reshape(-1) for s in surfaces]).astype(np.float32)
    PR = np.concatenate([s.PR.reshape(-1) for s in surfaces]).astype(np.float32)
    off = np.concatenate([[0], np.cumsum(E)[:-1]]).astype(np.int64) * nT
    ptr = {w: (off * w).astype(np.int64) for w in (4, 6, 2)}
    ob = np.ascontiguousarray(obs, np.float32).reshape(-1)
    c = acum.TonalCase()
    c.nSurf, c.nTime, c.nObs = len(surfaces), nT, ob.size // 3
    keep = dict(E=E, X=X, N=N, V=V, PR=PR, p4=ptr[4], p6=ptr[6], p2=ptr[2], obs=ob)
    i64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
    f32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    c.E = i64(E)
    c.X, c.N, c.V, c.PR = f32(X), f32(N), f32(V), f32(PR)
    c.XPtr, c.NPtr, c.VPtr, c.PRPtr = i64(ptr[4]), i64(ptr[6]), i64(ptr[6]), i64(ptr[2])
    c.rObs = f32(ob)
    c.dPsi, c.dTau, c.a0, c.rhoRef, c.oM = dPsi, dTau, a0, rhoRef, oM
    for i in range(3):
        c.vInf[i] = vinf[i]
    c.precision = precision
    c._keep = keep
    return c


def rotor_surface(nT, radii, chord_x, omega, dTau, P, phase=0.0, z=0.0, rng=None, cone=0.0):
    """One rotating blade: E points at the given radii (and chordwise offsets), rotating about z
    at omega with analytic velocity, acceleration, normal and normal rate.  P is [E][nT] or [E]."""
    E = len(radii)
    X = np.zeros((E, nT, 4))
    N = np.zeros((E, nT, 6))
    V = np.zeros((E, nT, 6))
    PR = np.zeros((E, nT, 2))
    P = np.asarray(P, np.float64)
    if P.ndim == 1:
        P = np.repeat(P[:, None], nT, axis=1)
    for j in range(E):
        for k in range(nT):
            psi = phase + omega * dTau * k
            cs, sn = math.cos(psi), math.sin(psi)
            x0, y0 = radii[j], chord_x[j]
            x, y = cs * x0 - sn * y0, sn * x0 + cs * y0
            X[j, k] = (x, y, z + cone * radii[j], 1e-3 * (1 + 0.1 * j))
            V[j, k, :3] = (-omega * y, omega * x, 0.0)
            V[j, k, 3:] = (-omega * omega * x, -omega * omega * y, 0.0)
            n0 = np.array([0.1 * math.sin(j + 1.0), 0.05, 1.0])
            n0 /= np.linalg.norm(n0)
            nx, ny = cs * n0[0] - sn * n0[1], sn * n0[0] + cs * n0[1]
            N[j, k, :3] = (nx, ny, n0[2])
            N[j, k, 3:] = (-omega * ny, omega * nx, 0.0)
    PR[:, :, 0] = P
    PR[:, :, 1] = np.gradient(P, dTau, axis=1) if nT > 2 else 0.0
    return Surface(X, N, V, PR)


# ---- literal restatement of the reference's per-eval steps (core.cu:18-355), FP64 ------------
def literal_reference(case):
    """pT, pL, pA [nT][nObs] and the per-eval (ipts, npi, n_down, n_up) of observer 0, computed
    the way findPressureTimeHistory does: general Farassat 1A terms with rho = rhodot = 1, U = V
    and the dotted normal velocities forced to zero (core.cu:68-154), observer time in radians
    (core.cu:156), findpts/bsearch (core.cu:290-355) and the two sweeps (core.cu:196-231) with
    the running grid time tObs2.  Pure Python, for tiny cases."""
    nT, nObs = case.nTime, case.nObs
    a0, oM, dPsi, dTau, rhoRef = case.a0, case.oM, case.dPsi, case.dTau, case.rhoRef
    vinf = np.array([case.vInf[0], case.vInf[1], case.vInf[2]])
    lpi = nT * dPsi
    tObs = np.arange(nT) * dPsi
    arr = acum.case_arrays(case)
    E = arr["E"]
    X = arr["X"].astype(np.float64)
    N = arr["N"].astype(np.float64)
    V = arr["V"].astype(np.float64)
    PR = arr["PR"].astype(np.float64)
    obs_all = arr["rObs"].astype(np.float64)
    pT = np.zeros((nT, nObs))
    pL = np.zeros((nT, nObs))
    pA = np.zeros((nT, nObs))
    bins0 = []

    def acous_dist(src, obs):
        a = vinf @ vinf - a0 * a0
        d = obs - src
        b = -2.0 * (d @ vinf)
        c = d @ d
        disc = b * b - 4.0 * a * c
        t = (-b - math.sqrt(disc)) / (2.0 * a)
        if t < 0.0:
            t = (-b + math.sqrt(disc)) / (2.0 * a)
        return t * a0

    def findpts(t):
        npi = math.floor((lpi + tObs[0] - t) / lpi)
        newx = t + npi * lpi
        lo, hi = 0, nT  # bsearch: largest k with tObs[k] <= newx
        while True:
            mid = (lo + hi) // 2
            if newx < tObs[mid]:
                hi = mid
            else:
                lo = mid
            if lo + 1 >= hi:
                break
        assert tObs[lo] <= newx
        return lo, npi

    base = 0
    for s in range(case.nSurf):
        for j in range(E[s]):
            e = base + j
            for io in range(nObs):
                obs = obs_all[io]
                for it in range(nT):
                    q = e * nT + it
                    src = X[4 * q:4 * q + 3]
                    ds = X[4 * q + 3]
                    Vv, Vd = V[6 * q:6 * q + 3], V[6 * q + 3:6 * q + 6]
                    Nn, Nd = N[6 * q:6 * q + 3], N[6 * q + 3:6 * q + 6]
                    P, Pd = PR[2 * q], PR[2 * q + 1]
                    rho, rhodot = 1.0, 1.0
                    r = acous_dist(src, obs)
                    rhat = (obs - src - vinf * r / a0) / r
                    Mr = (Vv @ rhat) / a0
                    M = math.sqrt(Vv @ Vv) / a0
                    Mdr = (Vd @ rhat) / a0
                    D1 = 1.0 / (4.0 * math.pi * a0)
                    T1 = D1 / (r * (1.0 - Mr) ** 2)
                    T2 = T1 / (1.0 - Mr)
                    T3 = T1 / r
                    T4 = T3 / (1.0 - Mr)
                    U, Ud = Vv, Vd
                    Vn = Vv @ Nn
                    Un = U @ Nn
                    Vn_dot = Un_dot = 0.0
                    Ma_n = (Un - Vn) * rho + rhoRef * Vn
                    Man_dot = (Un_dot - Vn_dot) * rho + (Un - Vn) * rhodot + rhoRef * Vn_dot
                    TF2 = Man_dot * T1 * a0 * ds + Ma_n * Mdr * T2 * a0 * ds
                    TN = a0 * Ma_n * (Mr - M * M) * T4 * a0 * ds
                    thick = TN + TF2
                    Li = P * Nn
                    Lid = Pd * Nn + P * Nd
                    LF1 = (Lid @ rhat) * T1 * ds + (Li @ rhat) * Mdr * T2 * ds
                    LN = ((Li @ rhat) - (Li @ Vv) / a0) * T3 * a0 * ds + (Li @ rhat) * (Mr - M * M) * T4 * a0 * ds
                    Mo = rhoRef * U * (Un - Vn)
                    Mod = rhoRef * Ud * (Un - Vn) + rhoRef * U * (Un_dot - Vn_dot)
                    LF2 = (Mod @ rhat) * T1 * ds + (Mo @ rhat) * Mdr * T2 * ds
                    LN = LN + ((Mo @ rhat) - (Mo @ Vv) / a0) * T3 * a0 * ds + (Mo @ rhat) * (Mr - M * M) * T4 * a0 * ds
                    load = LN + LF1 + LF2
                    ttemp = (it * dTau + r / a0) * oM
                    ipts, npi = findpts(ttemp)
                    mm, nm = (it - 1, 0) if it > 0 else (nT - 1, 1)
                    mp, np_ = (it + 1, 0) if it < nT - 1 else (0, 1)
                    qm, qp = e * nT + mm, e * nT + mp
                    ttem = (mm * dTau + acous_dist(X[4 * qm:4 * qm + 3], obs) / a0) * oM - nm * lpi
                    ttep = (mp * dTau + acous_dist(X[4 * qp:4 * qp + 3], obs) / a0) * oM + np_ * lpi
                    iptm, t2, nd = ipts, tObs[ipts], 0
                    while t2 > ttem + npi * lpi:
                        w = (ttem + npi * lpi - t2) / (ttem - ttemp)
                        pT[iptm, io] += thick * w
                        pL[iptm, io] += load * w
                        pA[iptm, io] += (thick + load) * w
                        t2 -= dPsi
                        iptm = iptm - 1 if iptm > 0 else nT - 1
                        nd += 1
                    iptp = ipts + 1 if ipts < nT - 1 else 0
                    t2 = tObs[iptp] + (lpi if iptp == 0 else 0.0)
                    nu = 0
                    while t2 < ttep + npi * lpi:
                        w = (ttep + npi * lpi - t2) / (ttep - ttemp)
                        pT[iptp, io] += thick * w
                        pL[iptp, io] += load * w
                        pA[iptp, io] += (thick + load) * w
                        t2 += dPsi
                        iptp = iptp + 1 if iptp < nT - 1 else 0
                        nu += 1
                    if io == 0:
                        bins0.append((ipts, npi, nd, nu))
        base += E[s]
    return pT, pL, pA, np.array(bins0, np.int64)
