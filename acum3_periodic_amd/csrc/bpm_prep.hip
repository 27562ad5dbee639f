// BPM source preparation on the GPU: the observer-independent per (surface, section, step) state
// of findBBnoise (bb_core.cu:60-440) and the section frames of its observer loop
// (bb_core.cu:457-513), one thread per (section, step), in FP64.
//
// Per source: blade-section kinematics with the quarter-chord y, z taken from section
// round(0.75 nsect) (bb_core.cu:79-164), relative flow by periodic central differences of the
// section position (bb_core.cu:166-206), Mach/Re/section brackets and the trilinear cl lookup
// (bb_core.cu:208-250, bb_utils.cu:124-168), cl slope, cl(0) and alpha(cl = 0)
// (bb_core.cu:252-287), the fixed-point effective angle (bb_core.cu:291-321) and the
// boundary-layer thickness models t_del = 0, 1, 2 (bb_core.cu:354-438).  Positions of the
// neighbouring steps and sections are recomputed from the section table rather than stored, so
// one launch writes the records the shape and observer kernels read, with no host pass.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include "bpm_types.hpp"

namespace acum {

struct BpmSurfDev {
  int64_t src0;                       // first section of the surface in the source numbering
  int64_t xoff, doff;                 // offsets into xyzcR and bbdata (xsPtr, dsPtr)
  int64_t soff, moff, roff, aoff, toff;  // table offsets (sPtr, mPtr, rPtr, aPtr, dPtr)
  int32_t ns, iq, ccw, trip;          // sections, quarter-chord section, rotation sense, trip
  int32_t nst, nm, nr, na;            // table sections, Mach, Re and AoA counts
  double span, omega, psi0;           // psi0 = psi_offset in radians
  double cx, sx, cy, sy, cz, sz;      // cosines and sines of the fixed rotations about x, y, z
  double trn[3];
};

struct BpmPrepArgs {
  const BpmSurfDev* surf;
  int32_t nS, nT, t_del, pad;
  int64_t nsec;                       // sections over all surfaces
  const float *xyzcR, *bbdata, *sect_t, *mach_t, *re_t, *aoa_t, *bbdata_t;
  double dTau, a0, vx, vy, vz;
  BpmShapeIn* in;
  BpmSrcDev* dev;
  uint32_t* err;  // bit 0: a table lacks a negative AoA, an AoA >= 6 deg or a cl sign change;
                  // bit 1: an effective angle below the table's AoA range (t_del = 1)
};

namespace {

constexpr double kD2R = M_PI / 180.0;
constexpr double kNu = 0.0000181206 / 1.22500;  // kinematic viscosity (bb_core.cu:47)

__device__ inline double sq(double x) { return x * x; }

struct Placer {
  const BpmSurfDev& S;
  const float* xc;
  // section is at azimuth psi (cos cp, sin sp), then the fixed rotations about x, y, z, then the
  // translation
  __device__ void place(const double* a, double cp, double sp, double* o) const {
    const double p[3] = {a[0] * cp - a[1] * sp, a[0] * sp + a[1] * cp, a[2]};
    const double q[3] = {p[0], p[1] * S.cx - p[2] * S.sx, p[1] * S.sx + p[2] * S.cx};
    const double r[3] = {q[0] * S.cy + q[2] * S.sy, q[1], -q[0] * S.sy + q[2] * S.cy};
    o[0] = r[0] * S.cz - r[1] * S.sz + S.trn[0];
    o[1] = r[0] * S.sz + r[1] * S.cz + S.trn[1];
    o[2] = r[2] + S.trn[2];
  }
  // blade-frame position of section is (its radius, the quarter-chord section's y and z)
  __device__ void base(int is, double* x0) const {
    x0[0] = xc[4 * is] * S.span;
    x0[1] = xc[4 * S.iq + 1] * S.span;
    x0[2] = xc[4 * S.iq + 2] * S.span;
    if (S.ccw == 0) x0[1] = -x0[1];
  }
  __device__ void azimuth(double dTau, int it, double* cp, double* sp) const {
    double psi = dTau * it * S.omega + S.psi0;
    if (S.ccw == 0) psi = -psi;
    sincos(psi, sp, cp);
  }
};

struct Table {
  const float *sect, *mach, *re, *aoa, *data;
  int ns, nm, nr, na;
  __device__ double at(int is, int im, int ir, int ia, int id) const {
    return data[(((int64_t)is * nm + im) * nr + ir) * na * 7 + (int64_t)ia * 7 + id];
  }
  // the first index whose axis value exceeds x, then the clamped bracket (bb_core.cu:208-234)
  __device__ static void bracket(const float* ax, int n, double x, int* lo, int* hi) {
    int k = 0;
    while (k < n && !(x < ax[k])) k++;
    *lo = k - 1 > 0 ? k - 1 : 0;
    *hi = k < n - 1 ? k : n - 1;
  }
};

// trilinear interpolation in (section, Mach, Re) of table column id at AoA index ia
struct Lerp3 {
  const Table& T;
  double ws, wm, wr;
  int s1, s2, m1, m2, r1, r2;
  __device__ double operator()(int id, int ia) const {
    auto sec = [&](int im, int ir) { return (1 - ws) * T.at(s1, im, ir, ia, id) + ws * T.at(s2, im, ir, ia, id); };
    return (1 - wm) * (1 - wr) * sec(m1, r1) + wm * (1 - wr) * sec(m2, r1) + (1 - wm) * wr * sec(m1, r2) +
           wm * wr * sec(m2, r2);
  }
};

__global__ void bpm_prep_kernel(BpmPrepArgs A) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= A.nsec * A.nT) return;
  const int64_t g = q / A.nT;
  const int it = (int)(q % A.nT);
  int s = 0;
  while (s + 1 < A.nS && A.surf[s + 1].src0 <= g) s++;
  const BpmSurfDev& S = A.surf[s];
  const int is = (int)(g - S.src0), ns = S.ns, nT = A.nT;
  const float* xc = A.xyzcR + S.xoff;
  const float* bd = A.bbdata + S.doff;
  const Placer pl{S, xc};
  const double ch = xc[4 * is + 3] * S.span;
  // positions: this section at this step and the neighbouring steps, the trailing edge, and the
  // neighbouring sections (bb_core.cu:79-164, 473-487)
  double cp, sp, cpm, spm, cpp, spp;
  pl.azimuth(A.dTau, it, &cp, &sp);
  pl.azimuth(A.dTau, it == 0 ? nT - 1 : it - 1, &cpm, &spm);
  pl.azimuth(A.dTau, it == nT - 1 ? 0 : it + 1, &cpp, &spp);
  double x0[3];
  pl.base(is, x0);
  const double th = bd[(int64_t)it * ns * 7 + is * 7 + 6] * kD2R;
  double xt0[3] = {x0[0], 0.0, x0[2] - ch * 0.75 * sin(th)};
  xt0[1] = (S.ccw == 1) ? x0[1] - ch * 0.75 * cos(th) : x0[1] + ch * 0.75 * cos(th);
  double P[3], Pt[3], Pm[3], Pp[3];
  pl.place(x0, cp, sp, P);
  pl.place(xt0, cp, sp, Pt);
  pl.place(x0, cpm, spm, Pm);
  pl.place(x0, cpp, spp, Pp);
  // relative flow along the chord (bb_core.cu:166-206)
  const double vinf[3] = {A.vx, A.vy, A.vz};
  double v[3], c1[3];
  for (int i = 0; i < 3; i++) v[i] = vinf[i] - (Pp[i] - Pm[i]) / (2 * A.dTau);
  for (int i = 0; i < 3; i++) c1[i] = Pt[i] - P[i];
  const double l1 = sqrt(sq(c1[0]) + sq(c1[1]) + sq(c1[2]));
  const double u = (v[0] * c1[0] + v[1] * c1[1] + v[2] * c1[2]) / l1;
  const double M = fabs(u / A.a0), Re = fabs(u * ch / kNu), rR = xc[4 * is];
  // table brackets and the cl curve (bb_core.cu:208-287)
  const Table T{A.sect_t + S.soff, A.mach_t + S.moff, A.re_t + S.roff, A.aoa_t + S.aoff, A.bbdata_t + S.toff,
                S.nst, S.nm, S.nr, S.na};
  Lerp3 L{T, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  Table::bracket(T.mach, T.nm, M, &L.m1, &L.m2);
  Table::bracket(T.re, T.nr, Re, &L.r1, &L.r2);
  Table::bracket(T.sect, T.ns, rR, &L.s1, &L.s2);
  L.ws = (L.s1 == L.s2) ? 0.5 : (rR - T.sect[L.s1]) / ((double)T.sect[L.s2] - T.sect[L.s1]);
  L.wm = (L.m1 == L.m2) ? 0.5 : (M - T.mach[L.m1]) / ((double)T.mach[L.m2] - T.mach[L.m1]);
  L.wr = (L.r1 == L.r2) ? 0.5 : (Re - T.re[L.r1]) / ((double)T.re[L.r2] - T.re[L.r1]);
  int a1 = -1, a2 = -1, a3 = -1;
  for (int i = 0; i < T.na && (a1 < 0 || a2 < 0 || a3 < 0); i++) {
    if (a1 < 0 && T.aoa[i] >= 0.0f) a1 = i;
    if (a2 < 0 && T.aoa[i] >= 6.0f) a2 = i;
    if (a3 < 0 && L(0, i) > 0) a3 = i;
  }
  if (a1 < 1 || a2 < 0 || a3 < 1) {
    atomicOr(A.err, 1u);
    return;
  }
  const double cla = (L(0, a2) - L(0, a1)) / ((double)T.aoa[a2] - T.aoa[a1]);
  double w = (0 - T.aoa[a1 - 1]) / ((double)T.aoa[a1] - T.aoa[a1 - 1]);
  const double cl0 = L(0, a1 - 1) * (1 - w) + L(0, a1) * w;
  const double c3m = L(0, a3 - 1), c3 = L(0, a3);
  w = (0 - c3m) / (c3 - c3m);
  const double acl0 = T.aoa[a3 - 1] * (1 - w) + T.aoa[a3] * w;
  // effective angle of attack by fixed-point iteration (bb_core.cu:291-321)
  const double cn = bd[(int64_t)it * ns * 7 + is * 7 + 4], ct = bd[(int64_t)it * ns * 7 + is * 7 + 5];
  double ae = 0.0, e = 1.0;
  for (int iter = 0; e > 1e-3;) {
    const double clsec = cn * cos(ae * kD2R) - ct * sin(ae * kD2R);
    const double ae1 = (clsec - cl0) / cla;
    e = fabs(ae1 - ae);
    ae = ae1;
    if (++iter > 50) break;
  }
  const double alpha = fabs(ae - acl0);
  // boundary-layer thicknesses (bb_core.cu:354-438)
  double bl[4];
  if (A.t_del == 0) {
    for (int i = 0; i < 4; i++) bl[i] = bd[(int64_t)it * ns * 7 + is * 7 + i];
  } else if (A.t_del == 1) {
    for (int i = 0; i < T.na; i++)
      if (T.aoa[i] >= ae) {
        a1 = i;
        break;
      }
    if (a1 < 1) {  // below the table's first AoA: no bracket (the reference reads before the array)
      atomicOr(A.err, 2u);
      return;
    }
    w = (ae - T.aoa[a1 - 1]) / ((double)T.aoa[a1] - T.aoa[a1 - 1]);
    for (int i = 0; i < 4; i++) bl[i] = ch * (L(i + 3, a1 - 1) * (1 - w) + L(i + 3, a1) * w);
  } else {
    const double lr = log10(Re), al = alpha;
    const double d0 = ch * pow(10.0, 1.6569 - 0.9045 * lr + 0.0596 * lr * lr);
    const double ds0 = ch * pow(10.0, 3.0187 - 1.5397 * lr + 0.1059 * lr * lr);
    bl[1] = d0 * pow(10.0, -0.04175 * al + 0.00106 * al * al);
    bl[3] = ds0 * pow(10.0, -0.0432 * al + 0.00113 * al * al);
    if (al <= 7.5) bl[0] = d0 * pow(10.0, 0.03114 * al);
    else if (al > 12.5) bl[0] = d0 * 12.0 * pow(10.0, 0.0258 * al);
    else bl[0] = d0 * 0.0303 * pow(10.0, 0.2336 * al);
    if (al <= 7.5) bl[2] = ds0 * pow(10.0, 0.0679 * al);
    else if (al > 12.5) bl[2] = ds0 * 52.42 * pow(10.0, 0.0258 * al);
    else bl[2] = ds0 * 0.0162 * pow(10.0, 0.3066 * al);
  }
  for (int i = 0; i < 4; i++) bl[i] = bl[i] < 0.0 ? 1.0e-10 : bl[i];
  // section length and frame (bb_core.cu:457-513)
  double ls;
  if (is == 0) ls = ((double)xc[4 * (is + 1)] - xc[4 * is]) / 2.0 * S.span;
  else if (is == ns - 1) ls = ((double)xc[4 * is] - xc[4 * (is - 1)]) / 2.0 * S.span;
  else ls = ((double)xc[4 * (is + 1)] - xc[4 * (is - 1)]) / 2.0 * S.span;
  const int km = (is == ns - 1) ? is - 1 : is, kp = (is == ns - 1) ? is : is + 1;
  double xk[3], Pkm[3], Pkp[3];
  pl.base(km, xk);
  pl.place(xk, cp, sp, Pkm);
  pl.base(kp, xk);
  pl.place(xk, cp, sp, Pkp);
  double t0[3], t1[3], t2[3];
  for (int i = 0; i < 3; i++) {
    t0[i] = Pt[i] - P[i];
    t1[i] = Pkp[i] - Pkm[i];
  }
  const double n0 = sqrt(sq(t0[0]) + sq(t0[1]) + sq(t0[2]));
  const double n1 = sqrt(sq(t1[0]) + sq(t1[1]) + sq(t1[2]));
  for (int i = 0; i < 3; i++) {
    t0[i] /= n0;
    t1[i] /= n1;
  }
  if (S.ccw == 1) {
    t2[0] = t0[1] * t1[2] - t0[2] * t1[1];
    t2[1] = t0[2] * t1[0] - t0[0] * t1[2];
    t2[2] = t0[0] * t1[1] - t0[1] * t1[0];
    const double n2 = sqrt(sq(t2[0]) + sq(t2[1]) + sq(t2[2]));
    for (int i = 0; i < 3; i++) t2[i] /= n2;
  } else {
    t2[0] = t1[1] * t0[2] - t1[2] * t0[1];
    t2[1] = t1[2] * t0[0] - t1[0] * t0[2];
    t2[2] = t1[0] * t0[1] - t1[1] * t0[0];
    const double n2 = sqrt(sq(t2[0]) + sq(t2[1]) + sq(t2[2]));
    for (int i = 0; i < 3; i++) t2[i] /= sqrt(n2);  // the reference's normalisation (bb_core.cu:512)
  }
  // records of the shape kernel (FP64) and the observer kernel (FP32)
  const int64_t o = g * nT + it;
  BpmShapeIn a;
  a.u = u;
  a.c = ch;
  a.lsect = ls;
  a.alpha = alpha;
  a.ds = bl[0];
  a.dp = bl[1];
  a.dss = bl[2];
  a.dsp = bl[3];
  a.R = S.span;
  a.trip = S.trip;
  a.tip = (is == ns - 1) ? 1 : 0;
  A.in[o] = a;
  BpmSrcDev d;
  for (int i = 0; i < 3; i++) d.xt[i] = (float)Pt[i];
  for (int i = 0; i < 3; i++) {
    d.tr[i] = (float)t0[i];
    d.tr[3 + i] = (float)t1[i];
    d.tr[6 + i] = (float)t2[i];
  }
  const double Ma = fabs(u) / A.a0;
  d.M = (float)Ma;
  d.Mmc = (float)(Ma - 0.6 * Ma);
  d.tipT0 = 0.0f;
  d.flags = 0;
  A.dev[o] = d;
}

}  // namespace

hipError_t launch_bpm_prep(hipStream_t s, const BpmPrepArgs& a) {
  const int64_t n = a.nsec * a.nT;
  const int grid = (int)((n + 127) / 128);
  hipLaunchKernelGGL(bpm_prep_kernel, dim3(grid), dim3(128), 0, s, a);
  return hipGetLastError();
}

}  // namespace acum
