// Tonal Farassat 1A (thickness + loading) kernels for gfx950.
//
// Work decomposition (the reference launches one thread per (step, observer, element) with
// float atomics on pT/pL/pA, core.cu:18-259, main.cu:194-236):
//   * one wave = 64 observers (one per lane); element records are wave-uniform, so every record
//     read from HBM serves 64 observers;
//   * a work item = (case, observer group, group of tiles); a tile is a run of elements that sit
//     close together; for each tile the wave walks the source steps in passes of K steps and, for
//     each element of the tile, the K steps of that pass in order -- one FP64 retarded solve per
//     (element, step, observer) instead of the reference's three (core.cu:84,189,193);
//   * each lane owns a column of an LDS window of W observer-time bins; tent-weighted
//     contributions (core.cu:196-231) go there with LDS atomic adds, which are ordered because no
//     other lane touches that column; after each pass the window is added into the item's row of
//     a partial slab in HBM;
//   * a reduction kernel sums the partial slabs of each (observer, bin) in fixed order with
//     FP64 Kahan compensation and writes pT, pL and pA = pT + pL.
// Results are deterministic: no operation depends on scheduling order.
//
// Precision: MIXED computes observer time, bin index and the bin time deltas in FP64 and the
// geometry, Doppler factor, integrands and weights in FP32 (FP32 window and slab); PLAIN does
// everything in FP64.  The FP64 observer-time/bin code is the canonical form the CPU oracle
// uses, compiled without contraction, so bin indices agree bit for bit.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "dev_types.hpp"

namespace acum {

// ---------------------------------------------------------------------------------------------
// Canonical FP64 observer time (acous_dist core.cu:261-279, ttemp core.cu:156) in bin units; the
// CPU oracle evaluates the same operations in the same order.
__device__ __forceinline__ double retarded_r(const TimeConst& k, double dx, double dy, double dz) {
  const double c = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
  if (k.still) return sqrt(c);
  const double b = -2.0 * __builtin_fma(dz, k.vz, __builtin_fma(dy, k.vy, dx * k.vx));
  const double disc = __builtin_fma(b, b, -(k.four_a * c));
  const double sq = sqrt(disc);
  double t = (-b - sq) * k.inv_2a;
  if (t < 0.0) t = (-b + sq) * k.inv_2a;
  return t * k.a0;
}

__device__ __forceinline__ double obs_tau(const TimeConst& k, int j, double r) {
  return __builtin_fma(r, k.r2b, (double)j);
}

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0; }

// Tent of one eval (findpts/bsearch core.cu:290-355 and the sweeps core.cu:196-231, in bin
// units).  The top rising bin is ic = floor(tau); the rising side holds nd = ic - floor(tau_prev)
// bins ic, ic-1, ... with weights (b - tau_prev) / (tau - tau_prev); the falling side holds
// nu = ceil(tau_next) - 1 - ic bins ic+1, ic+2, ... with weights (tau_next - b) / (tau_next - tau).
struct Tent {
  double fc, tp, tn;  // floor(tau), tau_prev, tau_next
  int ic, nd, nu;
};

__device__ __forceinline__ bool tent_of(double tp, double tc, double tn, Tent* T) {
  const double fc = floor(tc), fp = floor(tp), cn = ceil(tn);
  const bool ok = fabs(tc) < 1.0e9 && fabs(tp) < 1.0e9 && fabs(tn) < 1.0e9;  // finite, in range
  T->fc = fc;
  T->tp = tp;
  T->tn = tn;
  T->ic = ok ? (int)fc : 0;
  const int nd = ok ? (int)(fc - fp) : 0, nu = ok ? (int)(cn - fc) - 1 : 0;
  T->nd = nd > 0 ? nd : 0;
  T->nu = nu > 0 ? nu : 0;
  return ok;
}

// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float rcpT(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcpT(double x) { return 1.0 / x; }

// Records are read through global-address-space pointers: with a wave-uniform address each
// read is one vector memory request serving all 64 lanes, counted by vmcnt, so waiting for a
// record never waits for the LDS atomics of the bins (which count on lgkmcnt).
template <typename T>
using cptr = const __attribute__((address_space(1))) T*;

// Loads one record as whole dwords.
template <typename Rec>
__device__ __forceinline__ Rec load_rec(cptr<Rec> p) {
  static_assert(sizeof(Rec) % 4 == 0, "record must be whole dwords");
  constexpr int n = (int)(sizeof(Rec) / 4);
  cptr<uint32_t> w = (cptr<uint32_t>)p;
  uint32_t t[n];
#pragma unroll
  for (int i = 0; i < n; i++) t[i] = w[i];
  Rec r;
  __builtin_memcpy(&r, t, sizeof(Rec));
  return r;
}

template <bool MIXED>
struct Prec {
  using R = double;
  using R2 = double2;
  using Rec = Rec64;
};
template <>
struct Prec<true> {
  using R = float;
  using R2 = float2;
  using Rec = Rec32;
};

// Farassat 1A integrand of the impermeable surface (core.cu:68-154): rho = 1, U = V, the dotted
// normal velocities forced to zero (core.cu:117-118), so
//   pThick = rhoRef Vn dS a0 (dMr T2 + a0 (Mr - M^2) T4)
//   pLoad  = dS (dLr T1 + Lr dMr T2 + a0 (Lr - Lm) T3 + a0 Lr (Mr - M^2) T4)
// with T1 = 1/(4 pi a0 r (1-Mr)^2), T2 = T1/(1-Mr), T3 = T1/r, T4 = T3/(1-Mr).
template <typename R, typename Rec>
__device__ __forceinline__ void integrand(const Rec& q, R ox, R oy, R oz, R r, R inv_a0, R a0,
                                          R vx, R vy, R vz, R rhoRef, R D1, R* pth, R* plo,
                                          bool* bad) {
  const R rt = r * inv_a0;
  const R rvx = fmaT(-vx, rt, ox - (R)q.x[0]);
  const R rvy = fmaT(-vy, rt, oy - (R)q.x[1]);
  const R rvz = fmaT(-vz, rt, oz - (R)q.x[2]);
  const R ir = rcpT(r);
  const R hx = rvx * ir, hy = rvy * ir, hz = rvz * ir;
  const R Vr = fmaT((R)q.v[0], hx, fmaT((R)q.v[1], hy, (R)q.v[2] * hz));
  const R dVr = fmaT((R)q.vd[0], hx, fmaT((R)q.vd[1], hy, (R)q.vd[2] * hz));
  const R Nr = fmaT((R)q.n[0], hx, fmaT((R)q.n[1], hy, (R)q.n[2] * hz));
  const R dNr = fmaT((R)q.nd[0], hx, fmaT((R)q.nd[1], hy, (R)q.nd[2] * hz));
  const R Mr = Vr * inv_a0;
  const R dMr = dVr * inv_a0;
  const R om = (R)1 - Mr;
  const R iom = rcpT(om);
  const R T1 = D1 * ir * iom * iom;
  const R T2 = T1 * iom;
  const R T3 = T1 * ir;
  const R T4 = T3 * iom;
  const R MmM = Mr - (R)q.MM;
  const R ds = (R)q.ds, P = (R)q.P, dP = (R)q.dP, Vn = (R)q.Vn;
  *pth = rhoRef * Vn * ds * a0 * fmaT(dMr, T2, a0 * MmM * T4);
  const R Lr = P * Nr;
  const R dLr = fmaT(dP, Nr, P * dNr);
  const R Lm = P * Vn * inv_a0;
  R s = dLr * T1;
  s = fmaT(Lr * dMr, T2, s);
  s = fmaT(a0 * (Lr - Lm), T3, s);
  s = fmaT(a0 * Lr * MmM, T4, s);
  *plo = ds * s;
  *bad = !(r > (R)0) || !(om > (R)0);
}

template <typename Rec>
__device__ __forceinline__ double tau_at(const TimeConst& k, const Rec& q, int j, double ox,
                                         double oy, double oz, double* r_out) {
  const double r = retarded_r(k, ox - (double)q.x[0], oy - (double)q.x[1], oz - (double)q.x[2]);
  *r_out = r;
  return obs_tau(k, j, r);
}

// ---------------------------------------------------------------------------------------------
// Eval + binning.  One wave per workgroup, persistent over a queue of items.
template <bool MIXED, int W>
__global__ __launch_bounds__(64) void tonal_eval_kernel(const TonalCaseDev* __restrict__ cases,
                                                        const Tile* __restrict__ tiles,
                                                        const Item* __restrict__ items, int nItems,
                                                        int* __restrict__ counter,
                                                        uint32_t* __restrict__ flags_out) {
  using R = typename Prec<MIXED>::R;
  using Rec = typename Prec<MIXED>::Rec;
  using R2 = typename Prec<MIXED>::R2;
  // win[slot][lane] = (thickness, loading) partial of one observer-time bin; each lane owns its
  // column, so plain read-modify-writes are race free and keep a fixed order
  __shared__ R2 win[W * 64];
  const int lane = threadIdx.x;
  for (int s = 0; s < W; s++) win[s * 64 + lane] = R2{(R)0, (R)0};
  uint32_t flags = 0;
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(counter, 1);
    it = __builtin_amdgcn_readfirstlane(it);  // lane 0's value, wave-uniform
    if (it >= nItems) break;
    const Item item = items[it];
    const TonalCaseDev* Cp = cases + item.icase;
    const TimeConst k = Cp->k;
    const int nT = k.nT, nObs = Cp->nObs;
    const int o = item.obs0 + lane;
    const bool active = o < nObs;
    const int oc = active ? o : nObs - 1;
    const double ox = Cp->obs[3 * oc], oy = Cp->obs[3 * oc + 1], oz = Cp->obs[3 * oc + 2];
    const R oxr = (R)ox, oyr = (R)oy, ozr = (R)oz;
    R* slab = (R*)Cp->slab + ((int64_t)item.group * nObs + oc) * (int64_t)nT * 2;
    cptr<Rec> rec = (cptr<Rec>)(uintptr_t)Cp->rec;
    const R inv_a0 = (R)k.inv_a0, a0 = (R)k.a0, vx = (R)k.vx, vy = (R)k.vy, vz = (R)k.vz;
    const R rhoRef = (R)Cp->rhoRef, D1 = (R)Cp->D1;
    const double dnT = (double)nT;
    for (int ti = item.tile0; ti < item.tile0 + item.ntile; ti++) {
      const Tile tl = tiles[ti];
      for (int k0 = 0; k0 < nT; k0 += tl.K) {
        const int kend = min(k0 + tl.K, nT);
        const int km0 = (k0 > 0) ? k0 - 1 : nT - 1;
        int base;
        {
          double r;
          double tb = tau_at(k, load_rec<Rec>(rec + tl.e0 * nT + km0), km0, ox, oy, oz, &r);
          if (k0 == 0) tb = tb - dnT;
          base = (int)floor(tb) - tl.spread - 3;
        }
        int smin = W, smax = -1;
        for (int64_t e = tl.e0; e < tl.e0 + tl.ne; e++) {
          cptr<Rec> re = rec + e * nT;
          double r_cur, r_tmp;
          double t_prev = tau_at(k, load_rec<Rec>(re + km0), km0, ox, oy, oz, &r_tmp);
          if (k0 == 0) t_prev = t_prev - dnT;
          // two-deep record pipeline: qc = step kk (integrand), qn = step kk+1 (retarded solve),
          // qnn = step kk+2 in flight
          Rec qc = load_rec<Rec>(re + k0);
          const int kn0 = (k0 + 1 < nT) ? k0 + 1 : 0;
          Rec qn = load_rec<Rec>(re + kn0);
          double t_cur = tau_at(k, qc, k0, ox, oy, oz, &r_cur);
          for (int kk = k0; kk < kend; kk++) {
            const int kn = (kk + 1 < nT) ? kk + 1 : 0;
            const int knn = (kn + 1 < nT) ? kn + 1 : 0;
            const Rec qnn = load_rec<Rec>(re + knn);
            // the rising side of this eval's tent is known before the solve: start the window
            // reads of its footprint now
            const double fc = floor(t_cur);
            int sb = (int)fc - 1 - base;
            const bool inrange = sb >= 0 && sb + 3 < W;
            sb = inrange ? sb : (sb < 0 ? 0 : W - 4);
            R2* p = &win[sb * 64 + lane];
            R2 v0 = p[0], v1 = p[64], v2 = p[128], v3 = p[192];
            R pth, plo;
            bool bad;
            integrand<R, Rec>(qc, oxr, oyr, ozr, (R)r_cur, inv_a0, a0, vx, vy, vz, rhoRef, D1,
                              &pth, &plo, &bad);
            const double r_next =
                retarded_r(k, ox - (double)qn.x[0], oy - (double)qn.x[1], oz - (double)qn.x[2]);
            double t_next = obs_tau(k, kn, r_next);
            if (kk == nT - 1) t_next = t_next + dnT;
            Tent T;
            const bool ok = tent_of(t_prev, t_cur, t_next, &T) && !bad;
            if (!ok) flags |= kFlagNumeric;
            const R pthv = ok ? pth : (R)0, plov = ok ? plo : (R)0;
            const R idm = rcpT((R)(t_cur - t_prev));
            const R idp = rcpT((R)(t_next - t_cur));
            // footprint slots ic-1 .. ic+2: rising bins ic-1 (nd >= 2), ic (nd >= 1);
            // falling bins ic+1 (nu >= 1), ic+2 (nu >= 2)
            const R w0 = T.nd >= 2 ? (R)((fc - 1.0) - t_prev) * idm : (R)0;
            const R w1 = T.nd >= 1 ? (R)(fc - t_prev) * idm : (R)0;
            const R w2 = T.nu >= 1 ? (R)(t_next - (fc + 1.0)) * idp : (R)0;
            const R w3 = T.nu >= 2 ? (R)(t_next - (fc + 2.0)) * idp : (R)0;
            const bool any_on = (T.nd | T.nu) != 0;
            if (any_on && !inrange) flags |= kFlagWindow;
            v0.x = fmaT(pthv, w0, v0.x);
            v0.y = fmaT(plov, w0, v0.y);
            v1.x = fmaT(pthv, w1, v1.x);
            v1.y = fmaT(plov, w1, v1.y);
            v2.x = fmaT(pthv, w2, v2.x);
            v2.y = fmaT(plov, w2, v2.y);
            v3.x = fmaT(pthv, w3, v3.x);
            v3.y = fmaT(plov, w3, v3.y);
            p[0] = v0;
            p[64] = v1;
            p[128] = v2;
            p[192] = v3;
            if (any_on) {
              smin = min(smin, sb);
              smax = max(smax, sb + 3);
            }
            const bool tail = T.nd > 2 || T.nu > 2;
            if (wave_any(tail)) {  // tents wider than two bins a side
              if (tail) {
                if (T.nd > W || T.nu > W) flags |= kFlagSweep;
                const int nd = min(T.nd, W), nu = min(T.nu, W);
                auto put = [&](int b, R w) {
                  const int slot = b - base;
                  if (slot < 0 || slot >= W) {
                    flags |= kFlagWindow;
                    return;
                  }
                  R2 v = win[slot * 64 + lane];
                  v.x = fmaT(pthv, w, v.x);
                  v.y = fmaT(plov, w, v.y);
                  win[slot * 64 + lane] = v;
                  smin = min(smin, slot);
                  smax = max(smax, slot);
                };
                for (int j = 2; j < nd; j++) put(T.ic - j, (R)((fc - (double)j) - t_prev) * idm);
                for (int j = 2; j < nu; j++) put(T.ic + 1 + j, (R)(t_next - (fc + (double)(1 + j))) * idp);
              }
            }
            t_prev = t_cur;
            t_cur = t_next;
            r_cur = r_next;
            qc = qn;
            qn = qnn;
          }
        }
        // add the window into this item's slab row and clear it
        for (int s = smin; s <= smax; s++) {
          const R2 v = win[s * 64 + lane];
          win[s * 64 + lane] = R2{(R)0, (R)0};
          if (active) {
            int b = (base + s) % nT;
            if (b < 0) b += nT;
            slab[2 * b] += v.x;
            slab[2 * b + 1] += v.y;
          }
        }
      }
    }
  }
  if (flags) atomicOr(flags_out, flags);
}

// Ordered reduction of the partial slabs: out(b, o) = sum_g slab[g][o][b], FP64 Kahan.
// Block 32 x 8 covers 32 bins x 32 observers; LDS transposes to the time-major outputs.
template <typename A>
__global__ __launch_bounds__(256) void tonal_reduce_kernel(const TonalCaseDev* __restrict__ cases,
                                                           int icase) {
  const TonalCaseDev& C = cases[icase];
  const int nT = C.k.nT, nObs = C.nObs, G = C.G;
  const A* slab = (const A*)C.slab;
  const int b0 = blockIdx.x * 32, o0 = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;
  __shared__ float sT[32][33], sL[32][33], sA[32][33];
  for (int i = 0; i < 4; i++) {
    const int ol = ty + 8 * i, o = o0 + ol, b = b0 + tx;
    if (o < nObs && b < nT) {
      double sT_ = 0.0, cT = 0.0, sL_ = 0.0, cL = 0.0;
      const A* p = slab + ((int64_t)o * nT + b) * 2;
      const int64_t gstride = (int64_t)nObs * nT * 2;
      for (int g = 0; g < G; g++) {
        const double xT = (double)p[0], xL = (double)p[1];
        double y = xT - cT, t = sT_ + y;
        cT = (t - sT_) - y;
        sT_ = t;
        y = xL - cL;
        t = sL_ + y;
        cL = (t - sL_) - y;
        sL_ = t;
        p += gstride;
      }
      sT[ol][tx] = (float)sT_;
      sL[ol][tx] = (float)sL_;
      sA[ol][tx] = (float)(sT_ + sL_);
    }
  }
  __syncthreads();
  for (int i = 0; i < 4; i++) {
    const int bl = ty + 8 * i, b = b0 + bl, o = o0 + tx;
    if (b < nT && o < nObs) {
      C.pT[(int64_t)b * nObs + o] = sT[tx][bl];
      C.pL[(int64_t)b * nObs + o] = sL[tx][bl];
      C.pA[(int64_t)b * nObs + o] = sA[tx][bl];
    }
  }
}

// Reference AoS (X[4], N[6], V[6], PR[2] per element-step, surfaces concatenated) -> records.
template <typename Rec, typename H>
__global__ void tonal_prep_kernel(const float* __restrict__ X, const float* __restrict__ N,
                                  const float* __restrict__ V, const float* __restrict__ PR,
                                  int64_t n, double inv_a0, Rec* __restrict__ rec) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const float* x = X + 4 * q;
    const float* nn = N + 6 * q;
    const float* v = V + 6 * q;
    const float* p = PR + 2 * q;
    Rec o;
    for (int i = 0; i < 3; i++) {
      o.x[i] = x[i];
      o.v[i] = v[i];
      o.vd[i] = v[3 + i];
      o.n[i] = nn[i];
      o.nd[i] = nn[3 + i];
    }
    o.ds = x[3];
    o.P = p[0];
    o.dP = p[1];
    const double vv = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
    const double M = sqrt(vv) * inv_a0;
    o.MM = (H)(M * M);
    o.Vn = (H)((double)v[0] * nn[0] + (double)v[1] * nn[1] + (double)v[2] * nn[2]);
    rec[q] = o;
  }
}

template <typename T>
__global__ void fill_zero_kernel(T* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (T)0;
}

__global__ void reset_counter_kernel(int* c, uint32_t* flags) {
  *c = 0;
  *flags = 0u;
}

// Tent of every (element, step) for one observer, through the same tent code: the top rising
// bin ic (stored as ic mod nT, with its period index floor(ic / nT)) and the two side lengths.
template <typename Rec>
__global__ void tonal_bins_kernel(const Rec* __restrict__ rec, int64_t nE, TimeConst k, double ox,
                                  double oy, double oz, int32_t* ipts, int32_t* npi, int32_t* nd,
                                  int32_t* nu) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nE) return;
  const int nT = k.nT;
  const Rec* re = rec + e * nT;
  double r;
  double t_prev = tau_at(k, re[nT - 1], nT - 1, ox, oy, oz, &r) - (double)nT;
  double t_cur = tau_at(k, re[0], 0, ox, oy, oz, &r);
  for (int kk = 0; kk < nT; kk++) {
    double t_next = (kk + 1 < nT) ? tau_at(k, re[kk + 1], kk + 1, ox, oy, oz, &r)
                                  : tau_at(k, re[0], 0, ox, oy, oz, &r) + (double)nT;
    const int64_t idx = e * nT + kk;
    Tent T;
    if (tent_of(t_prev, t_cur, t_next, &T)) {
      int per = T.ic / nT, ph = T.ic % nT;
      if (ph < 0) {
        ph += nT;
        per -= 1;
      }
      ipts[idx] = ph;
      npi[idx] = per;
      nd[idx] = T.nd;
      nu[idx] = T.nu;
    } else {
      ipts[idx] = -1;
      npi[idx] = 0;
      nd[idx] = -1;
      nu[idx] = -1;
    }
    t_prev = t_cur;
    t_cur = t_next;
  }
}

// ---------------------------------------------------------------------------------------------
// Host-side launchers (called from the C-ABI layer).
hipError_t launch_tonal_eval(bool mixed, int W, int grid, hipStream_t s, const TonalCaseDev* cases,
                             const Tile* tiles, const Item* items, int nItems, int* counter,
                             uint32_t* flags) {
  dim3 b(64);
#define ACUM_EVAL(M, WW)                                                                     \
  if (mixed == M && W == WW) {                                                               \
    hipLaunchKernelGGL((tonal_eval_kernel<M, WW>), dim3(grid), b, 0, s, cases, tiles, items, \
                       nItems, counter, flags);                                              \
    return hipGetLastError();                                                                \
  }
  ACUM_EVAL(true, 32)
  ACUM_EVAL(true, 48)
  ACUM_EVAL(true, 64)
  ACUM_EVAL(true, 128)
  ACUM_EVAL(false, 32)
  ACUM_EVAL(false, 64)
#undef ACUM_EVAL
  return hipErrorInvalidValue;
}

hipError_t eval_occupancy(bool mixed, int W, int* blocks_per_cu) {
#define ACUM_OCC(M, WW)                                                                      \
  if (mixed == M && W == WW)                                                                 \
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,                       \
                                                        tonal_eval_kernel<M, WW>, 64, 0);
  ACUM_OCC(true, 32)
  ACUM_OCC(true, 48)
  ACUM_OCC(true, 64)
  ACUM_OCC(true, 128)
  ACUM_OCC(false, 32)
  ACUM_OCC(false, 64)
#undef ACUM_OCC
  return hipErrorInvalidValue;
}

hipError_t launch_tonal_reduce(bool mixed, hipStream_t s, const TonalCaseDev* cases, int icase,
                               int nT, int nObs) {
  dim3 g((nT + 31) / 32, (nObs + 31) / 32), b(32, 8);
  if (mixed)
    hipLaunchKernelGGL(tonal_reduce_kernel<float>, g, b, 0, s, cases, icase);
  else
    hipLaunchKernelGGL(tonal_reduce_kernel<double>, g, b, 0, s, cases, icase);
  return hipGetLastError();
}

hipError_t launch_tonal_prep(bool mixed, hipStream_t s, const float* X, const float* N,
                             const float* V, const float* PR, int64_t n, double inv_a0, void* rec) {
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
  if (mixed)
    hipLaunchKernelGGL((tonal_prep_kernel<Rec32, float>), dim3(grid), dim3(256), 0, s, X, N, V, PR, n,
                       inv_a0, (Rec32*)rec);
  else
    hipLaunchKernelGGL((tonal_prep_kernel<Rec64, double>), dim3(grid), dim3(256), 0, s, X, N, V, PR, n,
                       inv_a0, (Rec64*)rec);
  return hipGetLastError();
}

hipError_t launch_fill_zero(hipStream_t s, void* p, int64_t bytes) {
  const int64_t n = bytes / 16;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
  if (n > 0) hipLaunchKernelGGL(fill_zero_kernel<float4>, dim3(grid), dim3(256), 0, s, (float4*)p, n);
  const int64_t rem = bytes - n * 16;
  if (rem > 0)
    hipLaunchKernelGGL(fill_zero_kernel<unsigned char>, dim3(1), dim3(64), 0, s,
                       (unsigned char*)p + n * 16, rem);
  return hipGetLastError();
}

hipError_t launch_reset(hipStream_t s, int* counter, uint32_t* flags) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, counter, flags);
  return hipGetLastError();
}

hipError_t launch_tonal_bins(bool mixed, hipStream_t s, const void* rec, int64_t nE, const TimeConst& k,
                             double ox, double oy, double oz, int32_t* ipts, int32_t* npi,
                             int32_t* nd, int32_t* nu) {
  const int grid = (int)((nE + 63) / 64);
  if (mixed)
    hipLaunchKernelGGL(tonal_bins_kernel<Rec32>, dim3(grid), dim3(64), 0, s, (const Rec32*)rec, nE, k,
                       ox, oy, oz, ipts, npi, nd, nu);
  else
    hipLaunchKernelGGL(tonal_bins_kernel<Rec64>, dim3(grid), dim3(64), 0, s, (const Rec64*)rec, nE,
                       k, ox, oy, oz, ipts, npi, nd, nu);
  return hipGetLastError();
}

}  // namespace acum
