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
// Canonical FP64 observer time (acous_dist core.cu:261-279, ttemp core.cu:156).
__device__ __forceinline__ double retarded_r(const TimeConst& k, double dx, double dy, double dz) {
  double c = dx * dx;
  c = c + dy * dy;
  c = c + dz * dz;
  double b = 0.0;
  if (!k.still) {
    double bh = dx * k.vx;
    bh = bh + dy * k.vy;
    bh = bh + dz * k.vz;
    b = -2.0 * bh;
  }
  const double disc = b * b - k.four_a * c;
  const double sq = sqrt(disc);
  double t = (-b - sq) * k.inv_2a;
  if (t < 0.0) t = (-b + sq) * k.inv_2a;
  return t * k.a0;
}

__device__ __forceinline__ double obs_time(const TimeConst& k, int j, double r) {
  return ((double)j * k.dTau + r * k.inv_a0) * k.oM;
}

// Tent binning of one eval (findpts/bsearch core.cu:290-355, sweeps core.cu:196-231).
// Calls add(absolute_bin, t_num, t_den_inverse_index) for every bin: absolute bin a = shifted
// index - npi*nT, so a*dPsi is the unshifted observer time of the bin.  Returns the lookup
// (ipts, npi) and sweep lengths for diagnostics.  Weight of a bin is num/den with
//   down-sweep: num = ttem_s - tObs2, den = ttem - ttemp
//   up-sweep:   num = ttep_s - tObs2, den = ttep - ttemp
template <typename AddDown, typename AddUp>
__device__ __forceinline__ bool bin_sweeps(const TimeConst& k, double ttem, double ttemp,
                                           double ttep, int maxbins, int* ipts_o, int* npi_o,
                                           int* nd_o, int* nu_o, AddDown addd, AddUp addu) {
  const double q = floor((k.lpi - ttemp) * k.inv_lpi);
  if (!(fabs(q) < 1.0e9)) return false;  // non-finite observer time
  int npi = (int)q;
  double shift = (double)npi * k.lpi;
  double newx = ttemp + shift;
  if (newx < 0.0) {
    npi += 1;
    shift = (double)npi * k.lpi;
    newx = ttemp + shift;
  }
  const int nT = k.nT;
  double f = floor(newx * k.inv_dPsi);
  int ip = (f < 0.0) ? 0 : (f > (double)(nT - 1) ? nT - 1 : (int)f);
  while (ip + 1 < nT && (double)(ip + 1) * k.dPsi <= newx) ip++;
  while (ip > 0 && (double)ip * k.dPsi > newx) ip--;
  const int aoff = -npi * nT;
  // down-sweep
  const double ttem_s = ttem + shift;
  double tObs2 = (double)ip * k.dPsi;
  int sidx = ip, cnt = 0;
  bool ok = true;
  while (tObs2 > ttem_s) {
    addd(sidx + aoff, ttem_s - tObs2);
    tObs2 -= k.dPsi;
    sidx -= 1;
    if (++cnt > maxbins) {
      ok = false;
      break;
    }
  }
  *nd_o = cnt;
  // up-sweep
  const int iptp = (ip < nT - 1) ? ip + 1 : 0;
  tObs2 = (double)iptp * k.dPsi;
  if (iptp == 0) tObs2 += k.lpi;
  sidx = ip + 1;
  const double ttep_s = ttep + shift;
  cnt = 0;
  while (ok && tObs2 < ttep_s) {
    addu(sidx + aoff, ttep_s - tObs2);
    tObs2 += k.dPsi;
    sidx += 1;
    if (++cnt > maxbins) {
      ok = false;
      break;
    }
  }
  *nu_o = cnt;
  *ipts_o = ip;
  *npi_o = npi;
  return ok;
}

// ---------------------------------------------------------------------------------------------
// Arithmetic helpers per precision.
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float rcpT(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcpT(double x) { return 1.0 / x; }

// Farassat 1A integrand of the impermeable surface (core.cu:68-154): rho = 1, U = V, the dotted
// normal velocities forced to zero (core.cu:117-118), so
//   pThick = rhoRef Vn dS a0 (dMr T2 + a0 (Mr - M^2) T4)
//   pLoad  = dS (dLr T1 + Lr dMr T2 + a0 (Lr - Lm) T3 + a0 Lr (Mr - M^2) T4)
// with T1 = 1/(4 pi a0 r (1-Mr)^2), T2 = T1/(1-Mr), T3 = T1/r, T4 = T3/(1-Mr).
template <typename R>
__device__ __forceinline__ void integrand(const R* __restrict__ q, R dx, R dy, R dz, R r,
                                          R inv_a0, R a0, R vx, R vy, R vz, R rhoRef, R D1,
                                          R* pth, R* plo, bool* bad) {
  const R rt = r * inv_a0;
  const R rvx = fmaT(-vx, rt, dx), rvy = fmaT(-vy, rt, dy), rvz = fmaT(-vz, rt, dz);
  const R ir = rcpT(r);
  const R hx = rvx * ir, hy = rvy * ir, hz = rvz * ir;
  const R Vr = fmaT(q[4], hx, fmaT(q[5], hy, q[6] * hz));
  const R dVr = fmaT(q[7], hx, fmaT(q[8], hy, q[9] * hz));
  const R Nr = fmaT(q[10], hx, fmaT(q[11], hy, q[12] * hz));
  const R dNr = fmaT(q[13], hx, fmaT(q[14], hy, q[15] * hz));
  const R Mr = Vr * inv_a0;
  const R dMr = dVr * inv_a0;
  const R om = (R)1 - Mr;
  const R iom = rcpT(om);
  const R T1 = D1 * ir * iom * iom;
  const R T2 = T1 * iom;
  const R T3 = T1 * ir;
  const R T4 = T3 * iom;
  const R MmM = Mr - q[18];
  const R ds = q[3], P = q[16], dP = q[17], Vn = q[19];
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

template <bool MIXED>
struct Prec {
  using R = double;
};
template <>
struct Prec<true> {
  using R = float;
};

// ---------------------------------------------------------------------------------------------
// Eval + binning.  One wave per workgroup, persistent over a queue of items.
template <bool MIXED, int W>
__global__ __launch_bounds__(64) void tonal_eval_kernel(const TonalCaseDev* __restrict__ cases,
                                                        const Tile* __restrict__ tiles,
                                                        const Item* __restrict__ items, int nItems,
                                                        int* __restrict__ counter,
                                                        uint32_t* __restrict__ flags_out) {
  using R = typename Prec<MIXED>::R;
  __shared__ R winT[W * 64];
  __shared__ R winL[W * 64];
  const int lane = threadIdx.x;
  for (int s = 0; s < W; s++) {
    winT[s * 64 + lane] = (R)0;
    winL[s * 64 + lane] = (R)0;
  }
  uint32_t flags = 0;
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(counter, 1);
    it = __shfl(it, 0);
    if (it >= nItems) break;
    const Item item = items[it];
    const TonalCaseDev& C = cases[item.icase];
    const TimeConst k = C.k;
    const int nT = k.nT;
    const int o = item.obs0 + lane;
    const bool active = o < C.nObs;
    const int oc = active ? o : C.nObs - 1;
    const double ox = C.obs[3 * oc], oy = C.obs[3 * oc + 1], oz = C.obs[3 * oc + 2];
    R* slab = (R*)C.slab + ((int64_t)item.group * C.nObs + oc) * (int64_t)nT * 2;
    const R* rec = (const R*)C.rec;
    const R inv_a0 = (R)k.inv_a0, a0 = (R)k.a0, vx = (R)k.vx, vy = (R)k.vy, vz = (R)k.vz;
    const R rhoRef = (R)C.rhoRef, D1 = (R)C.D1;
    for (int ti = item.tile0; ti < item.tile0 + item.ntile; ti++) {
      const Tile tl = tiles[ti];
      for (int k0 = 0; k0 < nT; k0 += tl.K) {
        const int kend = min(k0 + tl.K, nT);
        const int km0 = (k0 > 0) ? k0 - 1 : nT - 1;
        int base;
        {
          const R* q = rec + (tl.e0 * nT + km0) * kRecFields;
          const double r = retarded_r(k, ox - (double)q[0], oy - (double)q[1], oz - (double)q[2]);
          double tb = obs_time(k, km0, r);
          if (k0 == 0) tb = tb - k.lpi;
          base = (int)floor(tb * k.inv_dPsi) - tl.spread - 2;
        }
        int smin = W, smax = -1;
        for (int64_t e = tl.e0; e < tl.e0 + tl.ne; e++) {
          const R* re = rec + e * nT * kRecFields;
          double t_prev, t_cur, r_cur, dxc, dyc, dzc;
          {
            const R* q = re + km0 * kRecFields;
            const double r = retarded_r(k, ox - (double)q[0], oy - (double)q[1], oz - (double)q[2]);
            t_prev = obs_time(k, km0, r);
            if (k0 == 0) t_prev = t_prev - k.lpi;
            const R* q0 = re + k0 * kRecFields;
            dxc = ox - (double)q0[0];
            dyc = oy - (double)q0[1];
            dzc = oz - (double)q0[2];
            r_cur = retarded_r(k, dxc, dyc, dzc);
            t_cur = obs_time(k, k0, r_cur);
          }
          for (int kk = k0; kk < kend; kk++) {
            const int kn = (kk + 1 < nT) ? kk + 1 : 0;
            const R* qn = re + kn * kRecFields;
            const double dxn = ox - (double)qn[0], dyn = oy - (double)qn[1], dzn = oz - (double)qn[2];
            const double r_next = retarded_r(k, dxn, dyn, dzn);
            double t_next = obs_time(k, kn, r_next);
            if (kk == nT - 1) t_next = t_next + k.lpi;
            R pth, plo;
            bool bad;
            integrand<R>(re + kk * kRecFields, (R)dxc, (R)dyc, (R)dzc, (R)r_cur, inv_a0, a0, vx, vy,
                         vz, rhoRef, D1, &pth, &plo, &bad);
            if (bad) flags |= kFlagNumeric;
            const R idm = rcpT((R)(t_prev - t_cur));
            const R idp = rcpT((R)(t_next - t_cur));
            int ip, np, nd, nu;
            auto put = [&](int a, R w) {
              const int slot = a - base;
              if (slot >= 0 && slot < W) {
                __hip_atomic_fetch_add(&winT[slot * 64 + lane], pth * w, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(&winL[slot * 64 + lane], plo * w, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                smin = min(smin, slot);
                smax = max(smax, slot);
              } else {
                flags |= kFlagWindow;
              }
            };
            const bool ok = bin_sweeps(
                k, t_prev, t_cur, t_next, W, &ip, &np, &nd, &nu,
                [&](int a, double num) { put(a, (R)num * idm); },
                [&](int a, double num) { put(a, (R)num * idp); });
            if (!ok) flags |= kFlagSweep | kFlagNumeric;
            t_prev = t_cur;
            t_cur = t_next;
            r_cur = r_next;
            dxc = dxn;
            dyc = dyn;
            dzc = dzn;
          }
        }
        // add the window into this item's slab row and clear it
        for (int s = smin; s <= smax; s++) {
          const R vT = winT[s * 64 + lane], vL = winL[s * 64 + lane];
          winT[s * 64 + lane] = (R)0;
          winL[s * 64 + lane] = (R)0;
          if (active) {
            int b = (base + s) % nT;
            if (b < 0) b += nT;
            slab[2 * b] += vT;
            slab[2 * b + 1] += vL;
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
template <typename R>
__global__ void tonal_prep_kernel(const float* __restrict__ X, const float* __restrict__ N,
                                  const float* __restrict__ V, const float* __restrict__ PR,
                                  int64_t n, double inv_a0, R* __restrict__ rec) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const float* x = X + 4 * q;
    const float* nn = N + 6 * q;
    const float* v = V + 6 * q;
    const float* p = PR + 2 * q;
    R* o = rec + q * kRecFields;
    for (int i = 0; i < 4; i++) o[i] = (R)x[i];
    for (int i = 0; i < 6; i++) o[4 + i] = (R)v[i];
    for (int i = 0; i < 6; i++) o[10 + i] = (R)nn[i];
    o[16] = (R)p[0];
    o[17] = (R)p[1];
    const double vv = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
    const double M = sqrt(vv) * inv_a0;
    o[18] = (R)(M * M);
    o[19] = (R)((double)v[0] * nn[0] + (double)v[1] * nn[1] + (double)v[2] * nn[2]);
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

// Bin lookup of every (element, step) for one observer, through the same sweep code.
template <typename R>
__global__ void tonal_bins_kernel(const R* __restrict__ rec, int64_t nE, TimeConst k, double ox,
                                  double oy, double oz, int32_t* ipts, int32_t* npi, int32_t* nd,
                                  int32_t* nu) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nE) return;
  const int nT = k.nT;
  const R* re = rec + e * nT * kRecFields;
  auto tob = [&](int j) {
    const R* q = re + j * kRecFields;
    return obs_time(k, j, retarded_r(k, ox - (double)q[0], oy - (double)q[1], oz - (double)q[2]));
  };
  double t_prev = tob(nT - 1) - k.lpi, t_cur = tob(0);
  for (int kk = 0; kk < nT; kk++) {
    double t_next = (kk + 1 < nT) ? tob(kk + 1) : tob(0) + k.lpi;
    int a, b, c, d;
    const bool ok = bin_sweeps(k, t_prev, t_cur, t_next, 1 << 20, &a, &b, &c, &d,
                               [](int, double) {}, [](int, double) {});
    const int64_t idx = e * nT + kk;
    ipts[idx] = ok ? a : -1;
    npi[idx] = b;
    nd[idx] = c;
    nu[idx] = d;
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
    hipLaunchKernelGGL(tonal_prep_kernel<float>, dim3(grid), dim3(256), 0, s, X, N, V, PR, n, inv_a0,
                       (float*)rec);
  else
    hipLaunchKernelGGL(tonal_prep_kernel<double>, dim3(grid), dim3(256), 0, s, X, N, V, PR, n, inv_a0,
                       (double*)rec);
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
    hipLaunchKernelGGL(tonal_bins_kernel<float>, dim3(grid), dim3(64), 0, s, (const float*)rec, nE, k,
                       ox, oy, oz, ipts, npi, nd, nu);
  else
    hipLaunchKernelGGL(tonal_bins_kernel<double>, dim3(grid), dim3(64), 0, s, (const double*)rec, nE,
                       k, ox, oy, oz, ipts, npi, nd, nu);
  return hipGetLastError();
}

}  // namespace acum
