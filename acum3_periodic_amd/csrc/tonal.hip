// Tonal Farassat 1A (thickness + loading) kernels for gfx950.
//
// Work decomposition (the reference launches one thread per (step, observer, element) with
// float atomics on pT/pL/pA, core.cu:18-259, main.cu:194-236):
//   * one wave = 64 observers (one per lane); element records are wave-uniform (positions through
//     scalar loads, attribute records spread over the lanes and broadcast by DPP), so every
//     record read from HBM serves 64 observers;
//   * a work item = (case, observer group, group of tiles), taken from one of eight queues (one
//     per XCD, so the observer groups reading the same records share an L2); a tile is a run of
//     elements that sit close together; for each tile the wave walks the source steps in passes
//     of K steps and, for each element of the tile, the K steps of that pass in order -- one FP64
//     retarded solve per (element, step, observer) instead of the reference's three
//     (core.cu:84,189,193);
//   * each lane owns a column of an LDS window of W observer-time bins; the tent-weighted
//     contributions of an eval (core.cu:196-231) are read-modify-written into the 4-slot
//     footprint ic-1 .. ic+2 around its top rising bin ic with weights clamped at zero (by the
//     FMA's clamp bit), rarer wider tents in a wave-uniform tail; no other lane touches the
//     column, so the order is fixed;
//   * after each pass the window is added into the item's row of a partial slab in HBM;
//   * a reduction kernel sums the partial slabs of each (observer, bin) in fixed order with
//     FP64 Kahan compensation and writes pT, pL and pA = pT + pL.
// Results are deterministic: no operation depends on scheduling order.
//
// Precision: MIXED computes observer time, bin index and the bin time deltas in FP64 and the
// geometry, Doppler factor, integrands and tent weights in FP32 (FP32 window and slab); the FP64
// variant does everything in FP64 (tonal_plain.hip keeps the reference's structure in FP64).
// The FP64 observer-time/bin code is the canonical form the CPU oracle uses, compiled without
// contraction, so bin indices agree bit for bit.
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <type_traits>
#include "dev_types.hpp"
#include "fp64_math.hpp"

namespace acum {

// Positions are read with wave-uniform addresses through constant-address-space pointers, as
// scalar loads into SGPRs; through global-address-space pointers they are vector loads (the next
// element's pass-start positions).  Attribute records reach the lanes through per-lane buffer
// loads (the mixed path's row per lane, the FP64 path's field per lane) and DPP broadcasts.
template <typename T>
using cptr = const __attribute__((address_space(1))) T*;
template <typename T>
using kptr = const __attribute__((address_space(4))) T*;

template <typename T>
__device__ __forceinline__ T load_words(cptr<T> p) {
  static_assert(sizeof(T) % 4 == 0, "record must be whole dwords");
  constexpr int n = (int)(sizeof(T) / 4);
  cptr<uint32_t> w = (cptr<uint32_t>)p;
  uint32_t t[n];
#pragma unroll
  for (int i = 0; i < n; i++) t[i] = w[i];
  T r;
  __builtin_memcpy(&r, t, sizeof(T));
  return r;
}

template <typename T>
__device__ __forceinline__ T load_words(kptr<T> p) {
  static_assert(sizeof(T) % 4 == 0, "record must be whole dwords");
  constexpr int n = (int)(sizeof(T) / 4);
  kptr<uint32_t> w = (kptr<uint32_t>)p;
  uint32_t t[n];
#pragma unroll
  for (int i = 0; i < n; i++) t[i] = w[i];
  T r;
  __builtin_memcpy(&r, t, sizeof(T));
  return r;
}

// ---------------------------------------------------------------------------------------------
// Canonical FP64 observer time in bins (acous_dist core.cu:261-279, ttemp core.cu:156; the form
// is described with TimeConst); the CPU oracle evaluates the same operations in the same order.
// `delay` returns T, the delay in bins.  EXACT selects the correctly rounded square root (the FP64
// path and the diagnostic bin lookup); the mixed path's weights are FP32, and there one residual
// correction of the root suffices.
struct LaneObs {
  double ox, oy, oz;  // observer - O
  double qo, wo;      // C2 |o|^2, w.o
};

__device__ __forceinline__ LaneObs lane_obs(const TimeConst& k, double obx, double oby, double obz) {
  LaneObs L;
  L.ox = obx - k.Ox;
  L.oy = oby - k.Oy;
  L.oz = obz - k.Oz;
  L.qo = __builtin_fma(L.oz, L.oz, __builtin_fma(L.oy, L.oy, L.ox * L.ox)) * k.C2;
  L.wo = __builtin_fma(L.oz, k.wz, __builtin_fma(L.oy, k.wy, L.ox * k.wx));
  return L;
}

template <bool MOVING, bool EXACT, typename P>
__device__ __forceinline__ double obs_time(const LaneObs& L, const P& p, double* delay) {
  const double cc = __builtin_fma(L.ox, p.mx, __builtin_fma(L.oy, p.my, __builtin_fma(L.oz, p.mz, L.qo + p.q)));
  double T;
  if (!MOVING) {
    T = EXACT ? sqrt_pos(cc) : sqrt_pos_fast(cc);
  } else {
    const double b = L.wo - p.wp;
    const double x = __builtin_fma(b, b, cc);
    T = (EXACT ? sqrt_pos(x) : sqrt_pos_fast(x)) - b;
  }
  *delay = T;
  return T + p.jt;
}

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0; }

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float rcpT(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcpT(double x) { return rcp_nr(x); }
// clamp into [0, 1]: folds into the clamp bit of the instruction that produces x
__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }
__device__ __forceinline__ double clamp01(double x) { return __builtin_fmin(__builtin_fmax(x, 0.0), 1.0); }
__device__ __forceinline__ float fractT(double x, float) { return (float)__builtin_amdgcn_fract(x); }
__device__ __forceinline__ double fractT(double x, double) { return __builtin_amdgcn_fract(x); }

template <bool MIXED>
struct Prec {
  using R = double;
  using R2 = double2;
  template <typename T>
  using ptr = cptr<T>;
};
template <>
struct Prec<true> {
  using R = float;
  using R2 = float2;
  template <typename T>
  using ptr = kptr<T>;
};

// Farassat 1A integrand of the impermeable surface (core.cu:68-154): rho = 1, U = V and the
// dotted normal velocities forced to zero (core.cu:117-118), so
//   pThick = rhoRef Vn dS a0 (dMr T2 + a0 (Mr - M^2) T4)
//   pLoad  = dS (dLr T1 + Lr dMr T2 + a0 (Lr - L_M) T3 + a0 Lr (Mr - M^2) T4)
// with T1 = D1/(r (1-Mr)^2), T2 = T1/(1-Mr), T3 = T1/r, T4 = T3/(1-Mr).  With the record
// products (AttrT) dotted into r_vec -- Mr = dV/r, dMr = dVd/r, Lr = dL/(cL r), dLr = dLd/(cL r)
// for dV = v.r_vec, dVd = vd.r_vec, dL = l.r_vec, dLd = ld.r_vec (l, ld carry the factor cL) --
// and g = 1/(r (1 - Mr)) = 1/(r - dV), every 1/r cancels:
//   pThick = cT g^3 (r (dVd - a0 M^2) + a0 dV) = cT g^2 (r t - a0)
//   pLoad  = g^2 (dL g (dVd - a0 M^2 + a0) + dLd - cL a0 L_M) = g^2 (dL t + dLd - aLM)
// with w = dVd - a0 (M^2 - 1) and t = g w (using 1 + g dV = r g and g (r - dV) = 1), one
// reciprocal per eval.  d is obs - src and `del` the delay T in bins: r = aB T and r_vec = d - u T
// (core.cu:86; aB = a0 / KB, u = vinf / KB; u = 0 in still air).  The records are stored scaled by
// s = 1 / aB (v, vd, aM1 and cT by s; l, ld and aLM by s^2), which makes r = T, a0 = KB and
// g = 1 / (T - dV) = aB / (r (1 - Mr)) in the same formulas with the same pThick and pLoad, and
// saves the product aB T.  Returns (T - dV) = r (1 - Mr) / aB: positive for a valid eval, zero
// for a source on the observer (T = 0 gives r_vec = 0 and dV = 0) and negative for Mr > 1.
// r_vec of an eval
template <typename R, bool MOVING>
__device__ __forceinline__ void radiation_vector(R dx, R dy, R dz, R del, R vx, R vy, R vz, R* rx,
                                                 R* ry, R* rz) {
  if (MOVING) {
    *rx = fmaT(-vx, del, dx);
    *ry = fmaT(-vy, del, dy);
    *rz = fmaT(-vz, del, dz);
  } else {
    *rx = dx;
    *ry = dy;
    *rz = dz;
  }
}

// the integrand from the delay T and the four record products (scaled records, a0 -> KB)
template <typename R>
__device__ __forceinline__ R integrand_tail(R aM1, R aLM, R cT, R del, R dV, R dVd, R dL, R dLd, R KB,
                                            R* pth, R* plo) {
  const R rm = del - dV;              // r (1 - Mr) / aB
  const R g = rcpT(rm);
  const R g2 = g * g;
  const R t = g * (dVd - aM1);        // g w, shared by both terms
  *pth = (cT * g2) * fmaT(del, t, -KB);
  *plo = g2 * fmaT(dL, t, dLd - aLM);
  return rm;
}

// FP64 path: what a step reads of its attribute record.  The sixteen doubles of the record
// (v[3], vd[3], l[3], ld[3], aM1, aLM, cT, unused) reach the lanes through one 8-byte vector load
// per lane: lane l holds field l % 16, so each 16-lane row of the wave holds the whole record, and
// a step reads field j from lane j of the lane's own row by DPP (row_newbcast:j, the one DPP
// control of FP64 instructions), on the operand of v_fmac_f64 for the record products and by
// v_mov_b64 for the three scalars (rec_products64).  Read as rows through wave-uniform vector
// loads (six 16-byte loads per step, each a 64-lane memory instruction) the record bound this
// path (this form is 7% faster), and through scalar loads it would not fit the SGPR file.
static_assert(offsetof(AttrT<double>, aM1) == 12 * sizeof(double), "FP64 attribute record layout");
static_assert(sizeof(AttrT<double>) == 16 * sizeof(double), "FP64 attribute record layout");

// integrand_tail's arithmetic with rm = del - dV out of the products, and -(dVd - aM1) and
// -(dLd - aLM) too, whose signs fold into the instructions that use them
template <bool MOVING>
__device__ __forceinline__ double integrand(double f, double dx, double dy, double dz, double del, double KB,
                                            double vx, double vy, double vz, double* pth, double* plo) {
  double rx, ry, rz;
  radiation_vector<double, MOVING>(dx, dy, dz, del, vx, vy, vz, &rx, &ry, &rz);
  double rm = del, nVd, dL = 0.0, nLd, cT;
  rec_products64(f, rx, ry, rz, rm, nVd, dL, nLd, cT);
  const double g = rcpT(rm);
  const double g2 = g * g;
  const double tn = g * nVd;  // -g w
  *pth = (cT * g2) * fmaT(-del, tn, -KB);
  *plo = g2 * -fmaT(dL, tn, nLd);
  return rm;
}

// Mixed path: the four record products of 64 observers by the matrix cores.  With
// v_mfma_f32_4x4x1f32 (16 blocks of four lanes, K = 1) lane l = 4b + j supplies row j of block b's
// A and column j of its B and receives column j of block b's result: so lane l supplies row
// l % 4 of the 4 x 3 record matrix (v, vd, l, ld) for A, its observer's r_vec component for B,
// and after three accumulating products (x, y, z) holds dV, dVd, dL, dLd of its own observer.
// The record reaches the lanes by one 16-byte vector load per lane: the mixed attribute record
// stores each row with one scalar field after it ({v, aM1}, {vd, aLM}, {l, cT}, {ld, 0}), lane l
// loads row l % 4, and the integrand reads aM1, aLM and cT from lanes 0, 1 and 2 of its block by
// DPP (quad_perm) on the instructions that use them.
typedef float v4f __attribute__((ext_vector_type(4)));
struct MixRec {
  float x, y, z, w;  // row l % 4 of the record and the scalar field stored after it
};
static_assert(sizeof(AttrT<float>) == 4 * sizeof(MixRec), "mixed attribute record layout");
// field w of lane J of the lane's block of four (J = 0, 1, 2: aM1, aLM, cT)
template <int J>
__device__ __forceinline__ float quad_field(float w) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(w), J * 0x55, 0xf, 0xf, true));
}

template <bool MOVING>
__device__ __forceinline__ float integrand_mfma(const MixRec& q, float dx, float dy, float dz, float del,
                                                float KB, float vx, float vy, float vz, float* pth,
                                                float* plo) {
  float rx, ry, rz;
  radiation_vector<float, MOVING>(dx, dy, dz, del, vx, vy, vz, &rx, &ry, &rz);
  v4f d = {0.0f, 0.0f, 0.0f, 0.0f};
  d = __builtin_amdgcn_mfma_f32_4x4x1f32(q.x, rx, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_4x4x1f32(q.y, ry, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_4x4x1f32(q.z, rz, d, 0, 0, 0);
  return integrand_tail<float>(quad_field<0>(q.w), quad_field<1>(q.w), quad_field<2>(q.w), del, d[0], d[1],
                               d[2], d[3], KB, pth, plo);
}

// The high word of a finite value read as a signed integer is positive exactly when the value is
// positive (for a double, at least 2^-1042): a running integer minimum of these words over r (1 - Mr) and the observer-time steps
// flags every finite invalid eval with one instruction.  Non-finite values (a NaN reads positive)
// are caught from the outputs by the reduction kernel.
__device__ __forceinline__ int sign_word(float x) { return __float_as_int(x); }
__device__ __forceinline__ int sign_word(double x) { return (int)(__double_as_longlong(x) >> 32); }
// one v_min3_i32 per step (the compiler's reassociation across the two-step loop made three
// instructions of two steps' minimums)
__device__ __forceinline__ int min3i(int a, int b, int c) {
  int r;
  __asm__("v_min3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
// x clamped into [lo, hi] (lo <= hi) in one instruction
__device__ __forceinline__ int med3i(int x, int lo, int hi) {
  int r;
  __asm__("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
  return r;
}

// ---------------------------------------------------------------------------------------------
// Eval + binning.  One wave per workgroup, persistent over a queue of items.
//
// Per eval (element e, step k, observer = lane) with tau_prev, tau, tau_next the observer times
// of steps k-1, k, k+1 (wrapped by -nT / +nT at the ends of the period, core.cu:173-194):
//   ic = floor(tau); the rising bins ic, ic-1, ... are the integers in (tau_prev, tau] with
//   weights (b - tau_prev)/(tau - tau_prev) = 1 - (fr + i)/(tau - tau_prev), fr = tau - ic; the
//   falling bins ic+1, ic+2, ... are the integers in (tau, tau_next) with weights
//   (tau_next - b)/(tau_next - tau) = 1 - (i + 1 - fr)/(tau_next - tau).
// The footprint ic-1 .. ic+2 takes the two nearest bins of each side with the weights clamped at
// zero: a weight is positive exactly when its bin lies inside the tent, so the clamp selects the
// tent's bins (up to the rounding of the weight itself, below 1e-6 in the mixed path).  Sides
// longer than two bins go to a wave-uniform tail that counts their bins from the FP64 times.
constexpr int kFlush = 10;  // window slots per round of the FP64 path's slab update

// Waves per SIMD that the LDS window allows (160 KB per CU, four SIMDs, one wave per workgroup):
// four for the mixed path's float2 window at W = 20, two for the FP64 path's double2 one.  The
// register budget is held to it.
constexpr int window_waves(bool mixed, int W) {
  const int per_cu = 160 * 1024 / (W * 64 * (mixed ? 8 : 16));
  return per_cu / 4 > 4 ? 4 : (per_cu / 4 < 1 ? 1 : per_cu / 4);
}

template <bool MIXED, bool MOVING, bool TAIL, int W>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(window_waves(MIXED, W))))
void tonal_eval_kernel(const TonalCaseDev* __restrict__ cases,
                                                        const Tile* __restrict__ tiles,
                                                        const Item* __restrict__ items,
                                                        const int* __restrict__ qoff,
                                                        int* __restrict__ counter) {
  using P = Prec<MIXED>;
  using R = typename P::R;
  using R2 = typename P::R2;
  using Attr = AttrT<R>;
  // what a step reads of its attribute record: the lane's field, broadcast over its row (FP64),
  // or the lane's row for the matrix-core products with one scalar field after it (mixed)
  using Rec = std::conditional_t<MIXED, MixRec, double>;
  // positions always through scalar loads (the FP64 path's went through vector loads, 25% slower
  // there: a wave-uniform record has no use for 64 lanes of vector addressing)
  using PosPtr = kptr<PosRec>;
  using AttrPtr = typename P::template ptr<Attr>;
  // the mixed path's weights are FP32: one residual correction of the observer time's root
  constexpr bool EXACT = !MIXED;
  // win[slot][lane] = (thickness, loading) partial of one observer-time bin
  __shared__ R2 win[W * 64];
  const int lane = threadIdx.x;
  // the lane's byte offset in an attribute record: row lane % 4 (mixed), field lane % 16 (FP64)
  const int loff = MIXED ? (lane & 3) * 16 : (lane & 15) * (int)sizeof(double);
  // the lane's byte addresses of window slots 0 and W - 4 (a footprint's first slot lies between)
  const int wlo = lane * (int)sizeof(R2), whi = ((W - 4) * 64 + lane) * (int)sizeof(R2);
  for (int s = 0; s < W; s++) win[s * 64 + lane] = R2{(R)0, (R)0};
  // Errors: `valid` is the minimum sign word of r (1 - Mr) and of every step tau_next - tau, so
  // it stays positive while every eval is valid and observer time increases.  Increasing time
  // makes the footprints of an element pass monotonic, so the window bound is checked on the
  // pass's first and last footprints (each footprint is also clamped into the window).
  int valid = 0x7fffffff;
  uint64_t winmask = 0, sweepmask = 0;  // lanes that raised each error
  // start on this XCD's queue (hardware register XCC_ID, bits 3:0), then drain the others
  int q = (int)(__builtin_amdgcn_s_getreg((3 << 11) | 20) & (kQueues - 1));
  int drained = 0;
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(&counter[q], 1);
    it = __builtin_amdgcn_readfirstlane(it);  // lane 0's value, wave-uniform
    const int idx = qoff[q] + it;
    if (idx >= qoff[q + 1]) {
      if (++drained == kQueues) break;
      q = (q + 1) & (kQueues - 1);
      continue;
    }
    const Item item = items[idx];
    const TonalCaseDev* Cp = cases + item.icase;
    const TimeConst k = Cp->k;
    const int nT = k.nT, nObs = Cp->nObs;
    const int o = item.obs0 + lane;
    const bool active = o < nObs;
    const int oc = active ? o : nObs - 1;
    const float ofx = Cp->obs[3 * oc], ofy = Cp->obs[3 * oc + 1], ofz = Cp->obs[3 * oc + 2];
    const LaneObs L = lane_obs(k, ofx, ofy, ofz);
    R* slab = (R*)Cp->slab + ((int64_t)item.group * nObs + oc) * (int64_t)nT * 2;
    const PosPtr pos = (PosPtr)(uintptr_t)Cp->pos;
    const AttrPtr attr = (AttrPtr)(uintptr_t)Cp->attr;
    const R KB = (R)k.KB, ux = (R)k.ux, uy = (R)k.uy, uz = (R)k.uz;
    const double dnT = (double)nT;
    // obs - src of the radiation vector: in FP32 from the parsed floats (mixed; the same bits as
    // the FP64 difference rounded to float), in FP64 as o + h2 m (FP64 path)
    // (h2 pinned in VGPRs: an FMA may read one scalar operand, and m comes through scalar loads)
    double h2 = k.h2;
    if constexpr (!MIXED) __asm__("" : "+v"(h2));
    auto diff = [&](const PosRec& p, R* dx, R* dy, R* dz) __attribute__((always_inline)) {
      if constexpr (MIXED) {
        *dx = ofx - p.x;
        *dy = ofy - p.y;
        *dz = ofz - p.z;
      } else {
        *dx = __builtin_fma(p.mx, h2, L.ox);
        *dy = __builtin_fma(p.my, h2, L.oy);
        *dz = __builtin_fma(p.mz, h2, L.oz);
      }
    };
    for (int ti = item.tile0; ti < item.tile0 + item.ntile; ti++) {
      const Tile tl = tiles[ti];
      for (int k0 = 0; k0 < nT; k0 += tl.K) {
        const int kend = min(k0 + tl.K, nT);
        const int km0 = (k0 > 0) ? k0 - 1 : nT - 1;
        const int kn0 = (k0 + 1 < nT) ? k0 + 1 : 0;
        int base;
        // the two positions an element's pass starts from (steps k0-1 and k0); the next
        // element's are loaded (through vector registers) while this element's steps run
        const cptr<PosRec> vpos = (cptr<PosRec>)(uintptr_t)Cp->pos;
        // the solve part of record i (the first 48 bytes of its 64)
        auto solve_part = [&](int64_t i) __attribute__((always_inline)) {
          return load_words<PosSolve>((cptr<PosSolve>)(vpos + i));
        };
        PosSolve nm = solve_part(tl.e0 * nT + km0);
        PosRec nc = load_words<PosRec>(vpos + tl.e0 * nT + k0);
        {
          double d;
          double tb = obs_time<MOVING, EXACT>(L, nm, &d);
          if (k0 == 0) tb = tb - dnT;
          base = (int)floor(tb) - tl.spread - 1;
        }
        // byte address in the window of a footprint's first slot ic - 1 - base, before clamping:
        // ic (64 sizeof(R2)) + wofs
        int wofs = (lane - 64 * (base + 1)) * (int)sizeof(R2);
        // kept opaque, so that each step forms the address in one shift-add (the compiler would
        // otherwise rebuild it from its parts with an extra instruction per step)
        __asm__("" : "+v"(wofs));
        for (int64_t e = tl.e0; e < tl.e0 + tl.ne; e++) {
          const PosPtr pe = pos + e * nT;
          const AttrPtr ae = attr + e * nT;
          // the element's attribute records as a buffer, so that a step's load takes the lane's
          // offset in a VGPR and the step's record offset in an SGPR (a global load would add a
          // 64-bit vector address every step); dword 3 is gfx9's raw-buffer word
          const __amdgpu_buffer_rsrc_t arec =
              __builtin_amdgcn_make_buffer_rsrc((void*)(uintptr_t)ae, (short)0, nT * (int)sizeof(Attr), 0x00020000);
          const PosSolve pm_e = nm;
          const PosRec pc_e = nc;
          if (e + 1 < tl.e0 + tl.ne) {
            nm = solve_part((e + 1) * nT + km0);
            nc = load_words<PosRec>(vpos + (e + 1) * nT + k0);
          }
          // Per-step state, kept twice and used in turn (cur is the step being evaluated, prv the
          // one before it), as are the record buffers: the loop runs two steps per iteration with
          // the roles swapped, so nothing is copied from one step to the next.
          struct Step {
            double t;  // observer time tau in bins
            int ic;    // floor(tau)
            R id;      // 1 / (tau - tau of the step before)
            R om;      // 1 - id
          };
          Step s0, s1;
          R dxc, dyc, dzc, delc;  // obs - src and the delay of the step being evaluated
          {
            double dl;
            double tp = obs_time<MOVING, EXACT>(L, pm_e, &dl);
            if (k0 == 0) tp = tp - dnT;  // the step before step 0 is step nT-1 of the previous period
            const double tc = obs_time<MOVING, EXACT>(L, pc_e, &dl);
            diff(pc_e, &dxc, &dyc, &dzc);
            delc = (R)dl;
            const R dp = (R)(tc - tp);
            s1 = Step{tp, (int)floor(tp), (R)0, (R)1};
            // tau >= 0 from here on (only the step before step 0 carries the -nT wrap), so
            // truncation is floor
            const R idp = rcpT(dp);
            s0 = Step{tc, (int)tc, idp, (R)1 - idp};
            valid = min(valid, sign_word(dp));
          }
          winmask |= __ballot(s0.ic - 1 - base < 0);  // the pass's first footprint
          // lanes whose current step has a rising side longer than two bins
          uint64_t longp = 0;
          if constexpr (TAIL) longp = __ballot(s1.ic < s0.ic - 1 && s0.t - s1.t > 1.999);
          // record pipeline: positions one step ahead of the solve, attributes one step ahead
          // of the integrand, each with one load in flight
          PosRec p0 = load_words<PosRec>(pe + kn0), p1;
          auto load_rec = [&](int kk) __attribute__((always_inline)) -> Rec {
            if constexpr (MIXED) {
              const auto v = __builtin_amdgcn_raw_buffer_load_b128(arec, loff, kk * (int)sizeof(Attr), 0);
              MixRec m;
              __builtin_memcpy(&m, &v, sizeof m);
              return m;
            } else {
              const auto v = __builtin_amdgcn_raw_buffer_load_b64(arec, loff, kk * (int)sizeof(Attr), 0);
              double f;
              __builtin_memcpy(&f, &v, sizeof f);
              return f;
            }
          };
          Rec a0r = load_rec(k0), a1r;
          // The window slots of the step to be evaluated.  In the variants without the tail they
          // are read at the end of the step before it (after that step's writes, which they may
          // overlap): the step's one wait, at its top, then covers them with the scalar position
          // loads issued a whole step earlier, and no wait in the middle of the step holds up the
          // loads issued at its top (mixed -1.7%, FP64 -3.4%).  The variants with the tail,
          // short of registers to keep the slots across a step, read them at the top of the
          // step, which waits for them (and for that step's record loads) before the update.
          constexpr bool AHEAD = !TAIL;
          R2* wp;
          R2 v0, v1, v2, v3;
          auto read_window = [&](int ic) __attribute__((always_inline)) {
            // (the footprint's first slot ic - 1 - base clamped into [0, W - 4], as a byte
            // address: the clamp bounds are the lane's addresses of slots 0 and W - 4)
            wp = (R2*)((char*)win + med3i(ic * (64 * (int)sizeof(R2)) + wofs, wlo, whi));
            v0 = wp[0];
            v1 = wp[64];
            v2 = wp[128];
            v3 = wp[192];
          };
          if constexpr (AHEAD) read_window(s0.ic);
          auto step = [&](int kk, bool wrap, Step& cur, Step& prv, const PosRec& pn, PosRec& pnn,
                          const Rec& ac, Rec& an) __attribute__((always_inline)) {
            const int kn = (kk + 1 < nT) ? kk + 1 : 0;
            const int knn = (kn + 1 < nT) ? kn + 1 : 0;
            // Wait for the loads of the step before (issued a step ago, so complete) before issuing
            // this step's: scalar loads return out of order, so a wait placed after the new loads
            // would wait for them too.  lgkmcnt(0), vmcnt and expcnt unlimited.
            __builtin_amdgcn_s_waitcnt(0xc07f);
            // (mixed: the pad of the position record read a step ago, used here where it has
            // arrived, keeps the compiler loading each record whole, in two scalar loads instead
            // of four; the FP64 path reads only the first 48 bytes)
            if constexpr (MIXED) __asm__("" ::"s"(pn.pad));
            pnn = load_words<PosRec>(pe + knn);
            an = load_rec(kn);
            const int ic = cur.ic;
            if constexpr (!AHEAD) read_window(ic);
            // keep the record loads (and the window reads) at the top of the step: their latency
            // then overlaps the solve and the integrand
            __builtin_amdgcn_sched_barrier(0);
            // observer time of the next step
            double dln;
            double t_next = obs_time<MOVING, EXACT>(L, pn, &dln);
            if (wrap) t_next = t_next + dnT;
            // integrand of this step
            R pth, plo;
            R rm;
            if constexpr (MIXED)
              rm = integrand_mfma<MOVING>(ac, dxc, dyc, dzc, delc, KB, ux, uy, uz, &pth, &plo);
            else
              rm = integrand<MOVING>(ac, dxc, dyc, dzc, delc, KB, ux, uy, uz, &pth, &plo);
            // tent weights of the footprint
            const R fr = fractT(cur.t, R());  // tau - ic, exact
            const R dn = (R)(t_next - cur.t);
            valid = min3i(valid, sign_word(rm), sign_word(dn));
            const R idp = cur.id, idn = rcpT(dn);
            const R omn = (R)1 - idn;
            const R frm = fr - (R)1;
            // (1 - idp carried from the step before)
            const R w0 = clamp01(fmaT(-fr, idp, cur.om));  // bin ic-1: 1 - (fr + 1) / dp
            const R w1 = clamp01(fmaT(-fr, idp, (R)1));        // bin ic:   1 - fr / dp
            const R w2 = clamp01(fmaT(frm, idn, (R)1));        // bin ic+1: 1 - (1 - fr) / dn
            const R w3 = clamp01(fmaT(frm, idn, omn));         // bin ic+2: 1 - (2 - fr) / dn
            v0.x = fmaT(pth, w0, v0.x);
            v0.y = fmaT(plo, w0, v0.y);
            v1.x = fmaT(pth, w1, v1.x);
            v1.y = fmaT(plo, w1, v1.y);
            v2.x = fmaT(pth, w2, v2.x);
            v2.y = fmaT(plo, w2, v2.y);
            v3.x = fmaT(pth, w3, v3.x);
            v3.y = fmaT(plo, w3, v3.y);
            wp[0] = v0;
            wp[64] = v1;
            wp[128] = v2;
            wp[192] = v3;
            if constexpr (TAIL) {
              // lanes with a side longer than two bins; a step's rising side is the previous
              // step's falling side, so one compare per step carries both
              const uint64_t longn = __ballot(dn > (R)1.999);
              const uint64_t tailm = longn | longp;
              longp = longn;
              if (tailm) {
                // keeps this wave-uniform test a branch of its own (the lane test below then
                // costs nothing on the common path)
                __asm__ volatile("");
                if ((tailm >> lane) & 1) {
                  const int nd = ic - prv.ic;
                  const int nu = (int)ceil(t_next) - 1 - ic;
                  if (nd > W || nu > W) sweepmask |= 1;
                  const int nd_ = min(nd, W), nu_ = min(nu, W);
                  auto put = [&](int b, R w) {
                    const int slot = b - base;
                    if (slot < 0 || slot >= W) {
                      winmask |= 1;
                      return;
                    }
                    R2 v = win[slot * 64 + lane];
                    v.x = fmaT(pth, w, v.x);
                    v.y = fmaT(plo, w, v.y);
                    win[slot * 64 + lane] = v;
                  };
                  for (int j = 2; j < nd_; j++) put(ic - j, fmaT(-(fr + (R)j), idp, (R)1));
                  for (int j = 2; j < nu_; j++) put(ic + 1 + j, fmaT(fr - (R)(1 + j), idn, (R)1));
                }
              }
            }
            // the step before becomes the next step
            prv.t = t_next;
            prv.ic = (int)t_next;
            prv.id = idn;
            prv.om = omn;
            diff(pn, &dxc, &dyc, &dzc);
            delc = (R)dln;
            if constexpr (AHEAD) read_window(prv.ic);  // the next step's (harmless after a pass's last)
          };
          // Steps run in pairs with the state and record buffers in alternating roles (A is the
          // first step's current state, B its next), so nothing is copied between steps.  An odd
          // step before the period's last runs first, singly, and the pairs then start with the
          // roles swapped; the period's last step (which wraps tau_next by +nT) runs singly after
          // them.  The pass's last footprint is then checked.
          const int kstop = (kend == nT) ? nT - 1 : kend;
          auto pairs = [&](int kk, Step& A, Step& B, PosRec& pA, PosRec& pB, Rec& rA, Rec& rB)
                           __attribute__((always_inline)) {
            for (; kk + 1 < kstop; kk += 2) {
              step(kk, false, A, B, pA, pB, rA, rB);
              step(kk + 1, false, B, A, pB, pA, rB, rA);
            }
            int last_ic = B.ic;  // the current state of the last step evaluated
            if (kk < kend) {
              step(kk, true, A, B, pA, pB, rA, rB);
              last_ic = A.ic;
            }
            winmask |= __ballot(last_ic - 1 - base > W - 4);
          };
          if ((kstop - k0) & 1) {
            step(k0, false, s0, s1, p0, p1, a0r, a1r);
            pairs(k0 + 1, s1, s0, p1, p0, a1r, a0r);
          } else {
            pairs(k0, s0, s1, p0, p1, a0r, a1r);
          }
        }
        // add the window into this item's slab row and clear it (untouched slots hold zeros)
        int b = base % nT;
        if (b < 0) b += nT;
        if (!MIXED && W <= nT) {
          // FP64 path (two waves per SIMD, 16-byte slots): the W bins are distinct, so each round
          // loads the touched slots' slab values before adding to any and waits for one memory
          // latency instead of one per slot (8% faster there with rounds of 10; the mixed path,
          // with four waves per SIMD to hide the latency, ran 1.3% slower this way and keeps the
          // slot-by-slot loop)
          R2* row = (R2*)slab;
          for (int s0 = 0; s0 < W; s0 += kFlush) {
            R2 v[kFlush], old[kFlush];
            int bs[kFlush];
            bool nz[kFlush];
#pragma unroll
            for (int j = 0; j < kFlush; j++) {
              const int s = s0 + j;
              nz[j] = false;
              if (s < W) {
                v[j] = win[s * 64 + lane];
                nz[j] = active && (v[j].x != (R)0 || v[j].y != (R)0);  // untouched slots stay out
                win[s * 64 + lane] = R2{(R)0, (R)0};
                bs[j] = (b + s >= nT) ? b + s - nT : b + s;
              }
            }
#pragma unroll
            for (int j = 0; j < kFlush; j++)
              if (nz[j]) old[j] = row[bs[j]];
#pragma unroll
            for (int j = 0; j < kFlush; j++)
              if (nz[j]) row[bs[j]] = R2{old[j].x + v[j].x, old[j].y + v[j].y};
          }
        } else {  // slot by slot (also where a period shorter than the window repeats bins)
          for (int s = 0; s < W; s++) {
            const R2 v = win[s * 64 + lane];
            if (v.x != (R)0 || v.y != (R)0) {
              win[s * 64 + lane] = R2{(R)0, (R)0};
              if (active) {
                slab[2 * b] += v.x;
                slab[2 * b + 1] += v.y;
              }
            }
            b = (b + 1 == nT) ? 0 : b + 1;
          }
        }
      }
    }
    // the item's errors go to its case, so that a case whose window bound broke can be run again
    const uint32_t flags = (wave_any(valid <= 0) ? kFlagNumeric : 0u) |
                           (winmask ? kFlagWindow : 0u) | (sweepmask ? kFlagSweep : 0u);
    if (flags && lane == 0) atomicOr(Cp->flags, flags);
    valid = 0x7fffffff;
    winmask = 0;
    sweepmask = 0;
  }
}

// Ordered reduction of the partial slabs: out(b, o) = sum_g slab[g][o][b], FP64 Kahan.
// A block of 256 threads covers 64 bins x 16 observers: thread (tx, ty) sums bins b0 + 2 tx and
// b0 + 2 tx + 1 (thickness and loading: two 2-vectors per partial row) of observers o0 + ty and
// o0 + ty + 8, with eight partial rows in flight; LDS transposes the sums to the time-major
// outputs, written as runs of 16 observers.
template <typename A>
struct Vec2;
template <>
struct Vec2<float> {
  using T = float2;
};
template <>
struct Vec2<double> {
  using T = double2;
};

template <typename A>
__global__ __launch_bounds__(256) void tonal_reduce_kernel(const TonalCaseDev* __restrict__ cases,
                                                           int icase) {
  using V2 = typename Vec2<A>::T;
  constexpr int kRows = 8;
  const TonalCaseDev& C = cases[icase];
  const int nT = C.k.nT, nObs = C.nObs, G = C.G;
  const A* slab = (const A*)C.slab;
  const int b0 = blockIdx.x * 64, o0 = blockIdx.y * 16;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  __shared__ float sT[16][65], sL[16][65], sA[16][65];
  const int64_t gstride = (int64_t)nObs * nT;  // in V2 units
  for (int i = 0; i < 2; i++) {
    const int ol = ty + 8 * i, o = o0 + ol, b = b0 + 2 * tx;
    if (o < nObs && b < nT) {
      const bool two = b + 1 < nT;
      const V2* p = (const V2*)slab + (int64_t)o * nT + b;
      double s[4] = {0.0, 0.0, 0.0, 0.0}, c[4] = {0.0, 0.0, 0.0, 0.0};
      auto add = [&](int k, double x) {
        const double y = x - c[k], t = s[k] + y;
        c[k] = (t - s[k]) - y;
        s[k] = t;
      };
      int g = 0;
      for (; g + kRows <= G; g += kRows) {
        V2 v0[kRows], v1[kRows];
#pragma unroll
        for (int r = 0; r < kRows; r++) {
          v0[r] = p[(g + r) * gstride];
          v1[r] = two ? p[(g + r) * gstride + 1] : V2{(A)0, (A)0};
        }
#pragma unroll
        for (int r = 0; r < kRows; r++) {
          add(0, (double)v0[r].x);
          add(1, (double)v0[r].y);
          add(2, (double)v1[r].x);
          add(3, (double)v1[r].y);
        }
      }
      for (; g < G; g++) {
        const V2 a0 = p[g * gstride];
        const V2 a1 = two ? p[g * gstride + 1] : V2{(A)0, (A)0};
        add(0, (double)a0.x);
        add(1, (double)a0.y);
        add(2, (double)a1.x);
        add(3, (double)a1.y);
      }
      const float out[6] = {(float)s[0], (float)s[1], (float)(s[0] + s[1]),
                            (float)s[2], (float)s[3], (float)(s[2] + s[3])};
      sT[ol][2 * tx] = out[0];
      sL[ol][2 * tx] = out[1];
      sA[ol][2 * tx] = out[2];
      sT[ol][2 * tx + 1] = out[3];
      sL[ol][2 * tx + 1] = out[4];
      sA[ol][2 * tx + 1] = out[5];
      // A non-finite output fails the case: an eval whose observer time or integrand is not
      // finite (a source on an observer makes the square root's argument zero) puts a NaN into
      // every slot of its footprint, whatever the weights, and the ordered sums carry it here;
      // a float overflow of a finite sum lands here too.  The eval kernel's sign test covers
      // the finite invalid evals (1 - Mr <= 0, observer time not increasing).
      bool bad = false;
#pragma unroll
      for (int k = 0; k < 6; k++) bad |= !__builtin_isfinite(out[k]);
      if (bad) atomicOr(C.flags, kFlagNumeric);
    }
  }
  __syncthreads();
  for (int i = 0; i < 4; i++) {
    const int idx = threadIdx.x + 256 * i, bl = idx >> 4, ol = idx & 15;
    const int b = b0 + bl, o = o0 + ol;
    if (b < nT && o < nObs) {
      C.pT[(int64_t)b * nObs + o] = sT[ol][bl];
      C.pL[(int64_t)b * nObs + o] = sL[ol][bl];
      C.pA[(int64_t)b * nObs + o] = sA[ol][bl];
    }
  }
}

// Reference AoS (X[4], N[6], V[6], PR[2] per element-step, surfaces concatenated, element-major
// with nT steps each) -> position and attribute records; the products are formed in FP64.
static_assert(sizeof(PosRec) == 64 && sizeof(PosSolve) == 48 && offsetof(PosRec, jt) == offsetof(PosSolve, jt),
              "position record layout");
__global__ void tonal_pos_kernel(const float* __restrict__ X, int64_t n, TimeConst k, PosRec* __restrict__ pos) {
  const double m2C2 = -2.0 * k.C2;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const float* x = X + 4 * q;
    const double px = (double)x[0] - k.Ox, py = (double)x[1] - k.Oy, pz = (double)x[2] - k.Oz;
    PosRec o;
    o.mx = px * m2C2;
    o.my = py * m2C2;
    o.mz = pz * m2C2;
    o.q = __builtin_fma(pz, pz, __builtin_fma(py, py, px * px)) * k.C2;
    o.wp = __builtin_fma(pz, k.wz, __builtin_fma(py, k.wy, px * k.wx));
    o.jt = (double)(int)(q % k.nT) * k.SB;
    o.x = x[0];
    o.y = x[1];
    o.z = x[2];
    o.pad = 0.0f;
    pos[q] = o;
  }
}

// The records are stored scaled for the eval kernel's integrand (which takes the delay in bins
// for r): v, vd, aM1 and cT times s = KB / a0 = 1 / aB, l, ld and aLM times s^2.
template <typename R>
__global__ void tonal_attr_kernel(const float* __restrict__ X, const float* __restrict__ N,
                                  const float* __restrict__ V, const float* __restrict__ PR,
                                  int64_t n, double a0, double rhoRef, double KB, AttrT<R>* __restrict__ attr,
                                  uint32_t* __restrict__ flags) {
  bool bad = false;
  const double sc = KB / a0, sc2 = sc * sc;
  const double inv_a0 = sc / a0;
  const double D1 = 1.0 / (4.0 * M_PI * a0);
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const float* x = X + 4 * q;
    const float* nn = N + 6 * q;
    const float* v = V + 6 * q;
    const double P = PR[2 * q], dP = PR[2 * q + 1], ds = x[3];
    bad |= !isfinite(P) || !isfinite(dP);
    for (int i = 0; i < 6; i++) bad |= !isfinite(nn[i]) || !isfinite(v[i]);
    for (int i = 0; i < 4; i++) bad |= !isfinite(x[i]);
    AttrT<R> o;
    double Vn = 0.0, VV = 0.0;
    const double cL = ds * D1 * sc2;
    for (int i = 0; i < 3; i++) {
      const double Vi = v[i], Vdi = v[3 + i], ni = nn[i], ndi = nn[3 + i];
      o.v[i] = (R)(Vi * inv_a0);
      o.vd[i] = (R)(Vdi * inv_a0);
      o.l[i] = (R)(cL * (P * ni));
      o.ld[i] = (R)(cL * (dP * ni + P * ndi));
      Vn += Vi * ni;
      VV += Vi * Vi;
    }
    o.aM1 = (R)((VV / a0 - a0) * sc);
    o.aLM = (R)(cL * (P * Vn));
    o.cT = (R)(rhoRef * Vn * ds * a0 * D1 * sc);
    o.unused = (R)0;
    if constexpr (sizeof(R) == sizeof(float)) {
      // the mixed record: each row with one scalar field after it (MixRec)
      const R w[16] = {o.v[0],  o.v[1],  o.v[2],  o.aM1, o.vd[0], o.vd[1], o.vd[2], o.aLM,
                       o.l[0],  o.l[1],  o.l[2],  o.cT,  o.ld[0], o.ld[1], o.ld[2], (R)0};
      __builtin_memcpy(&attr[q], w, sizeof w);
    } else {
      attr[q] = o;
    }
  }
  if (bad) atomicOr(flags, kFlagNumeric);  // non-finite element input
}

template <typename T>
__global__ void fill_zero_kernel(T* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (T)0;
}

__global__ void reset_counter_kernel(int* c, uint32_t* flags) {
  for (int i = 0; i < 2 * kQueues; i++) c[i] = 0;  // still-air queues, then moving-medium queues
  *flags = 0u;
}

// Tent of every (element, step) for one observer, from the same canonical observer time: the top
// rising bin ic (stored as ic mod nT, with its period index floor(ic / nT)) and the two side
// lengths.
template <bool MOVING>
__global__ void tonal_bins_kernel(const PosRec* __restrict__ pos, int64_t nE, TimeConst k,
                                  double ox, double oy, double oz, int32_t* ipts, int32_t* npi,
                                  int32_t* nd, int32_t* nu) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nE) return;
  const int nT = k.nT;
  const PosRec* pe = pos + e * nT;
  const LaneObs L = lane_obs(k, ox, oy, oz);
  auto tau = [&](int j) {
    const PosRec p = pe[j];
    double d;
    return obs_time<MOVING, true>(L, p, &d);
  };
  double t_prev = tau(nT - 1) - (double)nT;
  double t_cur = tau(0);
  for (int kk = 0; kk < nT; kk++) {
    const double t_next = (kk + 1 < nT) ? tau(kk + 1) : tau(0) + (double)nT;
    const int64_t idx = e * nT + kk;
    const bool ok = fabs(t_prev) < 1.0e9 && fabs(t_cur) < 1.0e9 && fabs(t_next) < 1.0e9;
    if (ok) {
      const double fc = floor(t_cur);
      const int ic = (int)fc;
      const int d = ic - (int)floor(t_prev);
      const int u = (int)ceil(t_next) - 1 - ic;
      int per = ic / nT, ph = ic % nT;
      if (ph < 0) {
        ph += nT;
        per -= 1;
      }
      ipts[idx] = ph;
      npi[idx] = per;
      nd[idx] = d > 0 ? d : 0;
      nu[idx] = u > 0 ? u : 0;
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
#define ACUM_EVAL_VARIANTS(X) \
  X(true, 16) X(true, 20) X(true, 24) X(true, 32) X(true, 48) X(true, 64) X(false, 16) X(false, 20) \
  X(false, 24) X(false, 32)

hipError_t launch_tonal_eval(bool mixed, bool moving, bool tail, int W, int grid, hipStream_t s,
                             const TonalCaseDev* cases, const Tile* tiles, const Item* items,
                             const int* qoff, int* counter) {
  dim3 b(64);
#define ACUM_EVAL1(M, MV, T, WW)                                                                  \
  if (mixed == M && moving == MV && tail == T && W == WW) {                                       \
    hipLaunchKernelGGL((tonal_eval_kernel<M, MV, T, WW>), dim3(grid), b, 0, s, cases, tiles, items, \
                       qoff, counter);                                                            \
    return hipGetLastError();                                                                     \
  }
#define ACUM_EVAL(M, WW) \
  ACUM_EVAL1(M, true, true, WW) ACUM_EVAL1(M, true, false, WW) ACUM_EVAL1(M, false, true, WW) ACUM_EVAL1(M, false, false, WW)
  ACUM_EVAL_VARIANTS(ACUM_EVAL)
#undef ACUM_EVAL
#undef ACUM_EVAL1
  return hipErrorInvalidValue;
}

// Occupancy of the tail-free variant (the register budget is held to the window's occupancy, so
// the tail does not change it).
hipError_t eval_occupancy(bool mixed, bool moving, int W, int* blocks_per_cu) {
#define ACUM_OCC(M, WW)                                                                        \
  if (mixed == M && W == WW)                                                                   \
    return moving ? hipOccupancyMaxActiveBlocksPerMultiprocessor(                              \
                        blocks_per_cu, tonal_eval_kernel<M, true, false, WW>, 64, 0)           \
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(                              \
                        blocks_per_cu, tonal_eval_kernel<M, false, false, WW>, 64, 0);
  ACUM_EVAL_VARIANTS(ACUM_OCC)
#undef ACUM_OCC
  return hipErrorInvalidValue;
}

bool eval_window_supported(bool mixed, int W) {
#define ACUM_HAS(M, WW) \
  if (mixed == M && W == WW) return true;
  ACUM_EVAL_VARIANTS(ACUM_HAS)
#undef ACUM_HAS
  return false;
}

hipError_t launch_tonal_reduce(bool mixed, hipStream_t s, const TonalCaseDev* cases, int icase,
                               int nT, int nObs) {
  dim3 g((nT + 63) / 64, (nObs + 15) / 16), b(256);
  if (mixed)
    hipLaunchKernelGGL(tonal_reduce_kernel<float>, g, b, 0, s, cases, icase);
  else
    hipLaunchKernelGGL(tonal_reduce_kernel<double>, g, b, 0, s, cases, icase);
  return hipGetLastError();
}

hipError_t launch_tonal_pos(hipStream_t s, const float* X, int64_t n, const TimeConst& k, void* pos) {
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(tonal_pos_kernel, dim3(grid), dim3(256), 0, s, X, n, k, (PosRec*)pos);
  return hipGetLastError();
}

hipError_t launch_tonal_attr(bool mixed, hipStream_t s, const float* X, const float* N,
                             const float* V, const float* PR, int64_t n, double a0, double rhoRef,
                             double KB, void* attr, uint32_t* flags) {
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
  if (mixed)
    hipLaunchKernelGGL(tonal_attr_kernel<float>, dim3(grid), dim3(256), 0, s, X, N, V, PR, n, a0,
                       rhoRef, KB, (AttrT<float>*)attr, flags);
  else
    hipLaunchKernelGGL(tonal_attr_kernel<double>, dim3(grid), dim3(256), 0, s, X, N, V, PR, n, a0,
                       rhoRef, KB, (AttrT<double>*)attr, flags);
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

hipError_t launch_tonal_bins(hipStream_t s, const void* pos, int64_t nE, const TimeConst& k,
                             double ox, double oy, double oz, int32_t* ipts, int32_t* npi,
                             int32_t* nd, int32_t* nu) {
  const int grid = (int)((nE + 63) / 64);
  if (k.still)
    hipLaunchKernelGGL(tonal_bins_kernel<false>, dim3(grid), dim3(64), 0, s, (const PosRec*)pos, nE,
                       k, ox, oy, oz, ipts, npi, nd, nu);
  else
    hipLaunchKernelGGL(tonal_bins_kernel<true>, dim3(grid), dim3(64), 0, s, (const PosRec*)pos, nE,
                       k, ox, oy, oz, ipts, npi, nd, nu);
  return hipGetLastError();
}

}  // namespace acum
