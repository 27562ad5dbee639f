// Types shared by the host planner and the HIP kernels of the tonal and broadband paths.
#pragma once
#include <cstdint>

namespace acum {

// Element-time records: one (element, source step), wave-uniform in the tonal kernel, read with
// scalar loads, element-major ([element][step]) like the reference AoS (io.cu:280-284).
// PosRec holds what the canonical observer time (below) needs of the source position, formed in
// FP64 from the parsed floats, and the parsed floats themselves (the FP32 radiation vector).
// AttrT holds the observer-independent factors of the Farassat 1A integrand (core.cu:68-154),
// formed in FP64 from the parsed floats and stored in the path's precision R:
//   v   = V / a0,  vd = dV/dt / a0                        (Mr, dMr of core.cu:93-95)
//   l   = cL P n,  ld = cL (dP/dt n + P dn/dt)             (L, dL/dt of core.cu:130-131)
//   aM1 = a0 (M^2 - 1) = |V|^2 / a0 - a0,  aLM = cL a0 L_M = cL P (V.n)   (core.cu:94,134)
//   cT  = rhoRef (V.n) dS a0 D1,  cL = dS D1,  D1 = 1/(4 pi a0)   (core.cu:53,123-124,138-141)
// (the loading factor cL is folded into l, ld and aLM, which only enter the loading term), each
// stored scaled by s = KB / a0 (v, vd, aM1, cT) or s^2 (l, ld, aLM), so that the eval kernel's
// integrand takes the delay in bins T where it would take r = (a0 / KB) T
struct PosRec {
  double mx, my, mz;  // (-2 C2) (p - O)
  double q;           // C2 |p - O|^2
  double wp;          // w.(p - O) (zero in still air)
  double jt;          // j SB: the step's emission time in bins
  float x, y, z;      // the parsed position
  float pad;
};  // 64 bytes
// The part of a PosRec the observer time reads (the first 48 bytes).
struct PosSolve {
  double mx, my, mz, q, wp, jt;
};
// An element's attributes at one step.  The FP64 records are stored in this order; the FP32 ones
// with each row followed by one scalar, {v, aM1, vd, aLM, l, cT, ld, 0}, for the mixed path's
// 16-byte row loads.
template <typename R>
struct AttrT {
  R v[3], vd[3], l[3], ld[3];
  R aM1, aLM, cT, unused;
};  // 64 bytes (float), 128 bytes (double)

// Canonical FP64 observer time, in units of the observer-time bin dPsi; the CPU oracle forms
// the same values with the same operations (no contraction other than the explicit fma).  With
// KB = oM / dPsi, na = a0^2 - |vinf|^2 > 0, C2 = KB (KB / na), w = vinf (KB / na), SB = dTau KB,
// aB = a0 / KB, u = vinf / KB and the origin O = the centre of the sources' bounding box:
//   per observer: o = obs - O, qo = C2 fma(oz,oz, fma(oy,oy, ox ox)), wo = fma(oz,wz, fma(oy,wy, ox wx))
//   per eval:     cc = fma(ox,mx, fma(oy,my, fma(oz,mz, qo + q)))        = C2 |obs - src|^2
//                 b = wo - wp;  T = sqrt(fma(b, b, cc)) - b;  tau = T + jt
//                 (still air: T = sqrt(cc));  r = aB T;  r_vec = (obs - src) - u T
// T = KB t is the delay in bins, t the positive root of a' t^2 + b' t + c = 0 with
// a' = |vinf|^2 - a0^2 < 0, b' = -2 d.vinf, c = |d|^2 (acous_dist, core.cu:261-279):
// t = (s - d.vinf) / na with s = sqrt((d.vinf)^2 + c na), scaled by KB inside the square root.
// tau is (j dTau + r / a0) oM / dPsi (core.cu:156).  The bins of an eval are the integers b in
// (tau_prev, tau] (rising side of the tent) and in (tau, tau_next) (falling side); bin b is
// observer time b dPsi, stored at b mod nT.
struct TimeConst {
  double a0, r2b, KB, SB;
  double wx, wy, wz, C2;  // moving medium: canonical observer time
  double aB, ux, uy, uz;  // r = aB T, r_vec = d - u T with u = vinf / KB
  double Ox, Oy, Oz;      // origin of the expansion: centre of the sources' bounding box
  double h2;              // 1 / (2 C2): p - O = -h2 m, for the FP64 radiation vector
  int32_t nT, still;      // still: vinf == 0
};

struct TonalCaseDev {
  const PosRec* pos;    // [nE][nT]
  const void* attr;     // [nE][nT] AttrT<float> or AttrT<double>
  const float* obs;     // [nObs][3]
  void* slab;           // [G][nObs][nT][2] partial sums (float or double)
  float* pT;            // outputs [nT][nObs]
  float* pL;
  float* pA;
  int64_t nE;
  int32_t nObs, G;
  uint32_t* flags;      // this case's error flags (kFlag*), cleared before each pass
  TimeConst k;
};

struct Tile {
  int64_t e0;     // first element (case-global index over all surfaces)
  int32_t ne;     // elements
  int32_t spread; // bins the tile's elements can sit from its first element, either side
  int32_t K;      // source steps per window pass
  int32_t pad;
};

struct Item {
  int32_t icase;
  int32_t obs0;   // first observer of the 64-lane group
  int32_t tile0;  // first tile of the group (index into the tile array)
  int32_t ntile;
  int32_t group;  // partial slab row
  int32_t pad;
  int64_t cost;   // evals, for longest-first ordering
};

// Work queues of one eval launch: one per XCD (an MI355X has 8), each holding whole tile groups
// (all observer groups of a group of tiles) so the waves that share those tiles' records share
// an L2.  A wave starts on its own XCD's queue and then drains the others.
constexpr int kQueues = 8;

// Error flags raised by the kernels (OR-ed into a device word).
enum : uint32_t {
  kFlagWindow = 1u,   // a contribution fell outside its tile window (planner bound broken)
  kFlagNumeric = 2u,  // r (1 - Mr) <= 0, observer time not increasing, or a non-finite output
  kFlagSweep = 4u     // a bin sweep ran longer than the window
};

}  // namespace acum
