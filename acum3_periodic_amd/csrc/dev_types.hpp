// Types shared by the host planner and the HIP kernels of the tonal and broadband paths.
#pragma once
#include <cstdint>

namespace acum {

// Element-time record: one (element, source step), wave-uniform in the tonal kernel.
// The 18 input values stay float (they are parsed floats); the two observer-independent
// products M^2 = (|V|/a0)^2 (core.cu:94) and V.n (core.cu:102) are formed in FP64 and kept
// in the path's precision.
struct Rec32 {
  float x[3], ds, v[3], vd[3], n[3], nd[3], P, dP;
  float MM, Vn;
};  // 80 bytes
struct Rec64 {
  float x[3], ds, v[3], vd[3], n[3], nd[3], P, dP;
  double MM, Vn;
};  // 88 bytes

// Canonical FP64 observer time, in units of the observer-time bin dPsi; the CPU oracle forms
// the same values with the same operations.
//   c = fma(dz,dz, fma(dy,dy, dx*dx))                         |obs - src|^2
//   still air: r = sqrt(c)
//   otherwise: b = -2 fma(dz,vz, fma(dy,vy, dx*vx)); disc = fma(b, b, -(four_a c));
//              t = (-b - sqrt(disc)) inv_2a, or (-b + sqrt(disc)) inv_2a if negative; r = t a0
//   tau(j) = fma(r, r2b, j)       r2b = (oM / a0) / dPsi: (j dTau + r / a0) oM / dPsi
// The bins of an eval are the integers b in (tau_prev, tau] (rising side of the tent) and in
// (tau, tau_next) (falling side); bin b is observer time b dPsi, stored at b mod nT.
struct TimeConst {
  double a0, r2b, four_a, inv_2a;
  double vx, vy, vz;
  double inv_a0;
  int32_t nT, still;  // still: vinf == 0
};

struct TonalCaseDev {
  const void* rec;      // [nE][nT] records (float or double)
  const float* obs;     // [nObs][3]
  void* slab;           // [G][nObs][nT][2] partial sums (float or double)
  float* pT;            // outputs [nT][nObs]
  float* pL;
  float* pA;
  int64_t nE;
  int32_t nObs, G;
  TimeConst k;
  double rhoRef, D1;    // D1 = 1/(4 pi a0)
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

// Error flags raised by the kernels (OR-ed into a device word).
enum : uint32_t {
  kFlagWindow = 1u,   // a contribution fell outside its tile window (planner bound broken)
  kFlagNumeric = 2u,  // non-finite observer time, r <= 0 or 1 - Mr <= 0
  kFlagSweep = 4u     // a bin sweep ran longer than the window
};

}  // namespace acum
