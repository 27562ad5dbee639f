// Types shared by the host planner and the HIP kernels of the tonal and broadband paths.
#pragma once
#include <cstdint>

namespace acum {

// Element-time record: one (element, source step), wave-uniform in the tonal kernel.
// Fields: 0-2 position, 3 dS, 4-6 V, 7-9 dV/dt, 10-12 n, 13-15 dn/dt, 16 P, 17 dP/dt,
// 18 M^2 = (|V|/a0)^2, 19 V.n   (the last two are observer independent, core.cu:94,102).
constexpr int kRecFields = 20;

// Canonical FP64 observer-time constants; the CPU oracle forms the same values.
struct TimeConst {
  double a0, inv_a0, dTau, oM, lpi, inv_lpi, dPsi, inv_dPsi, four_a, inv_2a;
  double vx, vy, vz;
  int32_t nT, still;  // still: vinf == 0, the b term vanishes exactly
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
