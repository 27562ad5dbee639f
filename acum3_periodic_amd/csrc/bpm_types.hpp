// Per-source records of the broadband kernels (host and device).
#pragma once
#include <cstdint>

namespace acum {

struct BpmShapeIn {   // FP64 per source (section, step), input of the spectral-shape kernel
  double u, c, lsect, alpha, ds, dp, dss, dsp, R;
  int32_t trip, tip;
};

struct BpmSrcDev {    // FP32 per source, read by the observer kernel
  float xt[3];
  float tr[9];
  float M, Mmc;       // Mach and Mach - 0.6 Mach (convection term of D_h)
  float tipT0;        // M^2 mm^3 tip_l^2 (tip-vortex level before directivity)
  uint32_t flags;     // 1: separation term uses D_l; 2: LBL-VS active; 4: tip section
};

}  // namespace acum
