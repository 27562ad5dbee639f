// Arguments of the GPU preprocessing kernel (preprocess.hip).
#pragma once
#include <cstdint>

namespace acum {

constexpr int kPrepMaxSteps = 4096;  // series held in LDS as doubles

struct PrepArgs {
  float* X;   // [q][4]  x, y, z, dS
  float* N;   // [q][6]  n, dn/dt
  float* V;   // [q][6]  v, dv/dt (written)
  float* PR;  // [q][2]  CP -> P, dP/dt
  int64_t nE;
  int nT;
  int periodic, lowpass;
  double C, pscale, dt;
  double g[13];  // normalised 13-point Gaussian weights (mathfunc.cu:350-370)
};

}  // namespace acum
