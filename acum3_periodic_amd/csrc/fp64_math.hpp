// FP64 helpers for the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace acum {

// Correctly rounded square root of a positive normal double, by Goldschmidt iteration from the
// hardware reciprocal square root (Markstein, "IA-64 and Elementary Functions", 2000, ch. 8):
//   y ~ 1/sqrt(x);  g = x y ~ sqrt(x);  h = y/2 ~ 1/(2 sqrt(x))
//   r = 1/2 - g h;  g += g r;  h += h r          (one Goldschmidt step)
//   twice: d = x - g^2 (exact by fma);  g += d h  (residual corrections; the last one rounds
//                                                  correctly)
// The argument ranges used here (squared distances in m^2, 1e-12 .. 1e12) need none of the
// denormal or overflow rescaling a general sqrt carries.
__device__ __forceinline__ double sqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = 0.5 * y;
  const double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return g;
}

// The same with one residual correction: before its final rounding the result is within about
// 2^-92 (relative) of sqrt(x), so it is the correctly rounded root except where sqrt(x) lies
// that close to a rounding boundary.  Used where the square root feeds FP32 arithmetic (the
// mixed path's observer time, whose weights are FP32): there the second correction changes
// nothing measurable.
__device__ __forceinline__ double sqrt_pos1(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = 0.5 * y;
  const double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  const double d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

// Reciprocal of a nonzero normal double to about 1 ulp: hardware estimate plus two
// Newton-Raphson steps e = 1 - x y, y += y e.
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  return y;
}

}  // namespace acum
