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

// The square root from the hardware estimate and one residual correction, without the Goldschmidt
// step: with y = 1/sqrt(x) (1 + e), g = x y and h = y/2, the result g + (x - g^2) h is
// sqrt(x) (1 - 3e^2/2 + O(e^3)), so its relative error is about 1.5 e^2.  Measured over squared
// distances 1e-6 .. 1e12 on gfx950: |e| <= 5.3e-8 = 2^-24.2 and the result is within
// 4.2e-15 = 2^-47.8 of sqrt(x).  Used where the square root feeds FP32 arithmetic (the mixed
// path's observer time, whose tent weights are FP32 and so carry errors near 2^-24 of a bin):
// 4.2e-15 of the delay, 4e-11 bins at a delay of 10^4 bins, is far below that.
__device__ __forceinline__ double sqrt_pos_fast(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double g = x * y;
  const double h = 0.5 * y;
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

// The FP64 tonal path's record products (tonal.hip).  From the lane's field f of a 16-double
// record held one field per lane of each 16-lane row (lane l holds field l % 16; fields 0-11 the
// rows v, vd, l, ld, 12-14 aM1, aLM, cT) and r = (rx, ry, rz): rm -= v.r (rm enters as the delay),
// nVd = aM1 - vd.r, dL += l.r (dL enters as zero), nLd = aLM - ld.r and cT, each product
// accumulated x, then y, then z, one rounding per FMA.  `s_nop 4` gives the wait states a DPP
// operand needs after a VALU write of its register or of EXEC.
__device__ __forceinline__ void rec_products64(double f, double rx, double ry, double rz, double& rm,
                                               double& nVd, double& dL, double& nLd, double& cT) {
  __asm__(
      "s_nop 4\n\t"
      "v_mov_b64_dpp %1, %5 row_newbcast:12 row_mask:0xf bank_mask:0xf\n\t"
      "v_mov_b64_dpp %3, %5 row_newbcast:13 row_mask:0xf bank_mask:0xf\n\t"
      "v_mov_b64_dpp %4, %5 row_newbcast:14 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, -%5, %6 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%5, %6 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %5, %6 row_newbcast:6 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, -%5, %6 row_newbcast:9 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, -%5, %7 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%5, %7 row_newbcast:4 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %5, %7 row_newbcast:7 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, -%5, %7 row_newbcast:10 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, -%5, %8 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%5, %8 row_newbcast:5 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %5, %8 row_newbcast:8 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, -%5, %8 row_newbcast:11 row_mask:0xf bank_mask:0xf"
      : "+v"(rm), "=&v"(nVd), "+v"(dL), "=&v"(nLd), "=&v"(cT)
      : "v"(f), "v"(rx), "v"(ry), "v"(rz));
}

}  // namespace acum
