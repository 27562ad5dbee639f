// preprocess_imper (preprocess.cu:228-413) on the GPU, in FP64, bit-identical to the host form in
// host/preprocess.cpp: the same operations in the same order, no contraction.  One workgroup per
// element; its series live in LDS as doubles while they are filtered and differentiated:
//   P = CP * 0.5 rhoRef Mref^2 a0^2 (preprocess.cu:279), the periodic cubic correction of
//   make_periodic (mathfunc.cu:165-206), optionally two rounds of 3-point median and 13-point
//   Gaussian (mathfunc.cu:208-421), dP/dt by periodic central (or end-biased) differences
//   (mathfunc.cu:132-159); dn/dt, V = d(x C)/dt and dV/dt by periodic central differences
//   (preprocess.cu:351-364); x, y, z scaled by C and dS by C^2 (preprocess.cu:275-278).
// The element arrays are updated in place (AoS of the reference, one element's steps contiguous).
#include <hip/hip_runtime.h>
#include <cstdint>
#include "prep_types.hpp"

namespace acum {

constexpr int kPrepThreads = 256;

__device__ __forceinline__ double central_at(const double* u, int i, int n, double dt) {
  const int ip = (i + 1 < n) ? i + 1 : 0, im = (i > 0) ? i - 1 : n - 1;
  return (u[ip] - u[im]) / (2.0 * dt);
}

__global__ __launch_bounds__(kPrepThreads) void preprocess_kernel(PrepArgs a) {
  __shared__ double p[kPrepMaxSteps];
  __shared__ double t[kPrepMaxSteps];
  const int n = a.nT;
  const int tid = threadIdx.x;
  for (int64_t e = blockIdx.x; e < a.nE; e += gridDim.x) {
    float* Xe = a.X + e * (int64_t)n * 4;
    float* Ne = a.N + e * (int64_t)n * 6;
    float* Ve = a.V + e * (int64_t)n * 6;
    float* Pe = a.PR + e * (int64_t)n * 2;
    // ---- pressure
    for (int k = tid; k < n; k += kPrepThreads) p[k] = (double)Pe[2 * k] * a.pscale;
    __syncthreads();
    {  // periodic cubic correction; f2 = (2 f0 - f1, f0 .. f[n-1]) as in the host form
      const double f0 = p[0], f1 = p[1], fm1 = p[n - 2], fn = p[n - 1];
      const double e0 = f0 - (f1 - f0);
      const double k2 = fn, k1 = e0;
      const double s2 = fn - fm1, s1 = f0 - e0;
      const double xm = n;
      const double d = k1, c = s1 - s2;
      const double b = -2.0 * (s1 - s2) / xm - 3.0 * (k1 - k2) / xm / xm;
      const double aa = -b / xm - (s1 - s2) / xm / xm - (k1 - k2) / xm / xm / xm;
      __syncthreads();  // every thread has read the ends before any update
      for (int k = tid; k < n; k += kPrepThreads) {
        const double x = k + 1;
        const double cub = aa * x * x * x + b * x * x + c * x + d;
        p[k] = p[k] - (cub - k2);
      }
      __syncthreads();
    }
    if (a.lowpass == 1) {
      for (int it = 0; it < 2; it++) {
        for (int k = tid; k < n; k += kPrepThreads) {  // 3-point median, strict tests
          const double av = p[k > 0 ? k - 1 : n - 1], bv = p[k], cv = p[k + 1 < n ? k + 1 : 0];
          double hi = av > bv ? av : bv;
          hi = hi > cv ? hi : cv;
          double lo = av < bv ? av : bv;
          lo = lo < cv ? lo : cv;
          t[k] = (av < hi && av > lo) ? av : ((bv < hi && bv > lo) ? bv : cv);
        }
        __syncthreads();
        for (int k = tid; k < n; k += kPrepThreads) {  // 13-point periodic Gaussian
          double acc = 0.0;
          for (int j = 0; j < 13; j++) {
            int q = k + j - 6;
            if (q < 0) q = n + q;
            else if (q > n - 1) q = q - n;
            acc += a.g[j] * t[q];
          }
          p[k] = acc;
        }
        __syncthreads();
      }
    }
    // dP/dt
    for (int k = tid; k < n; k += kPrepThreads) {
      double pd;
      if (a.periodic == 1) {
        pd = central_at(p, k, n, a.dt);
      } else if (k == 0) {
        pd = (p[1] - p[0]) / a.dt;
      } else if (k == n - 1) {
        pd = (p[n - 1] - p[n - 2]) / a.dt;
      } else {
        pd = (p[k + 1] - p[k - 1]) / (2.0 * a.dt);
      }
      t[k] = pd;
    }
    __syncthreads();
    if (a.periodic != 1 && tid == 0) {  // cubic extrapolation of the one-sided ends
      t[0] = 3 * t[1] - 3 * t[2] + t[3];
      t[n - 1] = 3 * t[n - 2] - 3 * t[n - 3] + t[n - 4];
    }
    __syncthreads();
    for (int k = tid; k < n; k += kPrepThreads) {
      Pe[2 * k] = (float)p[k];
      Pe[2 * k + 1] = (float)t[k];
    }
    __syncthreads();
    // ---- normals, positions, velocities
    for (int dcomp = 0; dcomp < 3; dcomp++) {
      for (int k = tid; k < n; k += kPrepThreads) p[k] = (double)Ne[6 * k + dcomp];
      __syncthreads();
      for (int k = tid; k < n; k += kPrepThreads) Ne[6 * k + 3 + dcomp] = (float)central_at(p, k, n, a.dt);
      __syncthreads();
      for (int k = tid; k < n; k += kPrepThreads) p[k] = (double)Xe[4 * k + dcomp] * a.C;
      __syncthreads();
      for (int k = tid; k < n; k += kPrepThreads) t[k] = central_at(p, k, n, a.dt);
      __syncthreads();
      for (int k = tid; k < n; k += kPrepThreads) {
        Xe[4 * k + dcomp] = (float)p[k];
        Ve[6 * k + dcomp] = (float)t[k];
        Ve[6 * k + 3 + dcomp] = (float)central_at(t, k, n, a.dt);
      }
      __syncthreads();
    }
    const double C2 = a.C * a.C;
    for (int k = tid; k < n; k += kPrepThreads) Xe[4 * k + 3] = (float)((double)Xe[4 * k + 3] * C2);
    __syncthreads();
  }
}

hipError_t launch_preprocess(hipStream_t s, const PrepArgs& a, int grid) {
  hipLaunchKernelGGL(preprocess_kernel, dim3(grid), dim3(kPrepThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace acum
