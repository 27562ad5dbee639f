// preprocess_imper (preprocess.cu:228-413) on the GPU, in FP64, bit-identical to the host form in
// host/preprocess.cpp: the same operations in the same order, no contraction.  One workgroup per
// element, in three phases:
//   1. kinematics, a thread per step, straight from the element arrays: V = d(x C)/dt and
//      dV/dt from the raw positions of steps k-2 .. k+2 and dn/dt from the normals of steps
//      k-1, k+1, all by periodic central differences (preprocess.cu:351-364).  V is a separate
//      array and dn/dt fills the other half of the normal record, so nothing read here is
//      written here;
//   2. after a barrier (every raw position of the element has been read), x, y, z scaled by C
//      and dS by C^2 in place (preprocess.cu:275-278), and P = CP * 0.5 rhoRef Mref^2 a0^2
//      (preprocess.cu:279) into LDS;
//   3. the pressure series in LDS as doubles (two series of nT doubles, sized per launch:
//      16 KB at 1024 steps): the periodic cubic correction of make_periodic
//      (mathfunc.cu:165-206), optionally two rounds of 3-point median and 13-point Gaussian
//      (mathfunc.cu:208-421), dP/dt by periodic central (or end-biased) differences
//      (mathfunc.cu:132-159).
// A series value is recomputed wherever it is needed (x C, or a central difference of them) by
// the same operations on the same inputs, so it is the same double each time.
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

__device__ __forceinline__ int wrap_step(int k, int n) { return k < 0 ? k + n : (k >= n ? k - n : k); }

__global__ __launch_bounds__(kPrepThreads) void preprocess_kernel(PrepArgs a) {
  extern __shared__ double lds[];
  const int n = a.nT;
  double* p = lds;
  double* t = lds + n;
  const int tid = threadIdx.x;
  const double C2 = a.C * a.C;
  for (int64_t e = blockIdx.x; e < a.nE; e += gridDim.x) {
    float4* Xe = (float4*)(a.X + e * (int64_t)n * 4);
    float* Ne = a.N + e * (int64_t)n * 6;
    float2* Ve = (float2*)(a.V + e * (int64_t)n * 6);
    float2* Pe = (float2*)(a.PR + e * (int64_t)n * 2);
    // ---- 1. kinematics
    for (int k = tid; k < n; k += kPrepThreads) {
      const float4 xm2 = Xe[wrap_step(k - 2, n)], xm1 = Xe[wrap_step(k - 1, n)], x0 = Xe[k];
      const float4 xp1 = Xe[wrap_step(k + 1, n)], xp2 = Xe[wrap_step(k + 2, n)];
      const float* nm = Ne + 6 * wrap_step(k - 1, n);
      const float* np = Ne + 6 * wrap_step(k + 1, n);
      const float2 nm01 = *(const float2*)nm, np01 = *(const float2*)np;
      const float nm2 = nm[2], np2 = np[2];
      const float xs[5][3] = {{xm2.x, xm2.y, xm2.z}, {xm1.x, xm1.y, xm1.z}, {x0.x, x0.y, x0.z},
                              {xp1.x, xp1.y, xp1.z}, {xp2.x, xp2.y, xp2.z}};
      const float nms[3] = {nm01.x, nm01.y, nm2}, nps[3] = {np01.x, np01.y, np2};
      float v[3], vd[3], dn[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        double q[5];
#pragma unroll
        for (int j = 0; j < 5; j++) q[j] = (double)xs[j][d] * a.C;
        const double vm = (q[2] - q[0]) / (2.0 * a.dt);  // V at k-1
        const double vc = (q[3] - q[1]) / (2.0 * a.dt);  // V at k
        const double vp = (q[4] - q[2]) / (2.0 * a.dt);  // V at k+1
        v[d] = (float)vc;
        vd[d] = (float)((vp - vm) / (2.0 * a.dt));
        dn[d] = (float)(((double)nps[d] - (double)nms[d]) / (2.0 * a.dt));
      }
      Ve[3 * k] = make_float2(v[0], v[1]);
      Ve[3 * k + 1] = make_float2(v[2], vd[0]);
      Ve[3 * k + 2] = make_float2(vd[1], vd[2]);
      Ne[6 * k + 3] = dn[0];
      *(float2*)(Ne + 6 * k + 4) = make_float2(dn[1], dn[2]);
    }
    __syncthreads();  // every raw position of the element has been read
    // ---- 2. scaled positions and areas; pressure into LDS
    for (int k = tid; k < n; k += kPrepThreads) {
      const float4 x = Xe[k];
      Xe[k] = make_float4((float)((double)x.x * a.C), (float)((double)x.y * a.C), (float)((double)x.z * a.C),
                          (float)((double)x.w * C2));
      p[k] = (double)Pe[k].x * a.pscale;
    }
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
    for (int k = tid; k < n; k += kPrepThreads) Pe[k] = make_float2((float)p[k], (float)t[k]);
    __syncthreads();  // LDS is refilled by the next element
  }
}

hipError_t launch_preprocess(hipStream_t s, const PrepArgs& a, int grid) {
  const size_t shmem = 2 * sizeof(double) * (size_t)a.nT;  // at most 64 KB (nT <= kPrepMaxSteps)
  hipLaunchKernelGGL(preprocess_kernel, dim3(grid), dim3(kPrepThreads), shmem, s, a);
  return hipGetLastError();
}

}  // namespace acum
