// Test-only probe of the FP64 helpers the tonal kernels use (acum3_periodic_amd/csrc/
// fp64_math.hpp): evaluates them on the device for arrays of arguments so tests can compare with
// the host's correctly rounded results.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "../../acum3_periodic_amd/csrc/fp64_math.hpp"

__global__ void probe_kernel(const double* __restrict__ x, double* __restrict__ s,
                             double* __restrict__ r, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    s[i] = acum::sqrt_pos(x[i]);
    r[i] = acum::rcp_nr(x[i]);
  }
}

// XCC_ID hardware register (bits 3:0) seen by each workgroup, the value the eval kernel uses to
// pick its work queue.
__global__ void xcc_kernel(int* __restrict__ out, int n) {
  if (threadIdx.x == 0 && (int)blockIdx.x < n)
    out[blockIdx.x] = (int)(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 15);
}

extern "C" int probe_xcc(int* ids, int nblocks) {
  int* d = nullptr;
  hipError_t e = hipMalloc(&d, nblocks * sizeof(int));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(xcc_kernel, dim3(nblocks), dim3(64), 0, 0, d, nblocks);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(ids, d, nblocks * sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return (int)e;
}

// x: host array of n positive doubles; s, r: host outputs (sqrt, reciprocal).  Returns 0 on
// success, a HIP error code otherwise.
extern "C" int probe_fp64(const double* x, double* s, double* r, int64_t n) {
  double *dx = nullptr, *ds = nullptr, *dr = nullptr;
  hipError_t e = hipMalloc(&dx, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&ds, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dr, n * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dx, ds, dr, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(s, ds, n * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(r, dr, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  (void)hipFree(ds);
  (void)hipFree(dr);
  return (int)e;
}
