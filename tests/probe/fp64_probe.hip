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

// One wave of v_mfma_f32_4x4x1f32 (16 blocks of 4x4, K = 1), accumulated over three products:
// d = a0 b0 + a1 b1 + a2 b2 with lane l supplying a_k[l] and b_k[l]; out[l][v] = d[v] of lane l.
// The test checks the lane layout the tonal kernel relies on.
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ void mfma_kernel(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out) {
  const int l = threadIdx.x;
  v4f d = {0.0f, 0.0f, 0.0f, 0.0f};
  d = __builtin_amdgcn_mfma_f32_4x4x1f32(a[l], b[l], d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_4x4x1f32(a[64 + l], b[64 + l], d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_4x4x1f32(a[128 + l], b[128 + l], d, 0, 0, 0);
  for (int v = 0; v < 4; v++) out[4 * l + v] = d[v];
}

// a, b: host arrays of 3 x 64 floats; out: 64 x 4 floats.
extern "C" int probe_mfma4x4(const float* a, const float* b, float* out) {
  float *da = nullptr, *db = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&da, 192 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&db, 192 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dout, 256 * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(da, a, 192 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, b, 192 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(mfma_kernel, dim3(1), dim3(64), 0, 0, da, db, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dout, 256 * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dout);
  return (int)e;
}

// One wave of rec_products64: lane l holds field f[l] (each 16-lane row its own record) and its
// own r = (r[l], r[64 + l], r[128 + l]) and delay d[l]; out[l][0..4] = rm, nVd, dL, nLd, cT of
// lane l.  The test checks the row broadcast the FP64 tonal kernel relies on.
__global__ void rec64_kernel(const double* __restrict__ f, const double* __restrict__ r,
                             const double* __restrict__ d, double* __restrict__ out) {
  const int l = threadIdx.x;
  double rm = d[l], nVd, dL = 0.0, nLd, cT;
  acum::rec_products64(f[l], r[l], r[64 + l], r[128 + l], rm, nVd, dL, nLd, cT);
  out[5 * l] = rm;
  out[5 * l + 1] = nVd;
  out[5 * l + 2] = dL;
  out[5 * l + 3] = nLd;
  out[5 * l + 4] = cT;
}

// f: 64 doubles, r: 3 x 64, d: 64; out: 64 x 5.
extern "C" int probe_rec64(const double* f, const double* r, const double* d, double* out) {
  double* dv = nullptr;  // f | r | d | out
  hipError_t e = hipMalloc(&dv, (64 + 192 + 64 + 320) * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(dv, f, 64 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dv + 64, r, 192 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dv + 256, d, 64 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(rec64_kernel, dim3(1), dim3(64), 0, 0, dv, dv + 64, dv + 256, dv + 320);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dv + 320, 320 * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dv);
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
