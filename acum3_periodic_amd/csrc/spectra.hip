// Tonal post-processing on the GPU: band levels and the overall level of the acoustic pressure
// history pA [nT][nObs] (time-major, as the tonal path writes it, core.cu:202).
//
//   spectra_stats_kernel  -- a workgroup of four waves per 64 observers: each wave sums a quarter
//                            of the steps (coalesced over observers, sixteen steps' loads in
//                            flight), the quarters are added in fixed order for the mean, and
//                            again for the variance (two passes, as the host overall level does);
//                            writes the mean and the overall level in dB re 20 uPa.
//   spectra_bins_kernel   -- one wave per 64 observers and kBins DFT bins: each bin's DFT of the
//                            mean-removed column is formed by the Goertzel-Reinsch recurrence in
//                            FP64.  Goertzel's s_n = x_n + 2 cos(w) s_{n-1} - s_{n-2} loses
//                            O(N^2 eps) near w = 0 and w = pi; Reinsch carries d_n = s_n - sigma
//                            s_{n-1} instead:
//                              d_n = sigma d_{n-1} + lambda s_{n-1} + x_n,  s_n = d_n + sigma s_{n-1}
//                            with sigma = +1, lambda = -4 sin^2(w/2) for w <= pi/2 and
//                            sigma = -1, lambda = 4 cos^2(w/2) above, so the error is O(N eps);
//                            |X_k|^2 = d_{N-1}^2 - lambda s_{N-1} s_{N-2}.  It stores the one-sided
//                            mean-square share w_k |X_k|^2 / nT^2 (w_k = 2, or 1 at k = nT/2),
//                            whose sum over k = 1 .. nT/2 is the variance (Parseval).  Sixteen
//                            samples are loaded before their recurrence steps.
//   spectra_bands_kernel  -- a thread per (observer, band): the band's bin energies summed in bin
//                            order and the level in dB re 20 uPa.
// The DFT is direct (three FP64 FMAs per sample and bin, 1.5 nT^2 per observer): at 2048 steps
// and 4096 observers that is 2.6e10 FMAs, with no FFT size restriction.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <mutex>
#include "spectra_types.hpp"

namespace acum {
namespace {

constexpr int kBins = 8;                      // DFT bins per thread
constexpr int kChunk = 16;                    // samples loaded ahead of their use
constexpr int kParts = 4;                     // step partitions of the statistics kernel
constexpr double kPref2 = 2e-5 * 2e-5;        // reference pressure squared, Pa^2
constexpr double kFloor = 1e-10 * 2e-5 * 2e-5;  // -100 dB

// sum over steps [n0, n1) of f(pA[n][o]) in step order
template <typename F>
__device__ __forceinline__ double column_sum(const float* __restrict__ pA, int nObs, int o, int n0, int n1, F f) {
  double s = 0.0;
  int n = n0;
  for (; n + kChunk <= n1; n += kChunk) {
    float v[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; u++) v[u] = pA[(int64_t)(n + u) * nObs + o];
#pragma unroll
    for (int u = 0; u < kChunk; u++) s += f(v[u]);
  }
  for (; n < n1; n++) s += f(pA[(int64_t)n * nObs + o]);
  return s;
}

__global__ __launch_bounds__(64 * kParts) void spectra_stats_kernel(const float* __restrict__ pA, int nT, int nObs,
                                                                    double* __restrict__ mean,
                                                                    float* __restrict__ oaspl) {
  const int x = threadIdx.x & 63, y = threadIdx.x >> 6;
  const int o = blockIdx.x * 64 + x;
  const int oc = o < nObs ? o : nObs - 1;
  const int n0 = (int)((int64_t)nT * y / kParts), n1 = (int)((int64_t)nT * (y + 1) / kParts);
  __shared__ double part[kParts][64];
  __shared__ double mean_s[64];
  part[y][x] = column_sum(pA, nObs, oc, n0, n1, [](float v) { return (double)v; });
  __syncthreads();
  if (y == 0) {
    double m = part[0][x];
    for (int p = 1; p < kParts; p++) m += part[p][x];
    mean_s[x] = m / (double)nT;
  }
  __syncthreads();
  const double m = mean_s[x];
  part[y][x] = column_sum(pA, nObs, oc, n0, n1, [m](float v) {
    const double d = (double)v - m;
    return d * d;
  });
  __syncthreads();
  if (y == 0 && o < nObs) {
    double ms = part[0][x];
    for (int p = 1; p < kParts; p++) ms += part[p][x];
    ms /= (double)nT;
    mean[o] = m;
    oaspl[o] = (float)(10.0 * log10(ms / kPref2));
  }
}

__global__ __launch_bounds__(64) void spectra_bins_kernel(const float* __restrict__ pA, const double* __restrict__ mean,
                                                          int nT, int nObs, int nbins, double* __restrict__ E) {
  const int o = blockIdx.x * 64 + threadIdx.x;
  const int oc = o < nObs ? o : nObs - 1;
  const int k0 = 1 + blockIdx.y * kBins;
  const double m = mean[oc];
  double lam[kBins], sg[kBins], d[kBins], sn[kBins];
#pragma unroll
  for (int j = 0; j < kBins; j++) {
    const int k = k0 + j;
    const double h = (double)k / (double)nT;  // w / (2 pi)
    const bool low = 4 * k <= nT;
    const double t = low ? sinpi(h) : cospi(h);
    lam[j] = low ? -4.0 * t * t : 4.0 * t * t;
    sg[j] = low ? 1.0 : -1.0;
    d[j] = 0.0;
    sn[j] = 0.0;
  }
  auto step = [&](float v) __attribute__((always_inline)) {
    const double x = (double)v - m;
#pragma unroll
    for (int j = 0; j < kBins; j++) {
      d[j] = fma(lam[j], sn[j], fma(sg[j], d[j], x));
      sn[j] = fma(sg[j], sn[j], d[j]);
    }
  };
  int n = 0;
  for (; n + kChunk <= nT; n += kChunk) {
    float v[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; u++) v[u] = pA[(int64_t)(n + u) * nObs + oc];
#pragma unroll
    for (int u = 0; u < kChunk; u++) step(v[u]);
  }
  for (; n < nT; n++) step(pA[(int64_t)n * nObs + oc]);
  if (o >= nObs) return;
  const double inv = 1.0 / ((double)nT * (double)nT);
#pragma unroll
  for (int j = 0; j < kBins; j++) {
    const int k = k0 + j;
    if (k > nbins) break;
    const double s2 = sg[j] * (sn[j] - d[j]);  // s_{N-2}
    const double x2 = fma(d[j], d[j], -lam[j] * sn[j] * s2);
    const double w = (2 * k == nT) ? 1.0 : 2.0;
    E[(int64_t)(k - 1) * nObs + o] = w * fmax(x2, 0.0) * inv;
  }
}

// The band table [lo_b, hi_b) comes from device memory (BandPtrs) or as a kernel argument
// (BandTable, the allocation-free entry point); the sums are the same either way.
struct BandPtrs {
  const int* lo;
  const int* hi;
};
__device__ __forceinline__ int band_lo(const BandPtrs& t, int b) { return t.lo[b]; }
__device__ __forceinline__ int band_hi(const BandPtrs& t, int b) { return t.hi[b]; }
__device__ __forceinline__ int band_lo(const BandTable& t, int b) { return t.lo[b]; }
__device__ __forceinline__ int band_hi(const BandTable& t, int b) { return t.hi[b]; }

template <typename Table>
__global__ __launch_bounds__(64) void spectra_bands_kernel(const double* __restrict__ E, int nObs, Table tab,
                                                           int nBands, float* __restrict__ spl) {
  const int o = blockIdx.x * 64 + threadIdx.x;
  if (o >= nObs) return;
  for (int b = blockIdx.y; b < nBands; b += gridDim.y) {
    const int k0 = band_lo(tab, b), k1 = band_hi(tab, b);
    double e = 0.0;
    int k = k0;
    for (; k + kChunk <= k1; k += kChunk) {
      double v[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; u++) v[u] = E[(int64_t)(k + u - 1) * nObs + o];
#pragma unroll
      for (int u = 0; u < kChunk; u++) e += v[u];
    }
    for (; k < k1; k++) e += E[(int64_t)(k - 1) * nObs + o];
    spl[(int64_t)o * nBands + b] = (float)(10.0 * log10(fmax(e, kFloor) / kPref2));
  }
}

template <typename Table>
hipError_t launch_all(hipStream_t s, const float* pA, int nT, int nObs, double* E, const Table& tab, int nBands,
                      float* spl, float* oaspl) {
  const int nbins = nT / 2;
  const int og = (nObs + 63) / 64;
  double* mean = E + (int64_t)nbins * nObs;
  hipLaunchKernelGGL(spectra_stats_kernel, dim3(og), dim3(64 * kParts), 0, s, pA, nT, nObs, mean, oaspl);
  if (nbins >= 1) {
    const dim3 g(og, (nbins + kBins - 1) / kBins);
    hipLaunchKernelGGL(spectra_bins_kernel, g, dim3(64), 0, s, pA, mean, nT, nObs, nbins, E);
  }
  const dim3 gb(og, nBands < 65535 ? nBands : 65535);
  hipLaunchKernelGGL(spectra_bands_kernel<Table>, gb, dim3(64), 0, s, E, nObs, tab, nBands, spl);
  return hipGetLastError();
}

// Scratch of the allocation-free entry point (acum_tonal_spectra_noalloc): the bin energies and
// observer means of one call, kSpectraScratchBytes per GPU, held from the library's first
// context on that GPU until the process ends.  Host code never uses this name as an address:
// the device address of each GPU's copy comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[kSpectraScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

// E holds nbins * nObs bin energies followed by nObs observer means.
hipError_t launch_spectra(hipStream_t s, const float* pA, int nT, int nObs, double* E, const int* lo, const int* hi,
                          int nBands, float* spl, float* oaspl) {
  return launch_all(s, pA, nT, nObs, E, BandPtrs{lo, hi}, nBands, spl, oaspl);
}

// Resolves the device address of the scratch array on `device`, which must be the current
// device; called by acum_ctx_create for each context's device.
hipError_t spectra_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = p;
  return hipSuccess;
}

// The same kernels on the scratch array of `device` (resolved by spectra_scratch_init) with the
// band table passed by value: no allocation and no copy.  The caller has checked that
// (nT / 2 + 1) nObs doubles fit kSpectraScratchBytes and that nBands <= kMaxBandsNoAlloc.
hipError_t launch_spectra_scratch(int device, hipStream_t s, const float* pA, int nT, int nObs,
                                  const BandTable& tab, int nBands, float* spl, float* oaspl) {
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device >= 0 && device < kMaxDevices) base = g_scratch[device];
  }
  if (!base) return hipErrorNotInitialized;
  return launch_all(s, pA, nT, nObs, (double*)base, tab, nBands, spl, oaspl);
}

}  // namespace acum
