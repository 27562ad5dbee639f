// Tonal post-processing on the GPU: band levels and the overall level of the acoustic pressure
// history pA [nT][nObs] (time-major, as the tonal path writes it, core.cu:202).
//
//   spectra_bins_kernel   -- one wave per 64 observers and kBins DFT bins: the observer's mean is
//                            removed and each bin's DFT is formed by the Goertzel-Reinsch
//                            recurrence in FP64.  Goertzel's s_n = x_n + 2 cos(w) s_{n-1} -
//                            s_{n-2} loses O(N^2 eps) near w = 0 and w = pi; Reinsch carries
//                            d_n = s_n - sigma s_{n-1} instead:
//                              d_n = sigma d_{n-1} + lambda s_{n-1} + x_n,  s_n = d_n + sigma s_{n-1}
//                            with sigma = +1, lambda = -4 sin^2(w/2) for w <= pi/2 and
//                            sigma = -1, lambda = 4 cos^2(w/2) above, so the error is O(N eps);
//                            |X_k|^2 = d_{N-1}^2 - lambda s_{N-1} s_{N-2}.  It stores the one-sided
//                            mean-square share w_k |X_k|^2 / nT^2 (w_k = 2, or 1 at k = nT/2),
//                            whose sum over k = 1 .. nT/2 is the variance (Parseval).
//   spectra_bands_kernel  -- one thread per observer: the variance (two passes, as the host
//                            overall level does), the band sums over contiguous bin ranges in bin
//                            order, and the levels in dB re 20 uPa.
// The DFT is direct (three FP64 FMAs per sample and bin, 1.5 nT^2 per observer): at 2048 steps
// and 4096 observers that is 2.6e10 FMAs, a few milliseconds, with no FFT size restriction.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace acum {
namespace {

constexpr int kBins = 16;                     // DFT bins per thread
constexpr double kPref2 = 2e-5 * 2e-5;        // reference pressure squared, Pa^2
constexpr double kFloor = 1e-10 * 2e-5 * 2e-5;  // -100 dB

__global__ __launch_bounds__(64) void spectra_bins_kernel(const float* __restrict__ pA, int nT, int nObs,
                                                          int nbins, double* __restrict__ E) {
  const int o = blockIdx.x * 64 + threadIdx.x;
  const int oc = o < nObs ? o : nObs - 1;
  const int k0 = 1 + blockIdx.y * kBins;
  double mean = 0.0;
  for (int n = 0; n < nT; n++) mean += pA[(int64_t)n * nObs + oc];
  mean /= (double)nT;
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
  for (int n = 0; n < nT; n++) {
    const double x = (double)pA[(int64_t)n * nObs + oc] - mean;
#pragma unroll
    for (int j = 0; j < kBins; j++) {
      d[j] = fma(lam[j], sn[j], fma(sg[j], d[j], x));
      sn[j] = fma(sg[j], sn[j], d[j]);
    }
  }
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

__global__ void spectra_bands_kernel(const float* __restrict__ pA, int nT, int nObs, const double* __restrict__ E,
                                     const int* __restrict__ lo, const int* __restrict__ hi, int nBands,
                                     float* __restrict__ spl, float* __restrict__ oaspl) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nObs) return;
  double mean = 0.0;
  for (int n = 0; n < nT; n++) mean += pA[(int64_t)n * nObs + o];
  mean /= (double)nT;
  double ms = 0.0;
  for (int n = 0; n < nT; n++) {
    const double d = pA[(int64_t)n * nObs + o] - mean;
    ms += d * d;
  }
  ms /= (double)nT;
  oaspl[o] = (float)(10.0 * log10(ms / kPref2));
  for (int b = 0; b < nBands; b++) {
    double e = 0.0;
    for (int k = lo[b]; k < hi[b]; k++) e += E[(int64_t)(k - 1) * nObs + o];
    spl[(int64_t)o * nBands + b] = (float)(10.0 * log10(fmax(e, kFloor) / kPref2));
  }
}

}  // namespace

int spectra_bins_per_thread() { return kBins; }

hipError_t launch_spectra(hipStream_t s, const float* pA, int nT, int nObs, double* E, const int* lo, const int* hi,
                          int nBands, float* spl, float* oaspl) {
  const int nbins = nT / 2;
  if (nbins >= 1) {
    const dim3 g((nObs + 63) / 64, (nbins + kBins - 1) / kBins);
    hipLaunchKernelGGL(spectra_bins_kernel, g, dim3(64), 0, s, pA, nT, nObs, nbins, E);
  }
  hipLaunchKernelGGL(spectra_bands_kernel, dim3((nObs + 127) / 128), dim3(128), 0, s, pA, nT, nObs, E, lo, hi,
                     nBands, spl, oaspl);
  return hipGetLastError();
}

}  // namespace acum
