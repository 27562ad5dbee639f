// C-ABI of the tonal band spectra (post-processing of pA; no counterpart in the reference).
#include <hip/hip_runtime.h>
#include <cmath>
#include <string>
#include <vector>
#include "abi_internal.hpp"
#include "spectra_types.hpp"

namespace acum {
hipError_t launch_spectra(hipStream_t s, const float* pA, int nT, int nObs, double* E, const int* lo, const int* hi,
                          int nBands, float* spl, float* oaspl);
hipError_t launch_spectra_scratch(int device, hipStream_t s, const float* pA, int nT, int nObs,
                                  const BandTable& tab, int nBands, float* spl, float* oaspl);

namespace {

constexpr double kSixthDown = 0.8908987181403393;  // 2^(-1/6)
constexpr double kSixthUp = 1.122462048309373;     // 2^(1/6)

}  // namespace

// Bin range [lo_b, hi_b) of every band: bin k (frequency k df, df = 1 / (nT dt)) belongs to band
// b when edge_b <= k df < edge_{b+1}.
// The edges are computed one at a time, so no buffer is needed (the allocation-free entry point
// fills fixed arrays with it).
void spectra_band_bins(int64_t nT, double dt, int nBands, const float* fc, int* lo, int* hi) {
  const int nbins = (int)(nT / 2);
  const double df = 1.0 / ((double)nT * dt);
  auto edge = [&](int b) {
    if (b == 0) return (double)fc[0] * kSixthDown;
    if (b == nBands) return (double)fc[nBands - 1] * kSixthUp;
    return std::sqrt((double)fc[b - 1] * (double)fc[b]);
  };
  auto first_at_or_above = [&](double f) {
    int k = 1;
    while (k <= nbins && (double)k * df < f) k++;
    return k;
  };
  for (int b = 0; b < nBands; b++) {
    lo[b] = first_at_or_above(edge(b));
    hi[b] = first_at_or_above(edge(b + 1));
  }
}

namespace {

int check_spectra_args(acum_ctx* ctx, int64_t nT, int64_t nObs, double dt, int nBands, const float* fc) {
  if (nT < 2 || nT > 65536 || nObs < 1 || nObs > (1 << 26) || nBands < 1 || !fc)
    return ctx_fail(ctx, ACUM_ERR_ARG, "spectra need 2 <= nTime <= 65536, nObs >= 1, nBands >= 1 and outputs");
  if (!(dt > 0) || !std::isfinite(dt)) return ctx_fail(ctx, ACUM_ERR_ARG, "dt must be positive");
  for (int b = 0; b < nBands; b++)
    if (!(fc[b] > 0) || !std::isfinite(fc[b]) || (b > 0 && !(fc[b] > fc[b - 1])))
      return ctx_fail(ctx, ACUM_ERR_ARG, "band centres must be positive and strictly increasing");
  return ACUM_OK;
}

}  // namespace

int spectra_device(acum_ctx* ctx, const float* d_pA, int64_t nT, int64_t nObs, double dt, int nBands,
                   const float* fc, float* spl, float* oaspl) {
  if (!spl || !oaspl) return ctx_fail(ctx, ACUM_ERR_ARG, "null spectra output");
  int rc = check_spectra_args(ctx, nT, nObs, dt, nBands, fc);
  if (rc) return rc;
  std::vector<int> lo(nBands), hi(nBands);
  spectra_band_bins(nT, dt, nBands, fc, lo.data(), hi.data());
  const int nbins = (int)(nT / 2);
  hipStream_t s = ctx_stream(ctx);
  double* d_E;
  int *d_lo, *d_hi;
  float *d_spl, *d_oa;
#define CHK(call)                                                                             \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return ctx_fail(ctx, e_ == hipErrorOutOfMemory ? ACUM_ERR_NOMEM : ACUM_ERR_DEVICE,      \
                      std::string(#call) + ": " + hipGetErrorString(e_));                     \
  } while (0)
  WorkspaceLayout L;
  L.add(&d_E, (size_t)(nbins + 1) * nObs);  // bin energies, then the observer means
  L.add(&d_lo, nBands);
  L.add(&d_hi, nBands);
  L.add(&d_spl, (size_t)nObs * nBands);
  L.add(&d_oa, nObs);
  void* ws = nullptr;
  CHK(ctx_workspace(ctx, L.bytes(), &ws));
  L.assign(ws);
  CHK(hipMemcpyAsync(d_lo, lo.data(), nBands * sizeof(int), hipMemcpyHostToDevice, s));
  CHK(hipMemcpyAsync(d_hi, hi.data(), nBands * sizeof(int), hipMemcpyHostToDevice, s));
  CHK(hipEventRecord(ctx_event(ctx, 0), s));
  CHK(launch_spectra(s, d_pA, (int)nT, (int)nObs, d_E, d_lo, d_hi, nBands, d_spl, d_oa));
  CHK(hipEventRecord(ctx_event(ctx, 1), s));
  CHK(hipMemcpyAsync(spl, d_spl, (size_t)nObs * nBands * sizeof(float), hipMemcpyDeviceToHost, s));
  CHK(hipMemcpyAsync(oaspl, d_oa, (size_t)nObs * sizeof(float), hipMemcpyDeviceToHost, s));
  CHK(hipStreamSynchronize(s));
  float ms = 0;
  CHK(hipEventElapsedTime(&ms, ctx_event(ctx, 0), ctx_event(ctx, 1)));
#undef CHK
  acum_timing* t = ctx_timing(ctx);
  *t = acum_timing();
  t->eval_ms = ms;
  t->total_ms = ms;
  t->evals = (int64_t)nbins * nT * nObs;
  t->launches = 3;
  return ACUM_OK;
}

}  // namespace acum

using namespace acum;

extern "C" int acum_tonal_spectra(acum_ctx* ctx, int64_t nTime, int64_t nObs, const float* pA, double dt,
                                  int32_t nBands, const float* fc, float* spl, float* oaspl) {
  if (!ctx) return ACUM_ERR_ARG;
  if (!pA) return ctx_fail(ctx, ACUM_ERR_ARG, "null pA");
  if (nTime < 2 || nTime > 65536 || nObs < 1 || nObs > (1 << 26))
    return ctx_fail(ctx, ACUM_ERR_ARG, "spectra need 2 <= nTime <= 65536 and nObs >= 1");
  if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return ctx_fail(ctx, ACUM_ERR_DEVICE, "hipSetDevice failed");
  float* d_pA = nullptr;
  const size_t bytes = (size_t)nTime * nObs * sizeof(float);
  hipError_t e = hipMalloc(&d_pA, bytes);
  if (e != hipSuccess) return ctx_fail(ctx, ACUM_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemcpy(d_pA, pA, bytes, hipMemcpyHostToDevice);
  int rc = (e == hipSuccess) ? spectra_device(ctx, d_pA, nTime, nObs, dt, nBands, fc, spl, oaspl)
                             : ctx_fail(ctx, ACUM_ERR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
  (void)hipFree(d_pA);
  return rc;
}

// Allocation-free form: device pA, device outputs, the library's fixed device scratch.
extern "C" int acum_tonal_spectra_noalloc(acum_ctx* ctx, int64_t nTime, int64_t nObs, const float* d_pA, double dt,
                                          int32_t nBands, const float* fc, float* d_spl, float* d_oaspl) {
  if (!ctx) return ACUM_ERR_ARG;
  if (!d_pA || !d_spl || !d_oaspl) return ctx_fail(ctx, ACUM_ERR_ARG, "null device pointer");
  int rc = check_spectra_args(ctx, nTime, nObs, dt, nBands, fc);
  if (rc) return rc;
  if (nBands > kMaxBandsNoAlloc)
    return ctx_fail(ctx, ACUM_ERR_ARG, "acum_tonal_spectra_noalloc takes at most 128 bands");
  const size_t need = (size_t)(nTime / 2 + 1) * (size_t)nObs * sizeof(double);
  if (need > kSpectraScratchBytes)
    return ctx_fail(ctx, ACUM_ERR_ARG,
                    "acum_tonal_spectra_noalloc: (nTime/2 + 1) * nObs doubles exceed its 64 MiB device scratch");
  BandTable tab;
  spectra_band_bins(nTime, dt, nBands, fc, tab.lo, tab.hi);
  if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return ctx_fail(ctx, ACUM_ERR_DEVICE, "hipSetDevice failed");
  hipStream_t s = ctx_stream(ctx);
  hipError_t e = hipEventRecord(ctx_event(ctx, 0), s);
  if (e == hipSuccess)
    e = launch_spectra_scratch(ctx_device(ctx), s, d_pA, (int)nTime, (int)nObs, tab, nBands, d_spl, d_oaspl);
  if (e == hipSuccess) e = hipEventRecord(ctx_event(ctx, 1), s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx_event(ctx, 0), ctx_event(ctx, 1));
  if (e != hipSuccess) return ctx_fail(ctx, ACUM_ERR_DEVICE, std::string("spectra (no allocation): ") + hipGetErrorString(e));
  acum_timing* t = ctx_timing(ctx);
  *t = acum_timing();
  t->eval_ms = ms;
  t->total_ms = ms;
  t->evals = (int64_t)(nTime / 2) * nTime * nObs;
  t->launches = 3;
  return ACUM_OK;
}
