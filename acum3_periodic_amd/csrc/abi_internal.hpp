// Internal interface between the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <string>
#include "../../include/acum.h"
#include "dev_types.hpp"

namespace acum {
int ctx_fail(acum_ctx* c, int code, const std::string& msg);
hipStream_t ctx_stream(acum_ctx* c);
hipEvent_t ctx_event(acum_ctx* c, int i);
acum_timing* ctx_timing(acum_ctx* c);
int ctx_device(acum_ctx* c);
// Decay rate k of an exponential size profile (density e^(-kx) over x in [0, 1]) for work split
// into pieces that run about n to a workgroup slot, largest first: the first piece, k / (1 - e^-k)
// times the mean, is held to 0.8 n mean pieces, about one slot's share; equal sizes (k = 0) up to
// 1.25 pieces per slot.  Cutting the work this way lets a launch end on small pieces.
inline double size_decay(double n) {
  const double r = 0.8 * n;
  if (r <= 1.0) return 0.0;
  double lo = 1e-6, hi = 40.0;
  for (int i = 0; i < 60; i++) {
    const double m = 0.5 * (lo + hi);
    if (m / -std::expm1(-m) < r) lo = m;
    else hi = m;
  }
  return 0.5 * (lo + hi);
}
// Fraction of the work in the first x of the pieces under that profile.
inline double size_profile(double k, double x) { return k > 0.0 ? std::expm1(-k * x) / std::expm1(-k) : x; }

// A device workspace owned by the context for the buffers of one call (broadband, spectra):
// grown on demand, kept until the context is destroyed.  Calls on one context are serialised
// and each synchronises its stream before returning, so one call's buffers never outlive it.
hipError_t ctx_workspace(acum_ctx* c, size_t bytes, void** out);
// Host <-> device copies of large pageable buffers through the context's pinned staging buffers
// (abi_tonal.cpp); the upload is queued on the context's stream, the download waits for its data.
hipError_t ctx_upload(acum_ctx* c, void* dst, const void* src, size_t n);
hipError_t ctx_download(acum_ctx* c, void* dst, const void* src, size_t n);

// Lays out a call's device buffers (at most 32) in one workspace, each at a 256-byte boundary.
struct WorkspaceLayout {
  struct Req {
    void** ptr;
    size_t bytes;
  };
  Req req[32];
  int n = 0;
  template <typename T>
  void add(T** p, size_t count) {
    req[n++] = Req{(void**)p, (count ? count : 1) * sizeof(T)};
  }
  static size_t round(size_t b) { return (b + 255) / 256 * 256; }
  size_t bytes() const {
    size_t b = 0;
    for (int i = 0; i < n; i++) b += round(req[i].bytes);
    return b;
  }
  void assign(void* base) const {
    char* q = (char*)base;
    for (int i = 0; i < n; i++) {
      *req[i].ptr = q;
      q += round(req[i].bytes);
    }
  }
};
// band spectra of a device-resident pA [nT][nObs] into host outputs (abi_spectra.cpp)
int spectra_device(acum_ctx* ctx, const float* d_pA, int64_t nT, int64_t nObs, double dt, int nBands,
                   const float* fc, float* spl, float* oaspl);
}  // namespace acum
