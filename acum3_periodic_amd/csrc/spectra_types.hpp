// Types shared by the spectra kernels and the spectra C-ABI.
#pragma once
#include <cstddef>

namespace acum {

// Most bands of one call of the allocation-free spectra entry point (its band table travels as a
// kernel argument).
constexpr int kMaxBandsNoAlloc = 128;
// Device scratch of that entry point: (nT / 2 + 1) nObs doubles must fit (64 MiB, e.g. 2048
// steps x 8000 observers, or 4096 steps x 4000 observers).
constexpr size_t kSpectraScratchBytes = (size_t)64 << 20;

// Bin range [lo_b, hi_b) of each band, by value.
struct BandTable {
  int lo[kMaxBandsNoAlloc];
  int hi[kMaxBandsNoAlloc];
};

}  // namespace acum
