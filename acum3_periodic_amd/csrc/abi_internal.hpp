// Internal interface between the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include "../../include/acum.h"
#include "dev_types.hpp"

namespace acum {
int ctx_fail(acum_ctx* c, int code, const std::string& msg);
hipStream_t ctx_stream(acum_ctx* c);
hipEvent_t ctx_event(acum_ctx* c, int i);
acum_timing* ctx_timing(acum_ctx* c);
int ctx_device(acum_ctx* c);
void bpm_release(acum_ctx* c);
// band spectra of a device-resident pA [nT][nObs] into host outputs (abi_spectra.cpp)
int spectra_device(acum_ctx* ctx, const float* d_pA, int64_t nT, int64_t nObs, double dt, int nBands,
                   const float* fc, float* spl, float* oaspl);
}  // namespace acum
