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
}  // namespace acum
