// C-ABI of the broadband (BPM) path -- placeholder until the BPM kernels land.
#include "abi_internal.hpp"

namespace acum {
void bpm_release(acum_ctx*) {}
}  // namespace acum

extern "C" int acum_bpm(acum_ctx* ctx, const acum_bpm_case* c, float* BBspl, float* BBoaspl) {
  (void)c;
  (void)BBspl;
  (void)BBoaspl;
  return acum::ctx_fail(ctx, ACUM_ERR_ARG, "broadband path not built yet");
}
