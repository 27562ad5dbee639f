// C-ABI of the GPU preprocessing: preprocess_imper (preprocess.cu:228-413) of a run on the device.
#include <hip/hip_runtime.h>
#include <string>
#include "abi_internal.hpp"
#include "host/run.hpp"
#include "prep_types.hpp"

namespace acum {
hipError_t launch_preprocess(hipStream_t s, const PrepArgs& a, int grid);
}

using namespace acum;

namespace {

struct DevArrays {
  float* p[4] = {nullptr, nullptr, nullptr, nullptr};
  ~DevArrays() {
    for (auto q : p) (void)hipFree(q);
  }
};

}  // namespace

#define HIPCHK(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ctx_fail(ctx, e_ == hipErrorOutOfMemory ? ACUM_ERR_NOMEM : ACUM_ERR_DEVICE, \
                      std::string(#call) + ": " + hipGetErrorString(e_));                \
  } while (0)

extern "C" int acum_preprocess_gpu(acum_ctx* ctx, acum_run* r) {
  if (!ctx) return ACUM_ERR_ARG;
  if (!r) return ctx_fail(ctx, ACUM_ERR_ARG, "null run");
  if (r->preprocessed) return ctx_fail(ctx, ACUM_ERR_ARG, "run already preprocessed");
  const int n = (int)r->nTime;
  if (n < 4) return ctx_fail(ctx, ACUM_ERR_ARG, "nTime must be >= 4");
  if (n > kPrepMaxSteps)
    return ctx_fail(ctx, ACUM_ERR_ARG, "GPU preprocessing holds at most 4096 steps per series");
  int64_t nE = 0;
  for (int64_t s = 0; s < r->nSurf; s++) {
    if (r->XPtr[s] != nE * n * 4 || r->NPtr[s] != nE * n * 6 || r->PRPtr[s] != nE * n * 2)
      return ctx_fail(ctx, ACUM_ERR_ARG, "surface arrays are not contiguous");
    nE += r->E[s];
  }
  HIPCHK(hipSetDevice(ctx_device(ctx)));
  hipStream_t st = ctx_stream(ctx);
  const int64_t q = nE * n;
  const int64_t width[4] = {4, 6, 6, 2};
  float* host[4] = {r->X.data(), r->N.data(), r->V.data(), r->PR.data()};
  DevArrays d;
  for (int i = 0; i < 4; i++) HIPCHK(hipMalloc(&d.p[i], (size_t)q * width[i] * sizeof(float)));
  HIPCHK(hipEventRecord(ctx_event(ctx, 0), st));
  for (int i : {0, 1, 3})
    HIPCHK(ctx_upload(ctx, d.p[i], host[i], (size_t)q * width[i] * sizeof(float)));
  PrepArgs a{};
  a.X = d.p[0];
  a.N = d.p[1];
  a.V = d.p[2];
  a.PR = d.p[3];
  a.nE = nE;
  a.nT = n;
  a.periodic = r->in.periodic;
  a.lowpass = r->in.lowpass;
  a.C = r->in.CFDscale;
  a.pscale = 0.5 * r->in.rhoRef * (r->in.Mref * r->in.Mref) * (r->in.a0 * r->in.a0);
  a.dt = r->in.dTau;
  acum_host::gauss_weights13(a.g);
  HIPCHK(hipEventRecord(ctx_event(ctx, 1), st));
  const int grid = (int)std::min<int64_t>(nE, 1 << 20);
  HIPCHK(launch_preprocess(st, a, grid));
  HIPCHK(hipEventRecord(ctx_event(ctx, 2), st));
  for (int i = 0; i < 4; i++)
    HIPCHK(ctx_download(ctx, host[i], d.p[i], (size_t)q * width[i] * sizeof(float)));
  HIPCHK(hipEventRecord(ctx_event(ctx, 3), st));
  HIPCHK(hipStreamSynchronize(st));
  float up = 0, kern = 0, down = 0;
  HIPCHK(hipEventElapsedTime(&up, ctx_event(ctx, 0), ctx_event(ctx, 1)));
  HIPCHK(hipEventElapsedTime(&kern, ctx_event(ctx, 1), ctx_event(ctx, 2)));
  HIPCHK(hipEventElapsedTime(&down, ctx_event(ctx, 2), ctx_event(ctx, 3)));
  acum_timing* t = ctx_timing(ctx);
  *t = acum_timing();
  t->prep_ms = up;
  t->eval_ms = kern;
  t->reduce_ms = down;
  t->total_ms = up + kern + down;
  t->evals = q;
  t->launches = 1;
  r->preprocessed = true;
  return ACUM_OK;
}
