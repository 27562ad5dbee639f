// C-ABI of the broadband (BPM) path.  Replaces the host call findBBnoise (main.cu:150,
// bb_core.cu:11): host FP64 source preparation, then the shape, observer and final kernels.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>
#include "abi_internal.hpp"
#include "bpm_types.hpp"
#include "host/bpm_prep.hpp"

namespace acum {
hipError_t launch_bpm_shape(hipStream_t s, const BpmShapeIn* in, int64_t nsrc, const float* ff, int nf2,
                            double a0, float* S, BpmSrcDev* dev);
hipError_t launch_bpm_obs(hipStream_t s, const BpmSrcDev* src, const float* S, int64_t nsrc, int64_t chunk,
                          int nchunk, int nf2, const float* obs, int nObs, float vx, float vy, float vz,
                          float aMv, float* part);
hipError_t launch_bpm_final(hipStream_t s, const float* part, int nchunk, int nObs, int nf2, int nT, float* spl,
                            float* oaspl);
int bpm_occupancy_blocks(int nf2);
int bpm_bands_per_pass(int nf2);
void bpm_release(acum_ctx*) {}
}  // namespace acum

using namespace acum;

namespace {

struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename T>
  hipError_t alloc(T** out, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) p.push_back(q);
    *out = (T*)q;
    return e;
  }
};

int validate(acum_ctx* ctx, const acum_bpm_case* c) {
  if (!c) return ctx_fail(ctx, ACUM_ERR_ARG, "null case");
  if (c->nSurf_BB < 1 || c->nTime < 3 || c->nObs < 1 || c->nf2 < 1)
    return ctx_fail(ctx, ACUM_ERR_ARG, "need nSurf_BB >= 1, nTime >= 3, nObs >= 1, nf2 >= 1");
  if (!c->rObs || !c->ff || !c->span || !c->omega || !c->psi_offset || !c->ccw || !c->trip || !c->rotate ||
      !c->translate || !c->nsect || !c->xsPtr || !c->dsPtr || !c->xyzcR || !c->bbdata || !c->nsect_t ||
      !c->nmach_t || !c->naoa_t || !c->nre_t || !c->aPtr || !c->mPtr || !c->rPtr || !c->sPtr || !c->dPtr ||
      !c->sect_t || !c->mach_t || !c->aoa_t || !c->re_t || !c->bbdata_t)
    return ctx_fail(ctx, ACUM_ERR_ARG, "null array in broadband case");
  if (!(c->a0 > 0) || !(c->dTau > 0)) return ctx_fail(ctx, ACUM_ERR_ARG, "a0 and dTau must be positive");
  const double vv = c->vInf[0] * c->vInf[0] + c->vInf[1] * c->vInf[1] + c->vInf[2] * c->vInf[2];
  if (!(std::sqrt(vv) < c->a0)) return ctx_fail(ctx, ACUM_ERR_NUMERIC, "free-stream Mach >= 1");
  for (int i = 0; i < c->nf2; i++)
    if (!(c->ff[i] > 0)) return ctx_fail(ctx, ACUM_ERR_ARG, "band frequencies must be positive");
  for (int s = 0; s < c->nSurf_BB; s++)
    if (c->nsect[s] < 3 || c->naoa_t[s] < 2 || c->nmach_t[s] < 1 || c->nre_t[s] < 1 || c->nsect_t[s] < 1)
      return ctx_fail(ctx, ACUM_ERR_ARG, "broadband surface needs >= 3 sections and a table");
  return ACUM_OK;
}

}  // namespace

#define HIPCHK(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ctx_fail(ctx, e_ == hipErrorOutOfMemory ? ACUM_ERR_NOMEM : ACUM_ERR_DEVICE, \
                      std::string(#call) + ": " + hipGetErrorString(e_));                \
  } while (0)

extern "C" int acum_bpm(acum_ctx* ctx, const acum_bpm_case* c, float* BBspl, float* BBoaspl) {
  if (!ctx) return ACUM_ERR_ARG;
  if (!BBspl || !BBoaspl) return ctx_fail(ctx, ACUM_ERR_ARG, "null output");
  int rc = validate(ctx, c);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx_device(ctx)));
  std::vector<BpmSource> src;
  std::string err;
  rc = bpm_prepare(c, &src, &err);
  if (rc) return ctx_fail(ctx, rc, err);
  const int64_t nsrc = (int64_t)src.size();
  const int nf2 = c->nf2, nObs = c->nObs;
  std::vector<BpmShapeIn> hin(nsrc);
  std::vector<BpmSrcDev> hdev(nsrc);
  for (int64_t i = 0; i < nsrc; i++) {
    const BpmSource& q = src[i];
    BpmShapeIn& a = hin[i];
    a.u = q.u;
    a.c = q.c;
    a.lsect = q.lsect;
    a.alpha = q.alpha;
    a.ds = q.bl[0];
    a.dp = q.bl[1];
    a.dss = q.bl[2];
    a.dsp = q.bl[3];
    a.R = q.R;
    a.trip = q.trip;
    a.tip = q.tip;
    BpmSrcDev& d = hdev[i];
    for (int k = 0; k < 3; k++) d.xt[k] = (float)q.xyzt[k];
    for (int k = 0; k < 9; k++) d.tr[k] = (float)q.tr[k];
    const double M = std::fabs(q.u) / c->a0;
    d.M = (float)M;
    d.Mmc = (float)(M - 0.6 * M);
    d.tipT0 = 0.0f;
    d.flags = 0;
  }
  hipStream_t s = ctx_stream(ctx);
  const int obs_tiles = (nObs + 63) / 64;
  const int nb = bpm_bands_per_pass(nf2);
  const int passes = (nf2 + nb - 1) / nb;
  int bpc = std::max(1, bpm_occupancy_blocks(nf2));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, ctx_device(ctx)));
  const int64_t target = 4LL * bpc * prop.multiProcessorCount;
  int64_t nchunk = std::max<int64_t>(1, target / std::max(1, obs_tiles * passes));
  nchunk = std::min<int64_t>(nchunk, nsrc);
  const int64_t chunk = (nsrc + nchunk - 1) / nchunk;
  nchunk = (nsrc + chunk - 1) / chunk;
  DevBufs B;
  BpmShapeIn* d_in;
  BpmSrcDev* d_src;
  float *d_S, *d_ff, *d_obs, *d_part, *d_spl, *d_oa;
  HIPCHK(B.alloc(&d_in, nsrc));
  HIPCHK(B.alloc(&d_src, nsrc));
  HIPCHK(B.alloc(&d_S, (size_t)nsrc * nf2 * 6));
  HIPCHK(B.alloc(&d_ff, nf2));
  HIPCHK(B.alloc(&d_obs, (size_t)nObs * 3));
  HIPCHK(B.alloc(&d_part, (size_t)nchunk * nObs * nf2 * 6));
  HIPCHK(B.alloc(&d_spl, (size_t)nObs * nf2 * 7));
  HIPCHK(B.alloc(&d_oa, nObs));
  HIPCHK(hipMemcpyAsync(d_in, hin.data(), nsrc * sizeof(BpmShapeIn), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_src, hdev.data(), nsrc * sizeof(BpmSrcDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_ff, c->ff, nf2 * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_obs, c->rObs, (size_t)nObs * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  const double vx = c->vInf[0], vy = c->vInf[1], vz = c->vInf[2];
  const double aMv = c->a0 * c->a0 - vx * vx - vy * vy - vz * vz;
  HIPCHK(hipEventRecord(ctx_event(ctx, 0), s));
  HIPCHK(launch_bpm_shape(s, d_in, nsrc, d_ff, nf2, c->a0, d_S, d_src));
  HIPCHK(hipEventRecord(ctx_event(ctx, 1), s));
  HIPCHK(launch_bpm_obs(s, d_src, d_S, nsrc, chunk, (int)nchunk, nf2, d_obs, nObs, (float)vx, (float)vy,
                        (float)vz, (float)aMv, d_part));
  HIPCHK(hipEventRecord(ctx_event(ctx, 2), s));
  HIPCHK(launch_bpm_final(s, d_part, (int)nchunk, nObs, nf2, c->nTime, d_spl, d_oa));
  HIPCHK(hipEventRecord(ctx_event(ctx, 3), s));
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(BBspl, d_spl, (size_t)nObs * nf2 * 7 * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(BBoaspl, d_oa, (size_t)nObs * sizeof(float), hipMemcpyDeviceToHost));
  float a = 0, b = 0, d = 0;
  HIPCHK(hipEventElapsedTime(&a, ctx_event(ctx, 0), ctx_event(ctx, 1)));
  HIPCHK(hipEventElapsedTime(&b, ctx_event(ctx, 1), ctx_event(ctx, 2)));
  HIPCHK(hipEventElapsedTime(&d, ctx_event(ctx, 2), ctx_event(ctx, 3)));
  acum_timing* t = ctx_timing(ctx);
  *t = acum_timing();
  t->prep_ms = a;
  t->eval_ms = b;
  t->reduce_ms = d;
  t->total_ms = a + b + d;
  t->evals = nsrc * nObs;
  t->partial_bytes = (int64_t)nchunk * nObs * nf2 * 6 * 4;
  t->n_partials = (int32_t)nchunk;
  t->n_items = (int32_t)(obs_tiles * nchunk * passes);
  t->launches = 2 + passes;
  return ACUM_OK;
}
