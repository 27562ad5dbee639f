// C-ABI of the broadband (BPM) path.  Replaces the host call findBBnoise (main.cu:150,
// bb_core.cu:11): FP64 source preparation, spectral shapes, observer sums and the final levels,
// all on the GPU; the host only uploads the case arrays.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>
#include "abi_internal.hpp"
#include "bpm_types.hpp"

namespace acum {
struct BpmSurfDev {
  int64_t src0;
  int64_t xoff, doff;
  int64_t soff, moff, roff, aoff, toff;
  int32_t ns, iq, ccw, trip;
  int32_t nst, nm, nr, na;
  double span, omega, psi0;
  double cx, sx, cy, sy, cz, sz;
  double trn[3];
};
struct BpmPrepArgs {
  const BpmSurfDev* surf;
  int32_t nS, nT, t_del, pad;
  int64_t nsec;
  const float *xyzcR, *bbdata, *sect_t, *mach_t, *re_t, *aoa_t, *bbdata_t;
  double dTau, a0, vx, vy, vz;
  BpmShapeIn* in;
  BpmSrcDev* dev;
  uint32_t* err;
};
hipError_t launch_bpm_prep(hipStream_t s, const BpmPrepArgs& a);
hipError_t launch_bpm_shape(hipStream_t s, const BpmShapeIn* in, int64_t nsrc, const float* ff, int nf2,
                            double a0, float* S, BpmSrcDev* dev);
hipError_t launch_bpm_obs(hipStream_t s, const BpmSrcDev* src, const float* S, const int64_t* bounds, int nchunk,
                          int nf2, const float* obs, int nObs, float vx, float vy, float vz, float aMv,
                          float* part);
hipError_t launch_bpm_final(hipStream_t s, const float* part, int nchunk, int nObs, int nf2, int nT, float* spl,
                            float* oaspl);
int bpm_occupancy_blocks(int nf2);
int bpm_bands_per_pass(int nf2);
}  // namespace acum

using namespace acum;

namespace {

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
  for (int s = 0; s < c->nSurf_BB; s++) {
    if (c->nsect[s] < 3 || c->naoa_t[s] < 2 || c->nmach_t[s] < 1 || c->nre_t[s] < 1 || c->nsect_t[s] < 1)
      return ctx_fail(ctx, ACUM_ERR_ARG, "broadband surface needs >= 3 sections and a table");
    if (c->xsPtr[s] < 0 || c->dsPtr[s] < 0 || c->aPtr[s] < 0 || c->mPtr[s] < 0 || c->rPtr[s] < 0 ||
        c->sPtr[s] < 0 || c->dPtr[s] < 0)
      return ctx_fail(ctx, ACUM_ERR_ARG, "negative offset in broadband case");
    if ((int)std::lround(c->nsect[s] * 0.75) >= c->nsect[s])
      return ctx_fail(ctx, ACUM_ERR_ARG, "broadband surface needs at least 3 sections");
  }
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
  const int nS = c->nSurf_BB, nT = c->nTime;
  // per-surface constants and the extents of the case arrays
  const double d2r = M_PI / 180.0;
  std::vector<BpmSurfDev> surf(nS);
  int64_t nsec = 0, nx = 0, nd = 0, nst = 0, nm = 0, nr = 0, na = 0, ntab = 0;
  for (int s = 0; s < nS; s++) {
    BpmSurfDev& f = surf[s];
    f.src0 = nsec;
    f.ns = c->nsect[s];
    f.iq = (int)std::lround(c->nsect[s] * 0.75);
    f.ccw = c->ccw[s];
    f.trip = c->trip[s];
    f.xoff = c->xsPtr[s];
    f.doff = c->dsPtr[s];
    f.soff = c->sPtr[s];
    f.moff = c->mPtr[s];
    f.roff = c->rPtr[s];
    f.aoff = c->aPtr[s];
    f.toff = c->dPtr[s];
    f.nst = c->nsect_t[s];
    f.nm = c->nmach_t[s];
    f.nr = c->nre_t[s];
    f.na = c->naoa_t[s];
    f.span = c->span[s];
    f.omega = c->omega[s];
    f.psi0 = c->psi_offset[s] * d2r;
    const double thx = c->rotate[3 * s] * d2r, thy = c->rotate[3 * s + 1] * d2r, thz = c->rotate[3 * s + 2] * d2r;
    f.cx = std::cos(thx);
    f.sx = std::sin(thx);
    f.cy = std::cos(thy);
    f.sy = std::sin(thy);
    f.cz = std::cos(thz);
    f.sz = std::sin(thz);
    for (int i = 0; i < 3; i++) f.trn[i] = c->translate[3 * s + i];
    nsec += f.ns;
    nx = std::max<int64_t>(nx, f.xoff + 4LL * f.ns);
    nd = std::max<int64_t>(nd, f.doff + (int64_t)nT * f.ns * 7);
    nst = std::max<int64_t>(nst, f.soff + f.nst);
    nm = std::max<int64_t>(nm, f.moff + f.nm);
    nr = std::max<int64_t>(nr, f.roff + f.nr);
    na = std::max<int64_t>(na, f.aoff + f.na);
    ntab = std::max<int64_t>(ntab, f.toff + (int64_t)f.nst * f.nm * f.nr * f.na * 7);
  }
  const int64_t nsrc = nsec * nT;
  const int nf2 = c->nf2, nObs = c->nObs;
  hipStream_t s = ctx_stream(ctx);
  const int obs_tiles = (nObs + 63) / 64;
  const int nb = bpm_bands_per_pass(nf2);
  const int passes = (nf2 + nb - 1) / nb;
  int bpc = std::max(1, bpm_occupancy_blocks(nf2));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, ctx_device(ctx)));
  const int64_t target = 4LL * bpc * prop.multiProcessorCount;
  int64_t nchunk = std::max<int64_t>(1, target / std::max(1, obs_tiles * passes));
  nchunk = std::min<int64_t>(std::min<int64_t>(nchunk, nsrc), 65535);
  // source chunks of decreasing size (largest first), each at least one source
  std::vector<int64_t> bounds(nchunk + 1, 0);
  {
    const double k = size_decay((double)obs_tiles * passes * nchunk / ((double)bpc * prop.multiProcessorCount));
    for (int64_t i = 1; i < nchunk; i++) {
      const int64_t b = (int64_t)std::llround((double)nsrc * size_profile(k, (double)i / nchunk));
      bounds[i] = std::min(std::max(b, bounds[i - 1] + 1), nsrc - (nchunk - i));
    }
    bounds[nchunk] = nsrc;
  }
  BpmSurfDev* d_surf;
  BpmShapeIn* d_in;
  BpmSrcDev* d_src;
  uint32_t* d_err;
  float *d_x, *d_bd, *d_st, *d_mt, *d_rt, *d_at, *d_tt;
  float *d_S, *d_ff, *d_obs, *d_part, *d_spl, *d_oa;
  int64_t* d_bounds;
  // the observer kernel reads a whole pass of bands per source, past nf2 on the last source
  const size_t s_pad = (size_t)bpm_bands_per_pass(nf2) * 6;
  WorkspaceLayout L;
  L.add(&d_surf, nS);
  L.add(&d_err, 1);
  L.add(&d_x, nx);
  L.add(&d_bd, nd);
  L.add(&d_st, nst);
  L.add(&d_mt, nm);
  L.add(&d_rt, nr);
  L.add(&d_at, na);
  L.add(&d_tt, ntab);
  L.add(&d_in, nsrc);
  L.add(&d_src, nsrc);
  L.add(&d_S, (size_t)nsrc * nf2 * 6 + s_pad);
  L.add(&d_ff, nf2);
  L.add(&d_obs, (size_t)nObs * 3);
  L.add(&d_part, (size_t)nchunk * nObs * nf2 * 6);
  L.add(&d_bounds, (size_t)nchunk + 1);
  L.add(&d_spl, (size_t)nObs * nf2 * 7);
  L.add(&d_oa, nObs);
  void* ws = nullptr;
  HIPCHK(ctx_workspace(ctx, L.bytes(), &ws));
  L.assign(ws);
  HIPCHK(hipMemcpyAsync(d_surf, surf.data(), nS * sizeof(BpmSurfDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(d_err, 0, sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(d_S + (size_t)nsrc * nf2 * 6, 0, s_pad * sizeof(float), s));
  HIPCHK(hipMemcpyAsync(d_x, c->xyzcR, nx * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_bd, c->bbdata, nd * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_st, c->sect_t, nst * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_mt, c->mach_t, nm * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_rt, c->re_t, nr * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_at, c->aoa_t, na * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_tt, c->bbdata_t, ntab * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_ff, c->ff, nf2 * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_obs, c->rObs, (size_t)nObs * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_bounds, bounds.data(), (nchunk + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  const double vx = c->vInf[0], vy = c->vInf[1], vz = c->vInf[2];
  const double aMv = c->a0 * c->a0 - vx * vx - vy * vy - vz * vz;
  HIPCHK(hipEventRecord(ctx_event(ctx, 0), s));
  BpmPrepArgs pa{};
  pa.surf = d_surf;
  pa.nS = nS;
  pa.nT = nT;
  pa.t_del = c->t_del;
  pa.nsec = nsec;
  pa.xyzcR = d_x;
  pa.bbdata = d_bd;
  pa.sect_t = d_st;
  pa.mach_t = d_mt;
  pa.re_t = d_rt;
  pa.aoa_t = d_at;
  pa.bbdata_t = d_tt;
  pa.dTau = c->dTau;
  pa.a0 = c->a0;
  pa.vx = c->vInf[0];
  pa.vy = c->vInf[1];
  pa.vz = c->vInf[2];
  pa.in = d_in;
  pa.dev = d_src;
  pa.err = d_err;
  HIPCHK(launch_bpm_prep(s, pa));
  HIPCHK(launch_bpm_shape(s, d_in, nsrc, d_ff, nf2, c->a0, d_S, d_src));
  HIPCHK(hipEventRecord(ctx_event(ctx, 1), s));
  HIPCHK(launch_bpm_obs(s, d_src, d_S, d_bounds, (int)nchunk, nf2, d_obs, nObs, (float)vx, (float)vy, (float)vz,
                        (float)aMv, d_part));
  HIPCHK(hipEventRecord(ctx_event(ctx, 2), s));
  HIPCHK(launch_bpm_final(s, d_part, (int)nchunk, nObs, nf2, c->nTime, d_spl, d_oa));
  HIPCHK(hipEventRecord(ctx_event(ctx, 3), s));
  HIPCHK(hipStreamSynchronize(s));
  uint32_t perr = 0;
  HIPCHK(hipMemcpy(&perr, d_err, sizeof(perr), hipMemcpyDeviceToHost));
  if (perr & 1u)
    return ctx_fail(ctx, ACUM_ERR_ARG, "broadband table lacks a negative AoA, an AoA >= 6 deg or a cl sign change");
  if (perr & 2u)
    return ctx_fail(ctx, ACUM_ERR_ARG, "effective angle of attack below the broadband table's AoA range");
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
  t->launches = 4;
  return ACUM_OK;
}
