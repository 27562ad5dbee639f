// C-ABI of the tonal path: context, staging of cases into HBM, tile planning, the persistent
// eval launch, the ordered reduction, timing, graph replay and the autotuner.
// Replaces the device-side part of the reference driver, main.cu:98-247.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <omp.h>
#include <string>
#include <tuple>
#include <vector>
#include "abi_internal.hpp"
#include "host/run.hpp"
#include "prep_types.hpp"

namespace acum {
hipError_t launch_preprocess(hipStream_t s, const PrepArgs& a, int grid);
hipError_t launch_tonal_eval(bool mixed, bool moving, bool tail, int W, int grid, hipStream_t s,
                             const TonalCaseDev* cases, const Tile* tiles, const Item* items,
                             const int* qoff, int* counter);
hipError_t eval_occupancy(bool mixed, bool moving, int W, int* blocks_per_cu);
bool eval_window_supported(bool mixed, int W);
hipError_t launch_tonal_reduce(bool mixed, hipStream_t s, const TonalCaseDev* cases, int icase,
                               int nT, int nObs);
hipError_t launch_tonal_pos(hipStream_t s, const float* X, int64_t n, const TimeConst& k, void* pos);
hipError_t launch_tonal_attr(bool mixed, hipStream_t s, const float* X, const float* N,
                             const float* V, const float* PR, int64_t n, double a0, double rhoRef,
                             double KB, void* attr, uint32_t* flags);
hipError_t launch_fill_zero(hipStream_t s, void* p, int64_t bytes);
hipError_t launch_reset(hipStream_t s, int* counter, uint32_t* flags);
hipError_t launch_tonal_bins(hipStream_t s, const void* pos, int64_t nE, const TimeConst& k,
                             double ox, double oy, double oz, int32_t* ipts,
                             int32_t* npi, int32_t* nd, int32_t* nu);
hipError_t spectra_scratch_init(int device);
hipError_t launch_tonal_plain(hipStream_t s, const float* X, const float* N, const float* V, const float* PR,
                              const float* obs, int64_t nE, int nT, int nObs, double dTau, double dPsi,
                              double a0, double rhoRef, double oM, const double* vInf, double* acc,
                              uint32_t* flags);
hipError_t launch_plain_to_float(hipStream_t s, const double* acc, float* const* out, int64_t n, uint32_t* flags);
}  // namespace acum

using namespace acum;

// ---------------------------------------------------------------------------------------------
namespace {

struct StagedCase {
  int64_t nE = 0, nT = 0, nObs = 0;
  TimeConst k{};
  double rhoRef = 0, Minf = 0;
  std::vector<int64_t> surf0;  // first element of each surface
  // host positions [nE][nT][3] and per element the largest source displacement between two
  // steps, for planning; shared with the autotuner's views of the case
  std::shared_ptr<std::vector<float>> pos;
  std::shared_ptr<std::vector<double>> step;
  float* d_obs = nullptr;
  float* d_aos[4] = {nullptr, nullptr, nullptr, nullptr};  // X, N, V, PR
  PosRec* d_pos = nullptr;                                  // shared by both precisions
  void* d_attr[2] = {nullptr, nullptr};                     // [ACUM_FP64], [ACUM_MIXED]
  float* d_out[3] = {nullptr, nullptr, nullptr};
  int motion_level = 0;  // 0: the tuned source-motion bound; 1, 2: at least 0.6, 0.8 of the
                         // worst case; 3: the full bound (raised after a window overflow to the
                         // next level that widens the bound)
  bool forced = false;   // plan with `force` (the autotuner's candidates) instead of the
  acum_tune force{};     // context's shape for this case's bucket
};

// Default of acum_tune.motion_permille: on the four-blade rotor a bound of 0.6 holds for every
// tile and observer (0.45 overflows) and lengthens the passes from ~5 to ~6 steps.
constexpr int kMotionPermille = 600;

// The source-motion bound (per mille of the worst case) a case plans with at a motion level.
int motion_permille_at(const acum_tune& t, int level) {
  const int pm = t.motion_permille > 0 ? t.motion_permille : kMotionPermille;
  switch (level) {
    case 0: return pm;
    case 1: return std::max(pm, 600);
    case 2: return std::max(pm, 800);
    default: return 1000;
  }
}
constexpr int kMotionLevels = 4;

// One eval launch: the work items of the cases that use one kernel variant (still air or moving
// medium, window size), drained through kQueues queues; queue q holds items
// [qoff[q], qoff[q+1]) of the launch's item range.
struct Launch {
  bool moving = false;
  bool tail = false;  // the variant with the tail for tent sides longer than two bins
  int window = 0;
  int item0 = 0, nitems = 0;
  int grid = 0;
  int qoff[kQueues + 1] = {};
};

struct Plan {
  bool valid = false;
  std::vector<acum_tune> tunes;  // the tile shape each case was planned with
  std::vector<Tile> tiles;
  std::vector<Item> items;
  std::vector<int> G;
  std::vector<int64_t> slab_off;
  int64_t slab_bytes = 0;
  std::vector<Launch> launches;
  Tile* d_tiles = nullptr;
  Item* d_items = nullptr;
  TonalCaseDev* d_cases = nullptr;
  void* d_slab = nullptr;
  uint32_t* d_flags = nullptr;  // [cases] error flags of the last pass
  int* d_qoff = nullptr;        // [launches][kQueues + 1]
  int* d_counter = nullptr;     // [launches][kQueues] queue heads, cleared before each pass
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// Shape bucket of the autotuner's cache: precision, still air or moving medium, the rounded-up
// base-2 logarithms of the element, step and observer counts, and the device's compute units.
struct TuneKey {
  int prec, moving, lgE, lgT, lgO, cus;
  bool operator<(const TuneKey& o) const {
    return std::tie(prec, moving, lgE, lgT, lgO, cus) < std::tie(o.prec, o.moving, o.lgE, o.lgT, o.lgO, o.cus);
  }
};
struct TuneEntry {
  acum_tune tune;
  double ms;  // time per pass of the chosen shape on the observer subset it was tuned on
};

struct Stage {
  std::vector<StagedCase> cases;
  Plan plan[2];
};

}  // namespace

struct acum_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  hipEvent_t ev[6] = {};
  acum_timing timing{};
  acum_tune tune[2]{};
  int* d_counter = nullptr;
  uint32_t* d_flags = nullptr;
  int cu_count = 0;
  std::map<int, std::unique_ptr<Stage>> stages;
  int next_handle = 1;
  std::map<TuneKey, TuneEntry> tune_cache;  // autotuned shapes by bucket (acum_autotune)
  void* ws = nullptr;  // per-call workspace (ctx_workspace)
  size_t ws_bytes = 0;
  void* pin[2] = {nullptr, nullptr};  // pinned staging buffers (ctx_upload / ctx_download)
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
};

namespace acum {

int ctx_fail(acum_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

hipStream_t ctx_stream(acum_ctx* c) { return c->stream; }
hipEvent_t ctx_event(acum_ctx* c, int i) { return c->ev[i]; }
acum_timing* ctx_timing(acum_ctx* c) { return &c->timing; }
int ctx_device(acum_ctx* c) { return c->device; }

namespace {

constexpr size_t kStageChunk = (size_t)64 << 20;  // bytes per pinned staging buffer
constexpr size_t kStageMin = (size_t)8 << 20;     // smaller copies go straight from pageable memory

// host memcpy split over up to 16 threads
void par_memcpy(void* dst, const void* src, size_t n) {
  const size_t piece = (size_t)1 << 20;
  const int64_t np = (int64_t)((n + piece - 1) / piece);
  const int nt = (int)std::min<int64_t>(np, std::min(16, omp_get_max_threads()));
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < np; i++) {
    const size_t o = (size_t)i * piece;
    std::memcpy((char*)dst + o, (const char*)src + o, std::min(piece, n - o));
  }
}

hipError_t ensure_pinned(acum_ctx* c) {
  for (int b = 0; b < 2; b++) {
    if (!c->pin[b]) {
      const hipError_t e = hipHostMalloc(&c->pin[b], kStageChunk, hipHostMallocDefault);
      if (e != hipSuccess) {
        c->pin[b] = nullptr;
        return e;
      }
    }
    if (!c->pin_ev[b]) {
      hipError_t e = hipEventCreateWithFlags(&c->pin_ev[b], hipEventDisableTiming);
      if (e != hipSuccess) {
        c->pin_ev[b] = nullptr;
        return e;
      }
      if ((e = hipEventRecord(c->pin_ev[b], c->stream)) != hipSuccess) return e;  // a completed event
    }
  }
  return hipSuccess;
}

}  // namespace

// Large host-to-device copies from pageable memory go through the context's two pinned
// buffers: a parallel host copy into one buffer overlaps the DMA out of the other.  The source
// may be reused when the call returns; the DMAs complete in stream order.
hipError_t ctx_upload(acum_ctx* c, void* dst, const void* src, size_t n) {
  if (n < kStageMin) return hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream);
  hipError_t e = ensure_pinned(c);
  if (e != hipSuccess) return e;
  for (size_t off = 0, i = 0; off < n; off += kStageChunk, i++) {
    const int b = (int)(i & 1);
    const size_t len = std::min(kStageChunk, n - off);
    if ((e = hipEventSynchronize(c->pin_ev[b])) != hipSuccess) return e;  // its last DMA is done
    par_memcpy(c->pin[b], (const char*)src + off, len);
    if ((e = hipMemcpyAsync((char*)dst + off, c->pin[b], len, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
      return e;
    if ((e = hipEventRecord(c->pin_ev[b], c->stream)) != hipSuccess) return e;
  }
  return hipSuccess;
}

// Large device-to-host copies into pageable memory, the reverse way; returns when dst holds the
// data (the copy follows the work queued before it on the context's stream).
hipError_t ctx_download(acum_ctx* c, void* dst, const void* src, size_t n) {
  hipError_t e;
  if (n < kStageMin) {
    if ((e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, c->stream)) != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
  }
  if ((e = ensure_pinned(c)) != hipSuccess) return e;
  size_t prev_off = 0, prev_len = 0;
  for (size_t off = 0, i = 0; off < n + kStageChunk; off += kStageChunk, i++) {
    const int b = (int)(i & 1);
    if (off < n) {  // queue this chunk's DMA into buffer b (emptied by the host before)
      const size_t len = std::min(kStageChunk, n - off);
      if ((e = hipMemcpyAsync(c->pin[b], (const char*)src + off, len, hipMemcpyDeviceToHost, c->stream)) !=
          hipSuccess)
        return e;
      if ((e = hipEventRecord(c->pin_ev[b], c->stream)) != hipSuccess) return e;
    }
    if (prev_len) {  // empty the other buffer, filled by the previous chunk's DMA
      if ((e = hipEventSynchronize(c->pin_ev[1 - b])) != hipSuccess) return e;
      par_memcpy((char*)dst + prev_off, c->pin[1 - b], prev_len);
    }
    prev_off = off;
    prev_len = off < n ? std::min(kStageChunk, n - off) : 0;
  }
  return hipSuccess;
}

hipError_t ctx_workspace(acum_ctx* c, size_t bytes, void** out) {
  if (bytes > c->ws_bytes) {
    void* old = c->ws;
    c->ws = nullptr;
    c->ws_bytes = 0;
    if (old) {
      const hipError_t e = hipFree(old);
      if (e != hipSuccess) return e;
    }
    const hipError_t e = hipMalloc(&c->ws, bytes);
    if (e != hipSuccess) {
      c->ws = nullptr;
      return e;
    }
    c->ws_bytes = bytes;
  }
  *out = c->ws;
  return hipSuccess;
}

}  // namespace acum

#define HIPCHK(ctx, call)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ctx_fail(ctx, e_ == hipErrorOutOfMemory ? ACUM_ERR_NOMEM : ACUM_ERR_DEVICE,   \
                      std::string(#call) + ": " + hipGetErrorString(e_));                  \
  } while (0)

namespace {

void free_plan(Plan& p) {
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  if (p.graph) (void)hipGraphDestroy(p.graph);
  (void)hipFree(p.d_tiles);
  (void)hipFree(p.d_items);
  (void)hipFree(p.d_qoff);
  (void)hipFree(p.d_counter);
  (void)hipFree(p.d_cases);
  (void)hipFree(p.d_slab);
  (void)hipFree(p.d_flags);
  p = Plan();
}

void free_stage(Stage& s) {
  for (auto& c : s.cases) {
    (void)hipFree(c.d_obs);
    for (auto& p : c.d_aos) (void)hipFree(p);
    (void)hipFree(c.d_pos);
    for (auto& p : c.d_attr) (void)hipFree(p);
    for (auto& p : c.d_out) (void)hipFree(p);
  }
  for (auto& p : s.plan) free_plan(p);
}

TimeConst make_time_const(const acum_tonal_case* c) {
  // The CPU oracle builds the same constants with the same operations.
  TimeConst k{};
  k.a0 = c->a0;
  k.r2b = (c->oM / c->a0) / c->dPsi;
  k.KB = c->oM / c->dPsi;
  k.SB = c->dTau * k.KB;
  k.nT = (int32_t)c->nTime;
  const double vx = c->vInf[0], vy = c->vInf[1], vz = c->vInf[2];
  double vv = vx * vx;
  vv = vv + vy * vy;
  vv = vv + vz * vz;
  const double na = k.a0 * k.a0 - vv;
  const double kn = k.KB / na;
  k.wx = vx * kn;
  k.wy = vy * kn;
  k.wz = vz * kn;
  k.C2 = k.KB * kn;
  k.aB = k.a0 / k.KB;
  k.ux = vx / k.KB;
  k.uy = vy / k.KB;
  k.uz = vz / k.KB;
  k.h2 = 1.0 / (2.0 * k.C2);
  k.still = (vx == 0.0 && vy == 0.0 && vz == 0.0) ? 1 : 0;
  return k;
}

// A source on an observer: an observer within kCoincident metres of a source position at any
// step (|obs - src|^2 formed in FP64 from the parsed floats, as the oracle does) fails the case
// with ACUM_ERR_NUMERIC.  The canonical observer time expands |obs - src|^2 about the sources'
// centre, whose rounding does not give an exact zero there.  Bounding boxes (all sources, each
// surface, each element) leave only observers near a source to the position-by-position test.
constexpr double kCoincident = 1e-6;

int check_coincident(acum_ctx* ctx, const StagedCase& sc, const float* rObs) {
  const std::vector<float>& P = *sc.pos;
  const int64_t nT = sc.nT;
  const double r2 = kCoincident * kCoincident;
  struct Box {
    double lo[3], hi[3];
  };
  auto near = [&](const Box& b, const double* o) {
    double d2 = 0;
    for (int d = 0; d < 3; d++) {
      const double g = std::max({b.lo[d] - o[d], 0.0, o[d] - b.hi[d]});
      d2 += g * g;
    }
    return d2 <= 4.0 * r2;
  };
  std::vector<Box> eb(sc.nE);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < sc.nE; e++) {
    Box b{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
    for (int64_t k = 0; k < nT; k++)
      for (int d = 0; d < 3; d++) {
        const double v = P[(e * nT + k) * 3 + d];
        b.lo[d] = std::min(b.lo[d], v);
        b.hi[d] = std::max(b.hi[d], v);
      }
    eb[e] = b;
  }
  const size_t ns = sc.surf0.size();
  std::vector<Box> sb(ns);
  Box all{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
  for (size_t s = 0; s < ns; s++) {
    const int64_t a = sc.surf0[s], b = (s + 1 < ns) ? sc.surf0[s + 1] : sc.nE;
    Box x{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
    for (int64_t e = a; e < b; e++)
      for (int d = 0; d < 3; d++) {
        x.lo[d] = std::min(x.lo[d], eb[e].lo[d]);
        x.hi[d] = std::max(x.hi[d], eb[e].hi[d]);
      }
    sb[s] = x;
    for (int d = 0; d < 3; d++) {
      all.lo[d] = std::min(all.lo[d], x.lo[d]);
      all.hi[d] = std::max(all.hi[d], x.hi[d]);
    }
  }
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : bad)
  for (int64_t i = 0; i < sc.nObs; i++) {
    const double o[3] = {rObs[3 * i], rObs[3 * i + 1], rObs[3 * i + 2]};
    if (!near(all, o)) continue;
    for (size_t s = 0; s < ns && !bad; s++) {
      if (!near(sb[s], o)) continue;
      const int64_t a = sc.surf0[s], b = (s + 1 < ns) ? sc.surf0[s + 1] : sc.nE;
      for (int64_t e = a; e < b && !bad; e++) {
        if (!near(eb[e], o)) continue;
        for (int64_t k = 0; k < nT; k++) {
          const float* q = &P[(e * nT + k) * 3];
          const double dx = o[0] - (double)q[0], dy = o[1] - (double)q[1], dz = o[2] - (double)q[2];
          if (std::fma(dz, dz, std::fma(dy, dy, dx * dx)) < r2) {
            bad = 1;
            break;
          }
        }
      }
    }
  }
  if (bad) return ctx_fail(ctx, ACUM_ERR_NUMERIC, "an observer lies on a source (within 1e-6 m of a source position)");
  return ACUM_OK;
}

// Origin of the observer-time expansion: the centre of the bounding box of every source position
// (host copy [nE][nT][3] of the parsed floats); the oracle takes the same from the case arrays.
void set_origin(StagedCase& sc) {
  const std::vector<float>& P = *sc.pos;
  const int64_t n = (int64_t)P.size() / 3;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma omp parallel
  {
    double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma omp for schedule(static) nowait
    for (int64_t i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) {
        const double v = P[3 * i + d];
        l[d] = std::min(l[d], v);
        h[d] = std::max(h[d], v);
      }
#pragma omp critical
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], l[d]);
      hi[d] = std::max(hi[d], h[d]);
    }
  }
  sc.k.Ox = (lo[0] + hi[0]) * 0.5;
  sc.k.Oy = (lo[1] + hi[1]) * 0.5;
  sc.k.Oz = (lo[2] + hi[2]) * 0.5;
}

int validate(acum_ctx* ctx, const acum_tonal_case* c) {
  if (!c) return ctx_fail(ctx, ACUM_ERR_ARG, "null case");
  if (c->nSurf < 1 || c->nTime < 4 || c->nObs < 1)
    return ctx_fail(ctx, ACUM_ERR_ARG, "need nSurf >= 1, nTime >= 4, nObs >= 1");
  if (c->nTime > (1 << 20) || c->nObs > (1 << 26))
    return ctx_fail(ctx, ACUM_ERR_ARG, "nTime or nObs too large");
  if (!c->E || !c->X || !c->N || !c->V || !c->PR || !c->XPtr || !c->NPtr || !c->VPtr ||
      !c->PRPtr || !c->rObs)
    return ctx_fail(ctx, ACUM_ERR_ARG, "null array in case");
  for (int64_t s = 0; s < c->nSurf; s++)
    if (c->E[s] < 1) return ctx_fail(ctx, ACUM_ERR_ARG, "surface with no elements");
  for (int64_t i = 0; i < 3 * c->nObs; i++)
    if (!std::isfinite(c->rObs[i])) return ctx_fail(ctx, ACUM_ERR_ARG, "non-finite observer position");
  if (!(c->a0 > 0) || !(c->dPsi > 0) || !(c->dTau > 0) || !(c->oM > 0))
    return ctx_fail(ctx, ACUM_ERR_ARG, "a0, dPsi, dTau and oM must be positive");
  const double vv = c->vInf[0] * c->vInf[0] + c->vInf[1] * c->vInf[1] + c->vInf[2] * c->vInf[2];
  if (!(std::sqrt(vv) < c->a0))
    return ctx_fail(ctx, ACUM_ERR_NUMERIC, "free-stream Mach >= 1: retarded time undefined");
  if (c->precision != ACUM_FP64 && c->precision != ACUM_MIXED)
    return ctx_fail(ctx, ACUM_ERR_ARG, "precision must be ACUM_FP64 or ACUM_MIXED");
  return ACUM_OK;
}

int stage_case(acum_ctx* ctx, const acum_tonal_case* c, StagedCase& sc) {
  int rc = validate(ctx, c);
  if (rc) return rc;
  const int64_t nT = c->nTime;
  sc.nT = nT;
  sc.nObs = c->nObs;
  sc.k = make_time_const(c);
  sc.rhoRef = c->rhoRef;
  sc.Minf = std::sqrt(c->vInf[0] * c->vInf[0] + c->vInf[1] * c->vInf[1] + c->vInf[2] * c->vInf[2]) / c->a0;
  sc.surf0.resize(c->nSurf);
  int64_t nE = 0;
  for (int64_t s = 0; s < c->nSurf; s++) {
    sc.surf0[s] = nE;
    nE += c->E[s];
  }
  sc.nE = nE;
  const int64_t q = nE * nT;
  const int width[4] = {4, 6, 6, 2};
  const float* src[4] = {c->X, c->N, c->V, c->PR};
  const int64_t* off[4] = {c->XPtr, c->NPtr, c->VPtr, c->PRPtr};
  for (int a = 0; a < 4; a++) {
    HIPCHK(ctx, hipMalloc(&sc.d_aos[a], (size_t)q * width[a] * sizeof(float)));
    int64_t e0 = 0;
    for (int64_t s = 0; s < c->nSurf; s++) {
      const size_t n = (size_t)c->E[s] * nT * width[a];
      HIPCHK(ctx, ctx_upload(ctx, sc.d_aos[a] + e0 * nT * width[a], src[a] + off[a][s], n * sizeof(float)));
      e0 += c->E[s];
    }
  }
  sc.pos = std::make_shared<std::vector<float>>((size_t)q * 3);
  {
    int64_t e0 = 0;
    for (int64_t s = 0; s < c->nSurf; s++) {
      const float* X = c->X + c->XPtr[s];
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < c->E[s] * nT; i++)
        for (int d = 0; d < 3; d++) (*sc.pos)[(e0 * nT + i) * 3 + d] = X[i * 4 + d];
      e0 += c->E[s];
    }
  }
  set_origin(sc);
  rc = check_coincident(ctx, sc, c->rObs);
  if (rc) return rc;
  HIPCHK(ctx, hipMalloc(&sc.d_obs, (size_t)c->nObs * 3 * sizeof(float)));
  HIPCHK(ctx, hipMemcpyAsync(sc.d_obs, c->rObs, (size_t)c->nObs * 3 * sizeof(float),
                             hipMemcpyHostToDevice, ctx->stream));
  for (int i = 0; i < 3; i++) HIPCHK(ctx, hipMalloc(&sc.d_out[i], (size_t)nT * c->nObs * sizeof(float)));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return ACUM_OK;
}

// A run read from its files but not preprocessed: the raw element arrays are uploaded once,
// preprocess_imper (preprocess.cu:228-413) runs on the device in place (the same kernel as
// acum_preprocess_gpu, bit-identical to the host FP64 form), and the staged case keeps the
// device arrays, so the preprocessed element data never return to the host.
int stage_raw_run(acum_ctx* ctx, const acum_run* r, StagedCase& sc) {
  acum_tonal_case c{};
  c.nSurf = r->nSurf;
  c.nTime = r->nTime;
  c.nObs = r->nObs;
  c.E = r->E.data();
  c.X = r->X.data();
  c.N = r->N.data();
  c.V = r->V.data();
  c.PR = r->PR.data();
  c.XPtr = r->XPtr.data();
  c.NPtr = r->NPtr.data();
  c.VPtr = r->VPtr.data();
  c.PRPtr = r->PRPtr.data();
  c.rObs = r->rObs.data();
  c.dPsi = r->in.dPsi;
  c.dTau = r->in.dTau;
  c.a0 = r->in.a0;
  c.rhoRef = r->in.rhoRef;
  c.oM = r->in.oM;
  for (int i = 0; i < 3; i++) c.vInf[i] = r->in.vInf[i];
  c.precision = ACUM_MIXED;
  int rc = validate(ctx, &c);
  if (rc) return rc;
  const int64_t nT = c.nTime;
  if (nT > kPrepMaxSteps)
    return ctx_fail(ctx, ACUM_ERR_ARG, "GPU preprocessing holds at most 4096 steps per series");
  int64_t nE = 0;
  sc.surf0.resize(c.nSurf);
  for (int64_t s = 0; s < c.nSurf; s++) {
    if (c.XPtr[s] != nE * nT * 4 || c.NPtr[s] != nE * nT * 6 || c.VPtr[s] != nE * nT * 6 ||
        c.PRPtr[s] != nE * nT * 2)
      return ctx_fail(ctx, ACUM_ERR_ARG, "surface arrays are not contiguous");
    sc.surf0[s] = nE;
    nE += c.E[s];
  }
  sc.nT = nT;
  sc.nObs = c.nObs;
  sc.k = make_time_const(&c);
  sc.rhoRef = c.rhoRef;
  sc.Minf = std::sqrt(c.vInf[0] * c.vInf[0] + c.vInf[1] * c.vInf[1] + c.vInf[2] * c.vInf[2]) / c.a0;
  sc.nE = nE;
  const int64_t q = nE * nT;
  const int width[4] = {4, 6, 6, 2};
  const float* src[4] = {c.X, c.N, c.V, c.PR};
  for (int a = 0; a < 4; a++) {
    HIPCHK(ctx, hipMalloc(&sc.d_aos[a], (size_t)q * width[a] * sizeof(float)));
    if (a != 2)  // V is written by the preprocessing
      HIPCHK(ctx, ctx_upload(ctx, sc.d_aos[a], src[a], (size_t)q * width[a] * sizeof(float)));
  }
  PrepArgs pa{};
  pa.X = sc.d_aos[0];
  pa.N = sc.d_aos[1];
  pa.V = sc.d_aos[2];
  pa.PR = sc.d_aos[3];
  pa.nE = nE;
  pa.nT = (int)nT;
  pa.periodic = r->in.periodic;
  pa.lowpass = r->in.lowpass;
  pa.C = r->in.CFDscale;
  pa.pscale = 0.5 * r->in.rhoRef * (r->in.Mref * r->in.Mref) * (r->in.a0 * r->in.a0);
  pa.dt = r->in.dTau;
  acum_host::gauss_weights13(pa.g);
  HIPCHK(ctx, launch_preprocess(ctx->stream, pa, (int)std::min<int64_t>(nE, 1 << 20)));
  // host positions for the planner: the preprocessing scales them by CFDscale in FP64 and rounds
  // to float, which the host repeats exactly
  sc.pos = std::make_shared<std::vector<float>>((size_t)q * 3);
  const double C = r->in.CFDscale;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < q; i++)
    for (int d = 0; d < 3; d++) (*sc.pos)[i * 3 + d] = (float)((double)c.X[i * 4 + d] * C);
  set_origin(sc);
  rc = check_coincident(ctx, sc, c.rObs);
  if (rc) return rc;
  HIPCHK(ctx, hipMalloc(&sc.d_obs, (size_t)c.nObs * 3 * sizeof(float)));
  HIPCHK(ctx, hipMemcpyAsync(sc.d_obs, c.rObs, (size_t)c.nObs * 3 * sizeof(float), hipMemcpyHostToDevice,
                             ctx->stream));
  for (int i = 0; i < 3; i++) HIPCHK(ctx, hipMalloc(&sc.d_out[i], (size_t)nT * c.nObs * sizeof(float)));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return ACUM_OK;
}

int ensure_records(acum_ctx* ctx, StagedCase& sc, int prec) {
  if (sc.d_attr[prec]) return ACUM_OK;
  if (!sc.d_aos[0]) return ctx_fail(ctx, ACUM_ERR_ARG, "staged element data already released");
  const bool mixed = prec == ACUM_MIXED;
  const int64_t n = sc.nE * sc.nT;
  HIPCHK(ctx, launch_reset(ctx->stream, ctx->d_counter, ctx->d_flags));
  if (!sc.d_pos) {
    HIPCHK(ctx, hipMalloc(&sc.d_pos, (size_t)n * sizeof(PosRec)));
    HIPCHK(ctx, launch_tonal_pos(ctx->stream, sc.d_aos[0], n, sc.k, sc.d_pos));
  }
  const size_t bytes = (size_t)n * (mixed ? sizeof(AttrT<float>) : sizeof(AttrT<double>));
  HIPCHK(ctx, hipMalloc(&sc.d_attr[prec], bytes));
  HIPCHK(ctx, launch_tonal_attr(mixed, ctx->stream, sc.d_aos[0], sc.d_aos[1], sc.d_aos[2], sc.d_aos[3],
                                n, sc.k.a0, sc.rhoRef, sc.k.KB, sc.d_attr[prec], ctx->d_flags));
  {
    uint32_t f = 0;
    HIPCHK(ctx, hipMemcpyAsync(&f, ctx->d_flags, sizeof(f), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (f & kFlagNumeric) {
      (void)hipFree(sc.d_attr[prec]);
      sc.d_attr[prec] = nullptr;
      return ctx_fail(ctx, ACUM_ERR_NUMERIC, "non-finite value in the element data");
    }
  }
  if (sc.d_attr[0] && sc.d_attr[1]) {  // both record sets built: the AoS copy is no longer needed
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& p : sc.d_aos) {
      (void)hipFree(p);
      p = nullptr;
    }
  }
  return ACUM_OK;
}

constexpr int64_t kPlanChunk = 4096;  // elements planned together (tiles never cross a chunk)

// Largest source displacement between consecutive steps, per element (planning input, computed
// once per staged case).
void element_steps(StagedCase& sc) {
  if (sc.step) return;
  const int64_t nT = sc.nT;
  const float* P = sc.pos->data();
  auto st = std::make_shared<std::vector<double>>(sc.nE);
  std::vector<double>& step = *st;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < sc.nE; e++) {
    double m = 0;
    for (int64_t k = 0; k < nT; k++) {
      const float* a = P + (e * nT + k) * 3;
      const float* b = P + (e * nT + (k + 1) % nT) * 3;
      const double dx = (double)a[0] - b[0], dy = (double)a[1] - b[1], dz = (double)a[2] - b[2];
      m = std::max(m, dx * dx + dy * dy + dz * dz);
    }
    step[e] = std::sqrt(m);
  }
  sc.step = st;
}

// Tiles of one case: runs of consecutive elements of one surface (within chunks of kPlanChunk
// elements, planned in parallel) whose observer times stay within
// `spread` - 1 bins of the first element's at every step (Lipschitz bound of the convected
// retarded time: |dtau| <= |dx| L, L = oM / (a0 dPsi (1 - Minf)) bins per metre), with the pass
// length K chosen so every bin a pass can touch fits the window.  With the window based at
// floor(tau_e0(k0 - 1)) - spread - 1 and at most step_bins bins between consecutive steps:
//   lowest bin  >= floor(tau_e(k0)) - 1 >= base                       (footprint ic - 1)
//   highest bin <= tau_e0(k0 - 1) + max(K step_bins + 2, (K + 1) step_bins) + spread
// so K step_bins + 2 spread + 5 <= W and (K + 1) step_bins + 2 spread + 3 <= W.
int plan_tiles(acum_ctx* ctx, StagedCase& sc, const acum_tune& t, std::vector<Tile>& tiles) {
  const int64_t nT = sc.nT;
  // bins of observer time per metre of source motion: the full bound, or the reduced one
  const double motion = motion_permille_at(t, sc.motion_level) / 1000.0;
  const double L = sc.k.r2b / (1.0 - sc.Minf) * motion;
  const int W = t.window;
  const float* P = sc.pos->data();
  auto dist2 = [&](int64_t e1, int64_t k1, int64_t e2, int64_t k2) {
    const float* a = P + (e1 * nT + k1) * 3;
    const float* b = P + (e2 * nT + k2) * 3;
    const double dx = (double)a[0] - b[0], dy = (double)a[1] - b[1], dz = (double)a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
  };
  element_steps(sc);
  const std::vector<double>& step = *sc.step;
  auto pass_len = [&](int spread, double S) {
    const double sb = sc.k.SB + S * L;
    const int K = (int)std::min(std::floor((W - 2.0 * spread - 5.0) / sb),
                                std::floor((W - 2.0 * spread - 3.0) / sb) - 1);
    return std::min<int64_t>(std::min<int64_t>(K, t.max_chunk), nT);
  };
  // chunks of at most kPlanChunk elements of one surface, planned in parallel; the chunking
  // depends only on the case, so the tiles do not depend on the host's thread count
  struct Range {
    int64_t a, b;
  };
  std::vector<Range> chunks;
  const size_t nsurf = sc.surf0.size();
  for (size_t s = 0; s < nsurf; s++) {
    const int64_t a = sc.surf0[s], b = (s + 1 < nsurf) ? sc.surf0[s + 1] : sc.nE;
    for (int64_t c = a; c < b; c += kPlanChunk) chunks.push_back(Range{c, std::min(b, c + kPlanChunk)});
  }
  std::vector<std::vector<Tile>> per(chunks.size());
  int status = ACUM_OK;
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t ci = 0; ci < chunks.size(); ci++) {
    const int64_t b = chunks[ci].b;
    int64_t e0 = chunks[ci].a;
    while (e0 < b) {
      double R = 0, S = step[e0];
      int spread = 1;
      int K = (int)pass_len(spread, S);
      if (K < 1) {
#pragma omp critical
        status = ACUM_ERR_PLAN;
        break;
      }
      int64_t e = e0 + 1;
      for (; e < b && e - e0 < t.max_tile; e++) {
        double d = 0;
        for (int64_t k = 0; k < nT; k++) d = std::max(d, dist2(e, k, e0, k));
        const double Rn = std::max(R, std::sqrt(d)), Sn = std::max(S, step[e]);
        const int sp = (int)std::ceil(Rn * L) + 1;
        const int Kn = (int)pass_len(sp, Sn);
        if (sp > t.max_spread || Kn < 4) break;
        R = Rn;
        S = Sn;
        spread = sp;
        K = Kn;
      }
      Tile tl{};
      tl.e0 = e0;
      tl.ne = (int32_t)(e - e0);
      tl.spread = spread;
      tl.K = K;
      per[ci].push_back(tl);
      e0 = e;
    }
  }
  if (status) return ctx_fail(ctx, ACUM_ERR_PLAN, "observer-time window too small for the source speed");
  tiles.clear();
  for (auto& v : per) tiles.insert(tiles.end(), v.begin(), v.end());
  return ACUM_OK;
}

// The tile shape a case is planned with: the autotuner's candidate while it times one, else the
// cached shape of the case's bucket, else the context's shape for the precision.
TuneKey tune_key(const acum_ctx* ctx, const StagedCase& sc, int prec) {
  auto lg = [](int64_t n) {
    int l = 0;
    while (((int64_t)1 << l) < n) l++;
    return l;
  };
  return TuneKey{prec, sc.k.still ? 0 : 1, lg(sc.nE), lg(sc.nT), lg(sc.nObs), ctx->cu_count};
}

acum_tune case_tune(const acum_ctx* ctx, const StagedCase& sc, int prec) {
  if (sc.forced) return sc.force;
  auto it = ctx->tune_cache.find(tune_key(ctx, sc, prec));
  return it != ctx->tune_cache.end() ? it->second.tune : ctx->tune[prec];
}

int build_plan(acum_ctx* ctx, Stage& st, int prec) {
  Plan& p = st.plan[prec];
  const int nc = (int)st.cases.size();
  std::vector<acum_tune> tunes(nc);
  for (int c = 0; c < nc; c++) tunes[c] = case_tune(ctx, st.cases[c], prec);
  if (p.valid && p.tunes.size() == tunes.size() &&
      std::memcmp(p.tunes.data(), tunes.data(), nc * sizeof(acum_tune)) == 0)
    return ACUM_OK;
  free_plan(p);
  p.tunes = tunes;
  const bool mixed = prec == ACUM_MIXED;
  // one launch per kernel variant (still air or moving medium, window) in use; a case's
  // capacity is that of its own variant, so it gets the same groups alone or in a batch
  // A case whose observer time can advance by more than two bins in one step needs the kernel
  // variant with the tail for long tent sides; the bound is rigorous: |d tau / d x| <=
  // r2b / (1 - Minf) bins per metre of source motion (the convected delay's gradient), so a step
  // advances by at most SB + (largest source displacement per step) r2b / (1 - Minf).
  std::vector<std::tuple<int, int, int>> kinds;  // (moving, window, tail)
  std::vector<int> kind_of(nc), capacity(nc);
  for (int c = 0; c < nc; c++) {
    StagedCase& sc = st.cases[c];
    const int W = tunes[c].window;
    if (!eval_window_supported(mixed, W)) return ctx_fail(ctx, ACUM_ERR_ARG, "no eval kernel for this window size");
    int bpc = 0;
    HIPCHK(ctx, eval_occupancy(mixed, !sc.k.still, W, &bpc));
    if (bpc < 1) return ctx_fail(ctx, ACUM_ERR_DEVICE, "eval kernel does not fit a compute unit");
    capacity[c] = bpc * ctx->cu_count;
    element_steps(sc);
    const double smax = sc.step->empty() ? 0.0 : *std::max_element(sc.step->begin(), sc.step->end());
    const bool tail = !(sc.k.SB + smax * sc.k.r2b / (1.0 - sc.Minf) <= 1.99);
    kinds.emplace_back(sc.k.still ? 0 : 1, W, tail ? 1 : 0);
  }
  {
    std::vector<std::tuple<int, int, int>> u = kinds;
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    for (int c = 0; c < nc; c++) kind_of[c] = (int)(std::lower_bound(u.begin(), u.end(), kinds[c]) - u.begin());
    p.launches.assign(u.size(), Launch());
    for (size_t l = 0; l < u.size(); l++) {
      p.launches[l].moving = std::get<0>(u[l]) != 0;
      p.launches[l].window = std::get<1>(u[l]);
      p.launches[l].tail = std::get<2>(u[l]) != 0;
    }
  }
  std::vector<std::vector<Tile>> ct(nc);
  for (int c = 0; c < nc; c++) {
    int rc = plan_tiles(ctx, st.cases[c], tunes[c], ct[c]);
    if (rc) return rc;
  }
  const size_t asz = mixed ? sizeof(float) : sizeof(double);
  p.G.assign(nc, 1);
  p.slab_off.assign(nc, 0);
  p.slab_bytes = 0;
  for (int c = 0; c < nc; c++) {
    const StagedCase& sc = st.cases[c];
    const int nog = (int)((sc.nObs + 63) / 64);
    // G depends on this case alone, so a case gets the same tiles, groups and summation order
    // whether it runs alone or inside a batch (bit-identical results).
    const double target_items = 4.0 * capacity[c];
    int G = tunes[c].groups > 0 ? tunes[c].groups : (int)std::lround(target_items / nog);
    G = std::max(1, std::min<int>(G, (int)ct[c].size()));
    const int64_t row = sc.nObs * sc.nT * 2 * (int64_t)asz;
    // partial rows take at most a quarter of the case's element records, or 8 GiB if that is
    // more: the bound scales with the case, so several cases of a batch still fit together
    const int64_t rec_bytes = sc.nE * sc.nT * (int64_t)(sizeof(PosRec) + (mixed ? sizeof(AttrT<float>) : sizeof(AttrT<double>)));
    const int64_t cap = std::max<int64_t>((int64_t)8 << 30, rec_bytes / 4);
    while (G > 1 && G * row > cap) G--;
    p.G[c] = G;
    p.slab_off[c] = p.slab_bytes;
    p.slab_bytes += ((G * row + 255) / 256) * 256;
    // split the tiles into G runs of decreasing element count
    int64_t tot = 0;
    for (auto& tl : ct[c]) tot += tl.ne;
    const int base = (int)p.tiles.size();
    p.tiles.insert(p.tiles.end(), ct[c].begin(), ct[c].end());
    const int nt = (int)ct[c].size();
    int t0 = 0;
    int64_t acc = 0;
    // Group sizes fall off exponentially with the group index (density e^(-kx) over x in
    // [0, 1]; at four items per slot k = 3 and the first group is about 20 times the last).
    // The queues run large groups first, so the launch ends on small items and the workgroups
    // finish close together.  k depends on this case and the device alone.
    const double k = size_decay((double)G * nog / capacity[c]);
    auto cum = [k](double x) { return size_profile(k, x); };
    for (int g = 0; g < G; g++) {
      const double x = (double)(g + 1) / G;
      const int64_t goal = (g == G - 1) ? tot : (int64_t)std::llround((double)tot * cum(x));
      int t1 = t0;
      int64_t grp = 0;
      // at least one tile per group; stop before overshooting the goal; leave one tile for
      // each later group; the last group takes the rest
      while (t1 < nt && (t1 == t0 || g == G - 1 ||
                         (acc + ct[c][t1].ne <= goal && nt - (t1 + 1) >= G - 1 - g))) {
        acc += ct[c][t1].ne;
        grp += ct[c][t1].ne;
        t1++;
      }
      for (int og = 0; og < nog; og++) {
        Item it{};
        it.icase = c;
        it.obs0 = og * 64;
        it.tile0 = base + t0;
        it.ntile = t1 - t0;
        it.group = g;
        it.cost = grp * sc.nT * 64;
        if (it.ntile > 0) p.items.push_back(it);
      }
      t0 = t1;
    }
  }
  // Items grouped by launch.  Within a launch the tile groups (all observer groups of one group
  // of tiles) are spread over the kQueues queues, longest first onto the least loaded queue; a
  // queue lists its groups longest first with each group's items adjacent, so the waves of one
  // XCD work on the same records at the same time.
  {
    struct Unit {
      int icase, group, queue;
      int64_t cost;
    };
    std::map<std::pair<int, int>, int> uid;
    std::vector<Unit> units;
    for (const Item& it : p.items) {
      auto key = std::make_pair(it.icase, it.group);
      auto f = uid.find(key);
      if (f == uid.end()) {
        uid[key] = (int)units.size();
        units.push_back(Unit{it.icase, it.group, 0, 0});
        f = uid.find(key);
      }
      units[f->second].cost += it.cost;
    }
    const int nl = (int)p.launches.size();
    for (int l = 0; l < nl; l++) {
      std::vector<int> ids;
      for (int u = 0; u < (int)units.size(); u++)
        if (kind_of[units[u].icase] == l) ids.push_back(u);
      std::stable_sort(ids.begin(), ids.end(),
                       [&](int a, int b) { return units[a].cost > units[b].cost; });
      int64_t load[kQueues] = {};
      for (int u : ids) {
        int best = 0;
        for (int q = 1; q < kQueues; q++)
          if (load[q] < load[best]) best = q;
        units[u].queue = best;
        load[best] += units[u].cost;
      }
    }
    std::stable_sort(p.items.begin(), p.items.end(), [&](const Item& a, const Item& b) {
      const Unit& ua = units[uid[std::make_pair(a.icase, a.group)]];
      const Unit& ub = units[uid[std::make_pair(b.icase, b.group)]];
      const int la = kind_of[a.icase], lb = kind_of[b.icase];
      if (la != lb) return la < lb;
      if (ua.queue != ub.queue) return ua.queue < ub.queue;
      if (ua.cost != ub.cost) return ua.cost > ub.cost;
      if (a.icase != b.icase) return a.icase < b.icase;
      if (a.group != b.group) return a.group < b.group;
      return a.obs0 < b.obs0;
    });
    int i = 0;
    for (int l = 0; l < nl; l++) {
      Launch& L = p.launches[l];
      L.item0 = i;
      int cap = 0;
      for (int q = 0; q < kQueues; q++) {
        L.qoff[q] = i - L.item0;
        while (i < (int)p.items.size() && kind_of[p.items[i].icase] == l &&
               units[uid[std::make_pair(p.items[i].icase, p.items[i].group)]].queue == q)
          i++;
      }
      L.qoff[kQueues] = i - L.item0;
      L.nitems = i - L.item0;
      for (int c = 0; c < nc; c++)
        if (kind_of[c] == l) cap = capacity[c];
      L.grid = std::min(L.nitems, cap);
    }
  }
  const int nl = (int)p.launches.size();
  std::vector<int> qoff((size_t)nl * (kQueues + 1));
  for (int l = 0; l < nl; l++)
    for (int q = 0; q <= kQueues; q++) qoff[(size_t)l * (kQueues + 1) + q] = p.launches[l].qoff[q];
  HIPCHK(ctx, hipMalloc(&p.d_tiles, p.tiles.size() * sizeof(Tile)));
  HIPCHK(ctx, hipMalloc(&p.d_items, p.items.size() * sizeof(Item)));
  HIPCHK(ctx, hipMalloc(&p.d_qoff, qoff.size() * sizeof(int)));
  HIPCHK(ctx, hipMalloc(&p.d_counter, (size_t)nl * kQueues * sizeof(int)));
  HIPCHK(ctx, hipMalloc(&p.d_cases, nc * sizeof(TonalCaseDev)));
  HIPCHK(ctx, hipMalloc(&p.d_slab, p.slab_bytes));
  HIPCHK(ctx, hipMalloc(&p.d_flags, nc * sizeof(uint32_t)));
  std::vector<TonalCaseDev> hc(nc);
  for (int c = 0; c < nc; c++) {
    const StagedCase& sc = st.cases[c];
    TonalCaseDev& d = hc[c];
    d.pos = sc.d_pos;
    d.attr = sc.d_attr[prec];
    d.obs = sc.d_obs;
    d.slab = (char*)p.d_slab + p.slab_off[c];
    d.pT = sc.d_out[0];
    d.pL = sc.d_out[1];
    d.pA = sc.d_out[2];
    d.nE = sc.nE;
    d.nObs = (int32_t)sc.nObs;
    d.G = p.G[c];
    d.flags = p.d_flags + c;
    d.k = sc.k;
  }
  HIPCHK(ctx, hipMemcpyAsync(p.d_tiles, p.tiles.data(), p.tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(p.d_items, p.items.data(), p.items.size() * sizeof(Item), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(p.d_qoff, qoff.data(), qoff.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(p.d_cases, hc.data(), nc * sizeof(TonalCaseDev), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  p.valid = true;
  return ACUM_OK;
}

// Enqueues one full tonal pass: clear queue heads, flags and partials, the eval launches, the
// reductions.  Only kernels (no memset nodes), so the pass can be captured into a graph.
int enqueue_pass(acum_ctx* ctx, Stage& st, int prec, bool with_events) {
  Plan& p = st.plan[prec];
  const bool mixed = prec == ACUM_MIXED;
  hipStream_t s = ctx->stream;
  const int nl = (int)p.launches.size();
  if (with_events) HIPCHK(ctx, hipEventRecord(ctx->ev[0], s));
  HIPCHK(ctx, launch_fill_zero(s, p.d_counter, (int64_t)nl * kQueues * sizeof(int)));
  HIPCHK(ctx, launch_fill_zero(s, p.d_flags, (int64_t)st.cases.size() * sizeof(uint32_t)));
  HIPCHK(ctx, launch_fill_zero(s, p.d_slab, p.slab_bytes));
  if (with_events) HIPCHK(ctx, hipEventRecord(ctx->ev[1], s));
  for (int l = 0; l < nl; l++) {
    const Launch& L = p.launches[l];
    if (L.nitems > 0)
      HIPCHK(ctx, launch_tonal_eval(mixed, L.moving, L.tail, L.window, L.grid, s, p.d_cases, p.d_tiles,
                                    p.d_items + L.item0, p.d_qoff + (size_t)l * (kQueues + 1),
                                    p.d_counter + (size_t)l * kQueues));
  }
  if (with_events) HIPCHK(ctx, hipEventRecord(ctx->ev[2], s));
  for (int c = 0; c < (int)st.cases.size(); c++)
    HIPCHK(ctx, launch_tonal_reduce(mixed, s, p.d_cases, c, (int)st.cases[c].nT, (int)st.cases[c].nObs));
  if (with_events) HIPCHK(ctx, hipEventRecord(ctx->ev[3], s));
  return ACUM_OK;
}

// Reads the cases' error flags after a pass.  A case planned below the full source-motion bound
// whose window overflowed moves to the next wider bound (*retry); any other error fails the run.
int check_flags(acum_ctx* ctx, Stage& st, int prec, bool* retry) {
  const int nc = (int)st.cases.size();
  std::vector<uint32_t> f(nc, 0u);
  HIPCHK(ctx, hipMemcpy(f.data(), st.plan[prec].d_flags, nc * sizeof(uint32_t), hipMemcpyDeviceToHost));
  *retry = false;
  uint32_t all = 0;
  const Plan& p = st.plan[prec];
  for (int c = 0; c < nc; c++) {
    StagedCase& sc = st.cases[c];
    if ((f[c] & kFlagWindow) && sc.motion_level < kMotionLevels - 1) {
      const acum_tune& t = p.tunes[c];
      const int cur = motion_permille_at(t, sc.motion_level);
      do
        sc.motion_level++;
      while (sc.motion_level < kMotionLevels - 1 && motion_permille_at(t, sc.motion_level) <= cur);
      *retry = true;
      f[c] &= ~kFlagWindow;
    }
    all |= f[c];
  }
  if (all & (kFlagNumeric | kFlagSweep))
    return ctx_fail(ctx, ACUM_ERR_NUMERIC,
                    "non-finite observer time, source on an observer, or |Mr| >= 1");
  if (*retry) return ACUM_OK;
  if (all & kFlagWindow)
    return ctx_fail(ctx, ACUM_ERR_PLAN, "a contribution fell outside its tile window");
  return ACUM_OK;
}

Stage* get_stage(acum_ctx* ctx, int32_t handle) {
  auto it = ctx->stages.find(handle);
  return it == ctx->stages.end() ? nullptr : it->second.get();
}

int run_stage_once(acum_ctx* ctx, Stage& st, int prec, int repeats, bool use_graph, bool* retry) {
  if (prec != ACUM_FP64 && prec != ACUM_MIXED) return ctx_fail(ctx, ACUM_ERR_ARG, "bad precision");
  if (repeats < 1) repeats = 1;
  struct Events {  // destroyed on every return
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  hipEvent_t &e0 = ev.e[0], &e1 = ev.e[1];
  HIPCHK(ctx, hipEventCreate(&e0));
  HIPCHK(ctx, hipEventCreate(&e1));
  HIPCHK(ctx, hipEventRecord(e0, ctx->stream));
  for (auto& sc : st.cases) {
    int rc = ensure_records(ctx, sc, prec);
    if (rc) return rc;
  }
  HIPCHK(ctx, hipEventRecord(e1, ctx->stream));
  int rc = build_plan(ctx, st, prec);
  if (rc) return rc;
  Plan& p = st.plan[prec];
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float prep = 0;
  HIPCHK(ctx, hipEventElapsedTime(&prep, e0, e1));
  double eval = 0, red = 0, total = 0;
  if (use_graph) {
    if (!p.exec) {
      HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue_pass(ctx, st, prec, false);
      hipGraph_t g = nullptr;
      hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
      if (rc || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc) return rc;
        HIPCHK(ctx, ce);
      }
      p.graph = g;
      HIPCHK(ctx, hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    for (int r = 0; r < repeats; r++) HIPCHK(ctx, hipGraphLaunch(p.exec, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
    total = ms / repeats;
  } else {
    HIPCHK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    for (int r = 0; r < repeats; r++) {
      rc = enqueue_pass(ctx, st, prec, true);
      if (rc) return rc;
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      float a = 0, b = 0;
      HIPCHK(ctx, hipEventElapsedTime(&a, ctx->ev[1], ctx->ev[2]));
      HIPCHK(ctx, hipEventElapsedTime(&b, ctx->ev[2], ctx->ev[3]));
      eval += a;
      red += b;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
    total = ms / repeats;
    eval /= repeats;
    red /= repeats;
  }
  acum_timing& t = ctx->timing;
  t = acum_timing();
  t.prep_ms = prep;
  t.eval_ms = eval;
  t.reduce_ms = red;
  t.total_ms = total;
  for (auto& sc : st.cases) t.evals += sc.nE * sc.nT * sc.nObs;
  t.partial_bytes = p.slab_bytes;
  t.n_partials = 0;
  for (int g : p.G) t.n_partials = std::max(t.n_partials, g);
  t.n_items = (int32_t)p.items.size();
  t.launches = 3 + (int32_t)p.launches.size() + (int32_t)st.cases.size();
  t.eval_launches = (int32_t)p.launches.size();
  for (auto& sc : st.cases) t.motion_widened_cases += sc.motion_level > 0 ? 1 : 0;
  return check_flags(ctx, st, prec, retry);
}

// Tiles are planned with a reduced bound on how far a source moves in observer time per step
// (acum_tune.motion_permille), which lengthens the passes; the kernel checks every footprint
// against its window, so a case whose bound was too tight is planned again with a wider bound
// (the next of 0.6, 0.8 and the full bound that widens it) and the pass is run again.  The plan
// of a case depends only on the case, so results agree between a case run alone and in a batch.
int run_stage(acum_ctx* ctx, Stage& st, int prec, int repeats, bool use_graph) {
  bool retry = false;
  int rc = run_stage_once(ctx, st, prec, repeats, use_graph, &retry);
  for (int attempt = 0; rc == ACUM_OK && retry && attempt < kMotionLevels - 1; attempt++) {
    for (auto& p : st.plan) free_plan(p);  // both precisions' plans hold the cases' tiles
    rc = run_stage_once(ctx, st, prec, repeats, use_graph, &retry);
  }
  if (rc == ACUM_OK && retry) return ctx_fail(ctx, ACUM_ERR_PLAN, "a contribution fell outside its tile window");
  return rc;
}

acum_tune default_tune(int prec) {
  acum_tune t{};
  // measured on the four-blade rotor (4 x 20,000 elements, 1024 steps, 1024 observers): the
  // eval kernel's occupancy is bound by its LDS window (16 one-wave workgroups per CU at 20
  // bins), and a 20-bin window with 3-bin tiles ran fastest in both precisions (0.26 s mixed,
  // 0.59 s FP64 per pass, against 0.29 s and 0.66 s with 24 bins and 0.49 s and 0.80 s with 16)
  t.window = 20;
  t.max_spread = 3;
  t.max_tile = 256;
  t.groups = 0;
  t.max_chunk = 1 << 20;
  t.waves = 1;
  t.motion_permille = kMotionPermille;
  return t;
}

// Successive halving over tile shapes for one case, timed on a view of the case restricted to
// its first kTuneObs observers (the view shares the case's device data; only its plans are its
// own and freed here).  Round 0 times every window x spread x source-motion candidate (with the
// context's tile length and other limits) for one pass; the best quarter goes on, joined by
// each survivor with the other tile length, for two passes; the best quarter of those for four
// passes, and the fastest of them wins.  A plan's first pass is timed like the others (its
// kernels are loaded and its buffers allocated before the timed events).  Windows above 24 bins
// are left out: their LDS halves the kernel's occupancy, which has never paid.
// The winner then runs once on the whole case, which keeps its motion bound only if the bound
// holds for every observer.
// The view's size sets how well its timings rank shapes for the whole case: on the four-blade
// rotor and four of the seeded uneven cases, shapes tuned on 512 observers ran 1.5-6.7% faster
// than the default shape on every case, while 128 and 256 observers picked shapes up to 2.7% and
// 3.4% slower on some (their work items are too small to show the whole case's launch tail).
constexpr int kTuneObs = 512;

int tune_case(acum_ctx* ctx, const StagedCase& sc, int prec, TuneEntry* out, int* passes) {
  const bool mixed = prec == ACUM_MIXED;
  const acum_tune keep = ctx->tune[prec];
  Stage view;
  view.cases.push_back(sc);
  StagedCase& v = view.cases[0];
  v.nObs = std::min<int64_t>(sc.nObs, kTuneObs);
  v.forced = true;
  struct Cand {
    acum_tune t;
    double ms;
  };
  std::vector<Cand> cand;
  for (int W : {16, 20, 24}) {
    if (!eval_window_supported(mixed, W)) continue;
    for (int sp : {2, 3, 4})
      for (int pm : {500, 600, 800}) {
        acum_tune t = keep;
        t.window = W;
        t.max_spread = sp;
        t.motion_permille = pm;
        t.waves = 1;
        cand.push_back(Cand{t, 0.0});
      }
  }
  auto time_one = [&](Cand& c, int n) -> int {
    v.force = c.t;
    v.motion_level = 0;
    for (auto& p : view.plan) free_plan(p);
    int rc = run_stage(ctx, view, prec, n, false);
    c.ms = rc == ACUM_OK ? ctx->timing.total_ms : 1e300;
    *passes += n;
    return rc == ACUM_ERR_PLAN ? ACUM_OK : rc;  // a shape whose windows cannot hold the case loses
  };
  // one untimed pass first, so no candidate's time holds the loading of the kernels
  int rc = time_one(cand[0], 1);
  for (int round = 0, n = 1; round < 3 && rc == ACUM_OK; round++, n *= 2) {
    for (auto& c : cand)
      if ((rc = time_one(c, n)) != ACUM_OK) break;
    if (rc) break;
    std::stable_sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) { return a.ms < b.ms; });
    if (round == 2) break;
    cand.resize(std::max<size_t>(1, (cand.size() + 3) / 4));
    if (round == 0) {  // the other tile length for the survivors
      const size_t m = cand.size();
      for (size_t i = 0; i < m; i++) {
        Cand c = cand[i];
        c.t.max_tile = keep.max_tile > 64 ? 64 : 256;
        cand.push_back(c);
      }
    }
  }
  for (auto& p : view.plan) free_plan(p);
  if (rc) return rc;
  if (cand.empty() || !(cand[0].ms < 1e300)) return ctx_fail(ctx, ACUM_ERR_PLAN, "no candidate tile shape ran");
  // The winner's motion bound held for the view's observers; one pass of the whole case checks
  // it for all of them, and a bound that overflowed there is replaced by the one that held, so
  // later plans in the bucket do not pay the rerun.
  {
    Stage whole;
    whole.cases.push_back(sc);
    StagedCase& w = whole.cases[0];
    w.forced = true;
    w.force = cand[0].t;
    w.motion_level = 0;
    rc = run_stage(ctx, whole, prec, 1, false);
    for (auto& p : whole.plan) free_plan(p);
    *passes += 1 + w.motion_level;
    if (rc) return rc;
    if (w.motion_level > 0) cand[0].t.motion_permille = motion_permille_at(cand[0].t, w.motion_level);
  }
  out->tune = cand[0].t;
  out->tune.waves = 1;
  out->ms = cand[0].ms;
  return ACUM_OK;
}

}  // namespace

// =============================================================================================
extern "C" {

int acum_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int acum_ctx_create(acum_ctx** out, int device) {
  if (!out) return ACUM_ERR_ARG;
  *out = nullptr;
  int n = acum_device_count();
  if (device < 0 || device >= n) return ACUM_ERR_DEVICE;
  acum_ctx* c = new acum_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    acum_ctx_destroy(c);  // releases what was created so far
    return ACUM_ERR_DEVICE;
  }
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      acum_ctx_destroy(c);
      return ACUM_ERR_DEVICE;
    }
  if (hipMalloc(&c->d_counter, 2 * kQueues * sizeof(int)) != hipSuccess ||
      hipMalloc(&c->d_flags, sizeof(uint32_t)) != hipSuccess) {
    acum_ctx_destroy(c);
    return ACUM_ERR_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    acum_ctx_destroy(c);
    return ACUM_ERR_DEVICE;
  }
  c->cu_count = prop.multiProcessorCount;
  // the device address of this GPU's scratch array of acum_tonal_spectra_noalloc
  if (spectra_scratch_init(device) != hipSuccess) {
    acum_ctx_destroy(c);
    return ACUM_ERR_DEVICE;
  }
  c->tune[ACUM_FP64] = default_tune(ACUM_FP64);
  c->tune[ACUM_MIXED] = default_tune(ACUM_MIXED);
  *out = c;
  return ACUM_OK;
}

void acum_ctx_destroy(acum_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  for (auto& kv : c->stages) free_stage(*kv.second);
  (void)hipFree(c->ws);
  for (int b = 0; b < 2; b++) {
    if (c->pin_ev[b]) (void)hipEventSynchronize(c->pin_ev[b]);
    if (c->pin[b]) (void)hipHostFree(c->pin[b]);
    if (c->pin_ev[b]) (void)hipEventDestroy(c->pin_ev[b]);
  }
  (void)hipFree(c->d_counter);
  (void)hipFree(c->d_flags);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* acum_last_error(const acum_ctx* c) { return c ? c->err.c_str() : "null context"; }

int acum_tonal_stage(acum_ctx* ctx, const acum_tonal_case* cases, int32_t nCases, int32_t* handle) {
  if (!ctx || !cases || nCases < 1 || !handle) return ctx_fail(ctx, ACUM_ERR_ARG, "bad argument");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto st = std::make_unique<Stage>();
  st->cases.resize(nCases);
  for (int i = 0; i < nCases; i++) {
    int rc = stage_case(ctx, &cases[i], st->cases[i]);
    if (rc) {
      free_stage(*st);
      return rc;
    }
  }
  *handle = ctx->next_handle++;
  ctx->stages[*handle] = std::move(st);
  return ACUM_OK;
}

int acum_run_stage(acum_ctx* ctx, const acum_run* run, int32_t* handle) {
  if (!ctx || !run || !handle) return ctx_fail(ctx, ACUM_ERR_ARG, "bad argument");
  if (run->preprocessed) {
    acum_tonal_case c{};
    if (acum_run_tonal_case(run, ACUM_MIXED, &c) != ACUM_OK)
      return ctx_fail(ctx, ACUM_ERR_ARG, "run cannot describe a tonal case");
    return acum_tonal_stage(ctx, &c, 1, handle);
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto st = std::make_unique<Stage>();
  st->cases.resize(1);
  int rc = stage_raw_run(ctx, run, st->cases[0]);
  if (rc) {
    free_stage(*st);
    return rc;
  }
  *handle = ctx->next_handle++;
  ctx->stages[*handle] = std::move(st);
  return ACUM_OK;
}

int acum_tonal_run(acum_ctx* ctx, int32_t handle, int32_t precision, int32_t repeats, int32_t use_graph) {
  if (!ctx) return ACUM_ERR_ARG;
  Stage* st = get_stage(ctx, handle);
  if (!st) return ctx_fail(ctx, ACUM_ERR_ARG, "unknown handle");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return run_stage(ctx, *st, precision, repeats, use_graph != 0);
}

int acum_tonal_fetch(acum_ctx* ctx, int32_t handle, int32_t icase, float* pT, float* pL, float* pA) {
  if (!ctx) return ACUM_ERR_ARG;
  Stage* st = get_stage(ctx, handle);
  if (!st || icase < 0 || icase >= (int)st->cases.size())
    return ctx_fail(ctx, ACUM_ERR_ARG, "unknown handle or case");
  StagedCase& sc = st->cases[icase];
  const size_t n = (size_t)sc.nT * sc.nObs * sizeof(float);
  float* dst[3] = {pT, pL, pA};
  for (int i = 0; i < 3; i++)
    if (dst[i]) HIPCHK(ctx, ctx_download(ctx, dst[i], sc.d_out[i], n));
  return ACUM_OK;
}

int acum_tonal_fetch_spectra(acum_ctx* ctx, int32_t handle, int32_t icase, int32_t nBands, const float* fc,
                             float* spl, float* oaspl) {
  if (!ctx) return ACUM_ERR_ARG;
  Stage* st = get_stage(ctx, handle);
  if (!st || icase < 0 || icase >= (int)st->cases.size())
    return ctx_fail(ctx, ACUM_ERR_ARG, "unknown handle or case");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  StagedCase& sc = st->cases[icase];
  return spectra_device(ctx, sc.d_out[2], sc.nT, sc.nObs, 1.0 / sc.k.KB, nBands, fc, spl, oaspl);
}

int acum_tonal_release(acum_ctx* ctx, int32_t handle) {
  if (!ctx) return ACUM_ERR_ARG;
  auto it = ctx->stages.find(handle);
  if (it == ctx->stages.end()) return ctx_fail(ctx, ACUM_ERR_ARG, "unknown handle");
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  free_stage(*it->second);
  ctx->stages.erase(it);
  return ACUM_OK;
}

int acum_tonal(acum_ctx* ctx, const acum_tonal_case* c, float* pT, float* pL, float* pA) {
  if (!ctx || !c || !pT || !pL || !pA) return ctx_fail(ctx, ACUM_ERR_ARG, "bad argument");
  int32_t h = 0;
  int rc = acum_tonal_stage(ctx, c, 1, &h);
  if (rc) return rc;
  rc = acum_tonal_run(ctx, h, c->precision, 1, 0);
  if (!rc) rc = acum_tonal_fetch(ctx, h, 0, pT, pL, pA);
  acum_tonal_release(ctx, h);
  return rc;
}

int acum_tonal_plain(acum_ctx* ctx, const acum_tonal_case* c, float* pT, float* pL, float* pA) {
  if (!ctx || !c || !pT || !pL || !pA) return ctx_fail(ctx, ACUM_ERR_ARG, "bad argument");
  Stage st;  // one staged case (the reference AoS arrays on the device), freed on every return
  struct Release {
    Stage& s;
    double* acc = nullptr;
    ~Release() {
      (void)hipFree(acc);
      free_stage(s);
    }
  } rel{st};
  st.cases.resize(1);
  StagedCase& sc = st.cases[0];
  int rc = stage_case(ctx, c, sc);
  if (rc) return rc;
  const int64_t n = sc.nT * sc.nObs;
  HIPCHK(ctx, hipMalloc(&rel.acc, (size_t)n * 3 * sizeof(double)));
  HIPCHK(ctx, launch_reset(ctx->stream, ctx->d_counter, ctx->d_flags));
  HIPCHK(ctx, launch_fill_zero(ctx->stream, rel.acc, n * 3 * (int64_t)sizeof(double)));
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  HIPCHK(ctx, launch_tonal_plain(ctx->stream, sc.d_aos[0], sc.d_aos[1], sc.d_aos[2], sc.d_aos[3], sc.d_obs,
                                 sc.nE, (int)sc.nT, (int)sc.nObs, c->dTau, c->dPsi, c->a0, c->rhoRef, c->oM,
                                 c->vInf, rel.acc, ctx->d_flags));
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
  HIPCHK(ctx, launch_plain_to_float(ctx->stream, rel.acc, sc.d_out, n, ctx->d_flags));
  HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
  uint32_t f = 0;
  HIPCHK(ctx, hipMemcpyAsync(&f, ctx->d_flags, sizeof(f), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float ev = 0, cv = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ev, ctx->ev[1], ctx->ev[2]));
  HIPCHK(ctx, hipEventElapsedTime(&cv, ctx->ev[2], ctx->ev[3]));
  acum_timing& t = ctx->timing;
  t = acum_timing();
  t.eval_ms = ev;
  t.reduce_ms = cv;
  t.total_ms = ev + cv;
  t.evals = sc.nE * sc.nT * sc.nObs;
  t.launches = 5;
  t.eval_launches = 1;
  if (f & kFlagNumeric)
    return ctx_fail(ctx, ACUM_ERR_NUMERIC, "plain path: r (1 - Mr) <= 0, observer time not increasing or a non-finite result");
  float* out[3] = {pT, pL, pA};
  for (int i = 0; i < 3; i++) HIPCHK(ctx, ctx_download(ctx, out[i], sc.d_out[i], (size_t)n * sizeof(float)));
  return ACUM_OK;
}

int acum_tonal_batch(acum_ctx* ctx, const acum_tonal_case* cases, int32_t nCases, float* const* pT,
                     float* const* pL, float* const* pA) {
  if (!ctx || !cases || nCases < 1 || !pT || !pL || !pA) return ctx_fail(ctx, ACUM_ERR_ARG, "bad argument");
  for (int i = 1; i < nCases; i++)
    if (cases[i].precision != cases[0].precision)
      return ctx_fail(ctx, ACUM_ERR_ARG, "all cases of a batch must use one precision");
  int32_t h = 0;
  int rc = acum_tonal_stage(ctx, cases, nCases, &h);
  if (rc) return rc;
  rc = acum_tonal_run(ctx, h, cases[0].precision, 1, 0);
  for (int i = 0; i < nCases && !rc; i++) rc = acum_tonal_fetch(ctx, h, i, pT[i], pL[i], pA[i]);
  acum_tonal_release(ctx, h);
  return rc;
}

int acum_tonal_bins(acum_ctx* ctx, const acum_tonal_case* c, int64_t iobs, int32_t* ipts,
                    int32_t* npi, int32_t* ndown, int32_t* nup) {
  if (!ctx || !c || !ipts || !npi || !ndown || !nup) return ctx_fail(ctx, ACUM_ERR_ARG, "bad argument");
  if (iobs < 0 || iobs >= c->nObs) return ctx_fail(ctx, ACUM_ERR_ARG, "observer out of range");
  int32_t h = 0;
  int rc = acum_tonal_stage(ctx, c, 1, &h);
  if (rc) return rc;
  Stage& st = *get_stage(ctx, h);
  StagedCase& sc = st.cases[0];
  rc = ensure_records(ctx, sc, c->precision);
  const int64_t n = sc.nE * sc.nT;
  int32_t* d = nullptr;
  if (!rc && hipMalloc(&d, 4 * n * sizeof(int32_t)) != hipSuccess) rc = ctx_fail(ctx, ACUM_ERR_NOMEM, "hipMalloc");
  if (!rc) {
    const double ox = c->rObs[3 * iobs], oy = c->rObs[3 * iobs + 1], oz = c->rObs[3 * iobs + 2];
    hipError_t e = launch_tonal_bins(ctx->stream, sc.d_pos, sc.nE, sc.k, ox, oy, oz, d, d + n,
                                     d + 2 * n, d + 3 * n);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    int32_t* dst[4] = {ipts, npi, ndown, nup};
    for (int i = 0; i < 4 && e == hipSuccess; i++)
      e = hipMemcpy(dst[i], d + i * n, n * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = ctx_fail(ctx, ACUM_ERR_DEVICE, hipGetErrorString(e));
  }
  (void)hipFree(d);
  acum_tonal_release(ctx, h);
  return rc;
}

int acum_get_timing(const acum_ctx* ctx, acum_timing* t) {
  if (!ctx || !t) return ACUM_ERR_ARG;
  *t = ctx->timing;
  return ACUM_OK;
}

int acum_get_tuning(const acum_ctx* ctx, int32_t precision, acum_tune* t) {
  if (!ctx || !t || (precision != ACUM_FP64 && precision != ACUM_MIXED)) return ACUM_ERR_ARG;
  *t = ctx->tune[precision];
  return ACUM_OK;
}

int acum_set_tuning(acum_ctx* ctx, int32_t precision, const acum_tune* t) {
  if (!ctx || !t || (precision != ACUM_FP64 && precision != ACUM_MIXED)) return ACUM_ERR_ARG;
  const bool mixed = precision == ACUM_MIXED;
  if (!eval_window_supported(mixed, t->window))
    return ctx_fail(ctx, ACUM_ERR_ARG, "window must be 16, 20, 24 or 32 (or 48, 64 for the mixed path)");
  if (t->max_spread < 1 || t->max_tile < 1 || t->max_chunk < 1 || t->groups < 0)
    return ctx_fail(ctx, ACUM_ERR_ARG, "bad tile limits");
  if (t->motion_permille != 0 && (t->motion_permille < 100 || t->motion_permille > 1000))
    return ctx_fail(ctx, ACUM_ERR_ARG, "motion_permille must be 0 (default) or 100..1000");
  ctx->tune[precision] = *t;
  ctx->tune[precision].waves = 1;
  return ACUM_OK;
}

int acum_autotune(acum_ctx* ctx, int32_t handle, int32_t precision, acum_tune* best, double* best_ms) {
  if (!ctx || (precision != ACUM_FP64 && precision != ACUM_MIXED)) return ACUM_ERR_ARG;
  Stage* st = get_stage(ctx, handle);
  if (!st) return ctx_fail(ctx, ACUM_ERR_ARG, "unknown handle");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  for (auto& sc : st->cases) {
    int rc = ensure_records(ctx, sc, precision);  // before any view copies the case
    if (rc) return rc;
  }
  int tuned = 0, hits = 0, passes = 0;
  for (auto& sc : st->cases) {
    const TuneKey key = tune_key(ctx, sc, precision);
    if (ctx->tune_cache.count(key)) {
      hits++;
      continue;
    }
    TuneEntry e{};
    int rc = tune_case(ctx, sc, precision, &e, &passes);
    if (rc) return rc;
    ctx->tune_cache[key] = e;
    tuned++;
  }
  for (auto& p : st->plan) free_plan(p);  // the next run plans with the cached shapes
  const auto it = ctx->tune_cache.find(tune_key(ctx, st->cases[0], precision));
  if (best) *best = it->second.tune;
  if (best_ms) *best_ms = it->second.ms;
  ctx->timing = acum_timing();
  ctx->timing.tune_cases = tuned;
  ctx->timing.tune_hits = hits;
  ctx->timing.tune_passes = passes;
  return ACUM_OK;
}

int acum_get_case_tuning(acum_ctx* ctx, int32_t handle, int32_t icase, int32_t precision, acum_tune* t) {
  if (!ctx || !t || (precision != ACUM_FP64 && precision != ACUM_MIXED)) return ACUM_ERR_ARG;
  Stage* st = get_stage(ctx, handle);
  if (!st || icase < 0 || icase >= (int)st->cases.size())
    return ctx_fail(ctx, ACUM_ERR_ARG, "unknown handle or case");
  *t = case_tune(ctx, st->cases[icase], precision);
  return ACUM_OK;
}

int acum_tune_cache_save(const acum_ctx* ctx, const char* path) {
  if (!ctx || !path) return ACUM_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return ACUM_ERR_IO;
  std::fprintf(f, "# acum tune cache: prec moving lgE lgT lgO cus | window spread tile groups chunk motion | ms\n");
  for (const auto& kv : ctx->tune_cache) {
    const TuneKey& k = kv.first;
    const acum_tune& t = kv.second.tune;
    std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d %d %d %.6f\n", k.prec, k.moving, k.lgE, k.lgT, k.lgO, k.cus,
                 t.window, t.max_spread, t.max_tile, t.groups, t.max_chunk, t.motion_permille, kv.second.ms);
  }
  return std::fclose(f) == 0 ? ACUM_OK : ACUM_ERR_IO;
}

int acum_tune_cache_load(acum_ctx* ctx, const char* path) {
  if (!ctx || !path) return ACUM_ERR_ARG;
  FILE* f = std::fopen(path, "r");
  if (!f) return ctx_fail(ctx, ACUM_ERR_IO, std::string("cannot open tune cache ") + path);
  char line[512];
  std::map<TuneKey, TuneEntry> add;
  int rc = ACUM_OK;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    TuneKey k{};
    acum_tune t{};
    double ms = 0;
    if (std::sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d %d %lf", &k.prec, &k.moving, &k.lgE, &k.lgT, &k.lgO,
                    &k.cus, &t.window, &t.max_spread, &t.max_tile, &t.groups, &t.max_chunk,
                    &t.motion_permille, &ms) != 13 ||
        (k.prec != ACUM_FP64 && k.prec != ACUM_MIXED) || !eval_window_supported(k.prec == ACUM_MIXED, t.window) ||
        t.max_spread < 1 || t.max_tile < 1 || t.max_chunk < 1 || t.groups < 0 ||
        (t.motion_permille != 0 && (t.motion_permille < 100 || t.motion_permille > 1000))) {
      rc = ctx_fail(ctx, ACUM_ERR_IO, std::string("malformed tune cache line: ") + line);
      break;
    }
    t.waves = 1;
    add[k] = TuneEntry{t, ms};
  }
  std::fclose(f);
  if (rc) return rc;
  for (auto& kv : add) ctx->tune_cache[kv.first] = kv.second;
  for (auto& kv : ctx->stages)  // plans made before see the new shapes on their next run
    for (auto& p : kv.second->plan) free_plan(p);
  return ACUM_OK;
}

int acum_tune_cache_clear(acum_ctx* ctx) {
  if (!ctx) return ACUM_ERR_ARG;
  ctx->tune_cache.clear();
  for (auto& kv : ctx->stages)
    for (auto& p : kv.second->plan) free_plan(p);
  return ACUM_OK;
}

}  // extern "C"
