// Broadband (BPM, NASA RP-1218) kernels for gfx950.
//
// The reference evaluates core() once per (observer, section, step) on one CPU thread and
// accumulates each band's level in dB (bb_core.cu:449-557, 591-1043).  Here the model is split by
// what each factor depends on:
//   * bpm_shape_kernel -- per (section, step), all bands: the observer-independent spectral
//     shapes of the five mechanisms as linear energies S (FP64 arithmetic, stored FP32):
//       TBL-TE suction/pressure/separation (bb_core.cu:630-845), LBL-VS (865-915),
//       TEB-VS bluntness (917-971) and tip vortex (975-1039);
//   * bpm_obs_kernel   -- per (observer, section, step): retarded position, directivities D_h,
//     D_l and 1/r^2 (FP32), then per band E = max(G S, floor) energy sums in FP32 registers over a
//     chunk of sources; one lane per observer, sources wave-uniform;
//   * bpm_final_kernel -- ordered FP64 Kahan sums of the chunk partials, the time average and
//     the overall level (bb_core.cu:565-583).
// In energy form the reference's per-step dB round trips cancel exactly: with every cell
// starting at -100 dB (1e-10), band level = 10 log10((1e-10 (1 + nT) + sum_{sect,t} E) / nT).
#include <hip/hip_runtime.h>
#include <cstdint>
#include "bpm_types.hpp"

namespace acum {

enum : int { kComp = 6 };  // a (separation), s, p, LBL-VS, bluntness, tip

__device__ __forceinline__ double p2(double x) { return x * x; }

__device__ double a_min(double a) {  // bb_utils.cu:6-25
  if (a < 0.204) return sqrt(67.552 - 886.788 * p2(a)) - 8.219;
  if (a > 0.244) return -142.795 * a * a * a + 103.656 * p2(a) - 57.757 * a + 6.006;
  return -32.665 * a + 3.981;
}
__device__ double a_max(double a) {  // bb_utils.cu:27-46
  if (a < 0.13) return sqrt(67.552 - 886.788 * p2(a)) - 8.219;
  if (a > 0.321) return -4.669 * a * a * a + 3.491 * p2(a) - 16.699 * a + 1.149;
  return -15.901 * a + 1.098;
}
__device__ double b_min(double b) {  // bb_utils.cu:48-67
  if (b < 0.13) return sqrt(16.888 - 886.788 * p2(b)) - 4.109;
  if (b > 0.145) return -817.810 * b * b * b + 355.21 * p2(b) - 135.024 * b + 10.619;
  return -83.607 * b + 8.138;
}
__device__ double b_max(double b) {  // bb_utils.cu:69-88
  if (b < 0.1) return sqrt(16.888 - 886.788 * p2(b)) - 4.109;
  if (b > 0.187) return -80.541 * b * b * b + 44.174 * p2(b) - 39.381 * b + 2.344;
  return -31.313 * b + 1.854;
}
__device__ double g5comp(double h, double eta) {  // bb_utils.cu:90-121
  double mu, m;
  if (h <= 0.25) mu = 0.1211;
  else if (h <= 0.62) mu = -0.2175 * h + 0.1755;
  else if (h < 1.15) mu = -0.0308 * h + 0.0596;
  else mu = 0.0242;
  if (h <= 0.02) m = 0.0;
  else if (h < 0.5) m = 68.724 * h - 1.35;
  else if (h <= 0.62) m = 308.475 * h - 121.23;
  else if (h <= 1.15) m = 224.811 * h - 69.354;
  else if (h < 1.2) m = 1583.28 * h - 1631.592;
  else m = 268.344;
  if (m < 0.0) m = 0.0;
  const double mu2 = mu * mu;
  const double eta0 = -sqrt((m * m * mu2 * mu2) / (6.25 + m * m * mu2));
  const double k = 2.5 * sqrt(1.0 - p2(eta0 / mu)) - 2.5 - m * eta0;
  if (eta <= eta0) return m * eta + k;
  if (eta <= 0.0) return 2.5 * sqrt(1.0 - p2(eta / mu)) - 2.5;
  if (eta <= 0.03616) return sqrt(1.5625 - 1194.99 * p2(eta)) - 1.25;
  return -155.543 * eta + 4.375;
}
__device__ __forceinline__ double a0_of(double rc) {
  if (rc < 9.52e4) return 0.57;
  if (rc > 8.57e5) return 1.13;
  return -9.57e-13 * p2(rc - 8.57e5) + 1.13;
}
__device__ __forceinline__ double b0_of(double rc) {
  if (rc < 9.52e4) return 0.3;
  if (rc > 8.57e5) return 0.56;
  return -4.48e-13 * p2(rc - 8.57e5) + 0.56;
}
__device__ __forceinline__ double k1_of(double rc) {
  if (rc < 2.47e5) return -4.31 * log10(rc) + 156.3;
  if (rc > 8.0e5) return 128.5;
  return -9.0 * log10(rc) + 181.6;
}
__device__ __forceinline__ double db2e(double x) { return exp10(0.1 * x); }

// One thread per source: the spectral shapes of every band.
__global__ void bpm_shape_kernel(const BpmShapeIn* __restrict__ in, int64_t nsrc, const float* __restrict__ ff,
                                 int nf2, double a0, float* __restrict__ S, BpmSrcDev* __restrict__ dev) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nsrc) return;
  const BpmShapeIn q = in[i];
  const double nu = 0.0000181206 / 1.22500;
  const double u = fabs(q.u), M = u / a0, rc = u * q.c / nu, alpha = q.alpha;
  const double M5 = pow(M, 5.0), lsect = q.lsect;
  const double st1 = 0.02 * pow(M, -0.6);
  const double st2 = (alpha < 1.33) ? st1 : (alpha > 12.5 ? st1 * 4.72 : st1 * exp10(0.0054 * p2(alpha - 1.33)));
  double a_0 = a0_of(rc);
  double amn = a_min(a_0), amx = a_max(a_0);
  double a_r = (-20.0 - amn) / (amx - amn);
  const double b_0 = b0_of(rc);
  const double bmn = b_min(b_0), bmx = b_max(b_0);
  const double b_r = (-20.0 - bmn) / (bmx - bmn);
  double k_1 = k1_of(rc);
  const double rdp = u * q.dsp / nu;
  const double dk1 = (rdp <= 5000) ? alpha * (1.43 * log10(rdp) - 5.29) : 0.0;
  const double gamm = 27.094 * M + 3.31, gamm0 = 23.43 * M + 4.651;
  const double beta = 72.65 * M + 10.74, beta0 = -34.19 * M - 13.82;
  auto k2_of = [&](double k1) {
    if (alpha < gamm0 - gamm) return k1 - 1000.0;
    if (alpha > gamm0 + gamm) return k1 - 12.0;
    return k1 + sqrt(p2(beta) - p2(beta / gamm) * p2(alpha - gamm0)) + beta0;
  };
  const double k1_ps = k_1;
  double k_2 = k2_of(k_1);
  const bool sep = (alpha >= gamm0 || alpha > 12.5);
  double a_r2 = a_r;
  if (sep) {  // bb_core.cu:781-826
    const double rcn = 3.0 * rc;
    a_0 = a0_of(rcn);
    k_1 = k1_of(rcn);
    k_2 = k2_of(k_1);
    amn = a_min(a_0);
    amx = a_max(a_0);
    a_r2 = (-20.0 - amn) / (amx - amn);
  }
  // LBL-VS constants (bb_core.cu:867-887)
  const bool lam = (q.trip == 0);
  double stpk = 1.0, g23 = 0.0;
  if (lam) {
    double stl;
    if (rc < 1.3e5) stl = 0.18;
    else if (rc > 1.3e5 && rc <= 4e5) stl = 0.001756 * pow(rc, 0.3931);
    else stl = 0.28;
    stpk = exp10(-0.04 * alpha) * stl;
    const double rc0 = (alpha <= 3.0) ? exp10(0.215 * alpha + 4.978) : exp10(0.120 * alpha + 5.263);
    const double d = rc / rc0, ld = log10(d);
    double g2;
    if (d <= 0.3237) g2 = 77.852 * ld + 15.328;
    else if (d <= 0.5689) g2 = 65.188 * ld + 9.125;
    else if (d <= 1.7579) g2 = -114.052 * ld * ld;
    else if (d <= 3.0889) g2 = -65.188 * ld + 9.125;
    else g2 = -77.852 * ld + 15.328;
    g23 = g2 + 171.04 - 3.03 * alpha;
  }
  // TEB-VS constants (bb_core.cu:917-937)
  const double h = 0.0005, psi = 14.0;
  const double hdstar = h / ((q.dss + q.dsp) / 2.0), dstarh = 1.0 / hdstar;
  const double aterm = 0.212 - 0.0045 * psi;
  const double stpeak = (hdstar >= 0.2) ? aterm / (1.0 + 0.235 * dstarh - 0.0132 * p2(dstarh))
                                        : 0.1 * hdstar + 0.095 - 0.00243 * psi;
  const double g4 = (hdstar <= 5.0) ? 17.5 * log10(hdstar) + 157.5 - 1.114 * psi : 169.7 - 1.114 * psi;
  const double hdstarp = 6.724 * p2(hdstar) - 4.019 * hdstar + 1.107;
  // tip vortex (bb_core.cu:975-1024)
  double tip_l = 0.0, um = 1.0, T0 = 0.0;
  if (q.tip) {
    const double AR = q.R / q.c;
    const double AR0[6] = {2.0, 2.67, 4.0, 6.0, 12.0, 24.0};
    const double al0[6] = {0.54, 0.62, 0.71, 0.79, 0.89, 0.95};
    double alprat;
    if (AR < AR0[0]) alprat = al0[0];
    else if (AR >= AR0[5]) alprat = al0[5];
    else {
      int kk = 0;
      while (kk < 6 && !(AR < AR0[kk])) kk++;
      alprat = al0[kk - 1] * (AR0[kk] - AR) / (AR0[kk] - AR0[kk - 1]) +
               al0[kk] * (AR - AR0[kk - 1]) / (AR0[kk] - AR0[kk - 1]);
    }
    const double at = alpha * alprat;
    tip_l = (fabs(at) <= 2.0) ? (0.023 + 0.0169 * at) * q.c : (0.0378 + 0.0095 * at) * q.c;
    const double mm = (1.0 + 0.036 * at) * M;
    um = mm * a0;
    T0 = M * M * mm * mm * mm * tip_l * tip_l;
  }
  const double cs = q.dss * M5 * lsect, cp = q.dsp * M5 * lsect, cl = q.dp * M5 * lsect;
  const double cb = pow(M, 5.5) * h * lsect;
  float* o = S + i * (int64_t)nf2 * kComp;
  for (int b = 0; b < nf2; b++) {
    const double f = ff[b];
    const double sts = f * q.dss / u, stp = f * q.dsp / u;
    const double as = fabs(log10(sts / st1)), ap = fabs(log10(stp / st1)), an = fabs(log10(sts / st2));
    const double aas = a_min(as) + a_r * (a_max(as) - a_min(as));
    const double aap = a_min(ap) + a_r * (a_max(ap) - a_min(ap));
    double ya;
    if (sep) ya = a_min(an) + a_r2 * (a_max(an) - a_min(an)) + k_2;
    else ya = b_min(an) + b_r * (b_max(an) - b_min(an)) + k_2;
    o[b * kComp + 0] = (float)(cs * db2e(ya));
    o[b * kComp + 1] = (float)(cs * db2e(aas + (k1_ps - 3.0)));
    o[b * kComp + 2] = (float)(cp * db2e(aap + (k1_ps - 3.0) + dk1));
    double slam = 0.0;
    if (lam) {
      const double e = (f * q.dp / u) / stpk, le = log10(e);
      double g1;
      if (e <= 0.5974) g1 = 39.8 * le - 11.12;
      else if (e <= 0.8545) g1 = 98.409 * le + 2.0;
      else if (e <= 1.17) g1 = -5.076 + sqrt(2.484 - 506.25 * le * le);
      else if (e <= 1.674) g1 = -98.409 * le + 2.0;
      else g1 = -39.80 * le - 11.12;
      slam = cl * db2e(g1 + g23);
    }
    o[b * kComp + 3] = (float)slam;
    const double eta = log10((f * h / u) / stpeak);
    const double g514 = g5comp(hdstar, eta), g50 = g5comp(hdstarp, eta);
    double g5 = g50 + 0.0714 * psi * (g514 - g50);
    if (g5 > 0) g5 = 0.0;
    const double f4 = g5comp(0.25, eta);
    if (g5 > f4) g5 = f4;
    o[b * kComp + 4] = (float)(cb * db2e(g4 + g5));
    double stip = 0.0;
    if (q.tip) {
      const double stpp = f * tip_l / um;
      stip = db2e(126.0 - 30.5 * p2(log10(stpp) + 0.3));
    }
    o[b * kComp + 5] = (float)stip;
  }
  dev[i].tipT0 = (float)T0;
  dev[i].flags = (sep ? 1u : 0u) | (lam ? 2u : 0u) | (q.tip ? 4u : 0u);
}

// max of two non-negative floats as one integer max (their bit patterns order as the values do),
// without the canonicalising maxes that an IEEE fmaxf of a loaded value costs
__device__ __forceinline__ float max_nonneg(float a, float b) {
  return __int_as_float(max(__float_as_int(a), __float_as_int(b)));
}

// One lane per observer, a chunk of sources per workgroup, NB bands from band b0.
template <int NB>
__global__ __launch_bounds__(64) void bpm_obs_kernel(const BpmSrcDev* __restrict__ src, const float* __restrict__ S,
                                                     const int64_t* __restrict__ bounds, int nf2,
                                                     const float* __restrict__ obs, int nObs, float vx, float vy,
                                                     float vz, float aMv, float inv_aMv, float* __restrict__ part) {
  const int o = blockIdx.x * 64 + threadIdx.x;
  const int b0 = blockIdx.y * NB;  // first band of this pass
  const int ch = blockIdx.z;       // source chunk [bounds[ch], bounds[ch + 1])
  const int oc = o < nObs ? o : nObs - 1;
  const float ox = obs[3 * oc], oy = obs[3 * oc + 1], oz = obs[3 * oc + 2];
  const int nb = min(NB, nf2 - b0);
  float acc[NB][kComp];
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int c = 0; c < kComp; c++) acc[b][c] = 0.0f;
  const int64_t s0 = bounds[ch], s1 = bounds[ch + 1];
  const float floor_e = 1.0e-10f, floor_t = 2.2204460492503131e-16f;
  for (int64_t s = s0; s < s1; s++) {
    const BpmSrcDev q = src[s];
    // retarded position (bb_core.cu:515-523), aa = aMv
    float rx = ox - q.xt[0], ry = oy - q.xt[1], rz = oz - q.xt[2];
    const float bq = rx * vx + ry * vy + rz * vz;
    const float cq = rx * rx + ry * ry + rz * rz;
    const float tau = (__builtin_amdgcn_sqrtf(bq * bq + aMv * cq) - bq) * inv_aMv;
    rx -= vx * tau;
    ry -= vy * tau;
    rz -= vz * tau;
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float ir2 = __builtin_amdgcn_rcpf(r2);
    const float p0 = rx * q.tr[0] + ry * q.tr[1] + rz * q.tr[2];
    const float p1 = rx * q.tr[3] + ry * q.tr[4] + rz * q.tr[5];
    const float pz = rx * q.tr[6] + ry * q.tr[7] + rz * q.tr[8];
    // theta = acos|p0/r|, phi = acos(|p1|/re), re = |(p1, pz)| (bb_core.cu:529-530) through
    // their cosines; re = 0 gives 0 * inf = NaN, which fminf turns into 1 as 0/0 did
    const float ct = fminf(fabsf(p0) * __builtin_amdgcn_rsqf(r2), 1.0f);
    const float cph = fminf(fabsf(p1) * __builtin_amdgcn_rsqf(p1 * p1 + pz * pz), 1.0f);
    const float sph2 = (1.0f - cph) * (1.0f + cph);
    const float sth2 = (1.0f - ct) * (1.0f + ct);
    const float d1 = 1.0f + q.M * ct, d2 = 1.0f + q.Mmc * ct;
    const float Gh = (1.0f - ct) * sph2 * __builtin_amdgcn_rcpf(d1 * d2 * d2) * ir2;  // D_h / r^2 (bb_core.cu:620)
    const float dl1 = d1 * d1;
    const float Gl = sth2 * sph2 * __builtin_amdgcn_rcpf(dl1 * dl1) * ir2;            // D_l / r^2 (bb_core.cu:624)
    const float Ga = (q.flags & 1u) ? Gl : Gh;
    const float tipE = fmaxf(q.tipT0 * Gh, floor_t);
    // max(G S, floor) = G max(S, floor / G) for G > 0: one max and one FMA per clipped
    // component.  G is held at >= 1e-30, where floor / G stays finite and, with S below 1e20 (a
    // 200 dB band), G S < floor already, so the clipped value is the floor as before.
    const float Gae = fmaxf(Ga, 1.0e-30f), Ghe = fmaxf(Gh, 1.0e-30f);
    const float fa = floor_e * __builtin_amdgcn_rcpf(Gae), fh = floor_e * __builtin_amdgcn_rcpf(Ghe);
    // the LBL-VS and tip shapes are zero where their mechanism is off, so those components are
    // added unconditionally; bands past nb read padding or the next source and are not stored
    const float* Sq = S + (s * nf2 + b0) * kComp;
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const float* Sb = Sq + b * kComp;
      acc[b][0] = fmaf(Gae, max_nonneg(Sb[0], fa), acc[b][0]);
      acc[b][1] = fmaf(Ghe, max_nonneg(Sb[1], fh), acc[b][1]);
      acc[b][2] = fmaf(Ghe, max_nonneg(Sb[2], fh), acc[b][2]);
      acc[b][3] = fmaf(Gh, Sb[3], acc[b][3]);
      acc[b][4] = fmaf(Gh, Sb[4], acc[b][4]);
      acc[b][5] = fmaf(tipE, Sb[5], acc[b][5]);
    }
  }
  if (o < nObs) {
    float* out = part + ((int64_t)ch * nObs + o) * nf2 * kComp + (int64_t)b0 * kComp;
#pragma unroll
    for (int b = 0; b < NB; b++)
      if (b < nb)
#pragma unroll
        for (int c = 0; c < kComp; c++) out[b * kComp + c] = acc[b][c];
  }
}

// Ordered FP64 sums over chunks, time average and overall level (bb_core.cu:565-583).  One
// workgroup per observer, one thread per (band, component) of up to kFinalThreads / 6 bands at a
// time: each thread sums its cell over the chunks in chunk order (Kahan; a chunk's cells are
// contiguous, so a workgroup's loads are too, and eight chunks' loads are issued before their
// sums); a thread per band then adds its six sums in component order, and thread 0 adds the band
// totals in band order.
constexpr int kFinalThreads = 1024;
constexpr int kFinalBatch = 8;

__global__ __launch_bounds__(kFinalThreads) void bpm_final_kernel(const float* __restrict__ part,
                                                                  int nchunk, int nObs, int nf2, int nT,
                                                                  float* __restrict__ spl,
                                                                  float* __restrict__ oaspl) {
  const int o = blockIdx.x;
  const int bpi = blockDim.x / kComp;  // bands per iteration
  __shared__ double sums[kFinalThreads];
  __shared__ double band_db[kFinalThreads / kComp];
  const double flo = 1.0e-10 * (1.0 + (double)nT);
  const int64_t cstride = (int64_t)nObs * nf2 * kComp;  // floats between chunks
  double oa = 1.0e-10;
  for (int bb = 0; bb < nf2; bb += bpi) {
    const int nbi = min(bpi, nf2 - bb);
    const int j = threadIdx.x;  // cell (band bb + j / 6, component j % 6)
    if (j < nbi * kComp) {
      const float* p = part + ((int64_t)o * nf2 + bb) * kComp + j;
      double sum = 0.0, cmp = 0.0;
      int ch = 0;
      for (; ch + kFinalBatch <= nchunk; ch += kFinalBatch) {
        float v[kFinalBatch];
#pragma unroll
        for (int u = 0; u < kFinalBatch; u++) v[u] = p[(int64_t)(ch + u) * cstride];
#pragma unroll
        for (int u = 0; u < kFinalBatch; u++) {
          const double y = (double)v[u] - cmp, t = sum + y;
          cmp = (t - sum) - y;
          sum = t;
        }
      }
      for (; ch < nchunk; ch++) {
        const double y = (double)p[(int64_t)ch * cstride] - cmp, t = sum + y;
        cmp = (t - sum) - y;
        sum = t;
      }
      sums[j] = sum;
      spl[((int64_t)o * nf2 + bb + j / kComp) * 7 + 1 + j % kComp] = (float)(10.0 * log10((flo + sum) / nT));
    }
    __syncthreads();
    if (threadIdx.x < nbi) {
      const int bl = threadIdx.x;
      double tot = flo;
      for (int c = 0; c < kComp; c++) tot += sums[bl * kComp + c];
      const double t0 = 10.0 * log10(tot / nT);
      spl[((int64_t)o * nf2 + bb + bl) * 7] = (float)t0;
      band_db[bl] = t0;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int k = 0; k < nbi; k++) oa += exp10(0.1 * band_db[k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) oaspl[o] = (float)(10.0 * log10(oa));
}

hipError_t launch_bpm_shape(hipStream_t s, const BpmShapeIn* in, int64_t nsrc, const float* ff, int nf2,
                            double a0, float* S, BpmSrcDev* dev) {
  const int grid = (int)((nsrc + 127) / 128);
  hipLaunchKernelGGL(bpm_shape_kernel, dim3(grid), dim3(128), 0, s, in, nsrc, ff, nf2, a0, S, dev);
  return hipGetLastError();
}

// Bands per pass of the observer kernel: the geometry of a (source, observer) pair is recomputed
// once per pass, and each band holds six FP32 accumulators per lane; at most 17 bands a pass.
int bpm_bands_per_pass(int nf2) {
  const int passes = (nf2 + 16) / 17;
  const int nb = (nf2 + passes - 1) / passes;
  return nb <= 12 ? 12 : (nb <= 14 ? 14 : 17);
}

// One launch over (observer tile, band pass, source chunk), the chunk slowest: blocks dispatch in
// index order, so with chunks of decreasing size the launch ends on the smallest ones.
hipError_t launch_bpm_obs(hipStream_t s, const BpmSrcDev* src, const float* S, const int64_t* bounds, int nchunk,
                          int nf2, const float* obs, int nObs, float vx, float vy, float vz, float aMv,
                          float* part) {
  const float inv_aMv = (float)(1.0 / (double)aMv);
  const int NB = bpm_bands_per_pass(nf2);
  const dim3 g((nObs + 63) / 64, (nf2 + NB - 1) / NB, nchunk);
  if (NB == 17)
    hipLaunchKernelGGL(bpm_obs_kernel<17>, g, dim3(64), 0, s, src, S, bounds, nf2, obs, nObs, vx, vy, vz, aMv,
                       inv_aMv, part);
  else if (NB == 14)
    hipLaunchKernelGGL(bpm_obs_kernel<14>, g, dim3(64), 0, s, src, S, bounds, nf2, obs, nObs, vx, vy, vz, aMv,
                       inv_aMv, part);
  else
    hipLaunchKernelGGL(bpm_obs_kernel<12>, g, dim3(64), 0, s, src, S, bounds, nf2, obs, nObs, vx, vy, vz, aMv,
                       inv_aMv, part);
  return hipGetLastError();
}

hipError_t launch_bpm_final(hipStream_t s, const float* part, int nchunk, int nObs, int nf2, int nT, float* spl,
                            float* oaspl) {
  // six cells per band, rounded up to whole waves
  const int64_t cells = ((int64_t)nf2 * kComp + 63) / 64 * 64;
  const int threads = cells < kFinalThreads ? (int)cells : kFinalThreads;
  hipLaunchKernelGGL(bpm_final_kernel, dim3(nObs), dim3(threads), 0, s, part, nchunk, nObs, nf2, nT, spl,
                     oaspl);
  return hipGetLastError();
}

int bpm_occupancy_blocks(int nf2) {
  int b = 0;
  const int NB = bpm_bands_per_pass(nf2);
  hipError_t e = NB == 17   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, bpm_obs_kernel<17>, 64, 0)
                 : NB == 14 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, bpm_obs_kernel<14>, 64, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, bpm_obs_kernel<12>, 64, 0);
  return e == hipSuccess ? b : 4;
}

}  // namespace acum
