// ORACLE -- test infrastructure only.  CPU restatement of the reference's tonal path:
// findPressureTimeHistory (core.cu:18-259) with its helpers acous_dist (core.cu:261-279),
// sumprod/vecmag (core.cu:281-288), findpts (core.cu:290-322) and bsearch (core.cu:325-355).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this code.
//
// Two instances:
//   orc_tonal_f64  FP64 everywhere: the parity truth for the GPU paths.
//   orc_tonal_f32  float integrand and float accumulation with the same FP64 observer time as
//                  orc_tonal_f64 (the "reference-faithful" CPU baseline).
// Summation order (the reference's is undefined: unordered float atomics, core.cu:202-226):
//   per observer -> surface -> element -> source step, each bin accumulated in that order.
//
// Observer time and the tent are written in the canonical FP64 form shared with the GPU
// kernels (same operations, same order, explicit fma, no other contraction), so bin indices
// agree exactly.  Time is measured in observer-time bins (units of dPsi).  With KB = oM / dPsi,
// na = a0^2 - |vinf|^2, C2 = KB (KB / na), w = vinf (KB / na) (zero in still air), SB = dTau KB,
// aB = a0 / KB and the origin O = the centre of the bounding box of every source position:
//   per source (step j):  p = src - O,  q = C2 fma(pz,pz, fma(py,py, px px)),  m = (-2 C2) p,
//                         wp = fma(pz,wz, fma(py,wy, px wx)),  jt = j SB
//   per observer:         o = obs - O,  qo = C2 fma(oz,oz, fma(oy,oy, ox ox)),
//                         wo = fma(oz,wz, fma(oy,wy, ox wx))
//   per eval:             cc = fma(ox,mx, fma(oy,my, fma(oz,mz, qo + q)))      = C2 |obs - src|^2
//                         b = wo - wp;  T = sqrt(fma(b, b, cc)) - b;  tau = T + jt;  r = aB T
// (still air: b = 0, T = sqrt(cc)).  T = KB t is the delay in bins, t the root
// (-b' - sqrt(b'^2 - 4 a' c)) / (2 a') of a' t^2 + b' t + c = 0 with a' = |vinf|^2 - a0^2 < 0,
// b' = -2 d.vinf, c = |d|^2 that acous_dist takes (core.cu:261-279): t = (s - d.vinf) / na with
// s = sqrt((d.vinf)^2 + c na), scaled by KB inside the square root; tau = (j dTau + r/a0) oM/dPsi
// (core.cu:156).  |obs - src|^2 is expanded about the origin so that each eval needs three fmas
// and one add for it (the rounding error of the expansion is that of |o|^2 + |p|^2, which with
// the origin at the sources' centre is about |obs - src|^2 for observers outside the sources).
// findpts/bsearch (core.cu:290-355) pick the grid time at or below the shifted observer time
// and the two sweeps (core.cu:196-231) visit the grid times in (ttem, ttemp] (weights rising
// from 0 to 1) and in (ttemp, ttep) (weights falling from 1 to 0), wrapping periodically.  In
// bin units the grid times are the integers: the rising bins are the integers b in
// (tau_prev, tau] with weight (b - tau_prev)/(tau - tau_prev), the falling bins the integers in
// (tau, tau_next) with weight (tau_next - b)/(tau_next - tau), and bin b is stored at b mod nT.
// The reference's radian-domain steps (npi shift, repeated tObs2 -= dPsi) select the same bins
// in exact arithmetic; tests check the two forms agree on sampled cases.
//
// Error codes follow the library's contract (include/acum.h), not the reference, which computes
// on with whatever the arithmetic gives: a non-finite observer position is ACUM_ERR_ARG; an
// observer within 1e-6 m of a source position (a source on an observer; |obs - src|^2 from the
// parsed floats), an eval with r <= 0 or 1 - Mr <= 0, observer time that does not increase from
// one step to the next, or a non-finite observer time or integrand is ACUM_ERR_NUMERIC.
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>
#include <omp.h>
#include "../include/acum.h"

namespace {

struct TimeConst {
  double a0, r2b, KB, SB;
  double vx, vy, vz;
  double wx, wy, wz, C2, aB;  // moving medium: w = vinf KB / na; C2 = KB^2 / na; aB = a0 / KB
  double Ox, Oy, Oz;          // origin: centre of the sources' bounding box
  double m2C2;                // -2 C2
  int64_t nT;
  bool still;  // vinf == 0
};

TimeConst make_time_const(const acum_tonal_case* c) {
  TimeConst k;
  k.a0 = c->a0;
  k.r2b = (c->oM / c->a0) / c->dPsi;
  k.KB = c->oM / c->dPsi;
  k.SB = c->dTau * k.KB;
  k.nT = c->nTime;
  k.vx = c->vInf[0];
  k.vy = c->vInf[1];
  k.vz = c->vInf[2];
  double vv = k.vx * k.vx;
  vv = vv + k.vy * k.vy;
  vv = vv + k.vz * k.vz;
  const double na = k.a0 * k.a0 - vv;
  const double kn = k.KB / na;
  k.wx = k.vx * kn;
  k.wy = k.vy * kn;
  k.wz = k.vz * kn;
  k.C2 = k.KB * kn;
  k.aB = k.a0 / k.KB;
  k.m2C2 = -2.0 * k.C2;
  k.still = (k.vx == 0.0 && k.vy == 0.0 && k.vz == 0.0);
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t s = 0; s < c->nSurf; s++) {
    const float* X = c->X + c->XPtr[s];
    for (int64_t i = 0; i < c->E[s] * c->nTime; i++)
      for (int d = 0; d < 3; d++) {
        const double v = X[4 * i + d];
        lo[d] = std::min(lo[d], v);
        hi[d] = std::max(hi[d], v);
      }
  }
  k.Ox = (lo[0] + hi[0]) * 0.5;
  k.Oy = (lo[1] + hi[1]) * 0.5;
  k.Oz = (lo[2] + hi[2]) * 0.5;
  return k;
}

// Per-observer constants of the canonical observer time.
struct ObsConst {
  double ox, oy, oz, qo, wo;
};

inline ObsConst obs_const(const TimeConst& k, double obx, double oby, double obz) {
  ObsConst o;
  o.ox = obx - k.Ox;
  o.oy = oby - k.Oy;
  o.oz = obz - k.Oz;
  o.qo = std::fma(o.oz, o.oz, std::fma(o.oy, o.oy, o.ox * o.ox)) * k.C2;
  o.wo = std::fma(o.oz, k.wz, std::fma(o.oy, k.wy, o.ox * k.wx));
  return o;
}

// Observer time of step j in bins and the retarded distance r in metres (acous_dist,
// core.cu:261-279; ttemp, core.cu:156), in the canonical form above.
inline double obs_tau_r(const TimeConst& k, int64_t j, const ObsConst& o, const float* src, double* r) {
  const double px = (double)src[0] - k.Ox;
  const double py = (double)src[1] - k.Oy;
  const double pz = (double)src[2] - k.Oz;
  const double q = std::fma(pz, pz, std::fma(py, py, px * px)) * k.C2;
  const double mx = px * k.m2C2, my = py * k.m2C2, mz = pz * k.m2C2;
  const double wp = std::fma(pz, k.wz, std::fma(py, k.wy, px * k.wx));
  const double jt = (double)j * k.SB;
  const double cc = std::fma(o.ox, mx, std::fma(o.oy, my, std::fma(o.oz, mz, o.qo + q)));
  double T;
  if (k.still) {
    T = std::sqrt(cc);
  } else {
    const double b = o.wo - wp;
    T = std::sqrt(std::fma(b, b, cc)) - b;
  }
  *r = T * k.aB;
  return T + jt;
}

// an observer this close (metres) to a source position is a source on an observer
constexpr double kCoincident = 1e-6;

struct Eval {
  double pT, pL;
};

// Farassat 1A integrand of the impermeable surface (core.cu:68-154), arithmetic type R.
// rho = rhodot = 1 and U = V (core.cu:71-72,109-111), so Un - Vn = 0 exactly; the dotted normal
// velocities are forced to zero (core.cu:117-118).  The terms that vanish identically (Mo,
// Modot, LF2, the (Un-Vn) parts of Ma_n/Man_dot) add an exact zero and are left out.
// Returns false for an eval the library reports as ACUM_ERR_NUMERIC: r <= 0, 1 - Mr <= 0 or a
// non-finite result.
template <typename R>
inline bool integrand(const float* X, const float* Nn, const float* Vv, const float* PR,
                      double ox, double oy, double oz, double r_d, const TimeConst& k,
                      R rhoRef, R* pThick, R* pLoad) {
  const R a0 = (R)k.a0;
  const R r = (R)r_d;
  const R obs[3] = {(R)ox, (R)oy, (R)oz};
  const R vinf[3] = {(R)k.vx, (R)k.vy, (R)k.vz};
  R rvec[3], rhat[3];
  for (int i = 0; i < 3; i++) rvec[i] = obs[i] - (R)X[i] - vinf[i] * r / a0;  // core.cu:86
  for (int i = 0; i < 3; i++) rhat[i] = rvec[i] / r;                          // core.cu:87
  const R V[3] = {(R)Vv[0], (R)Vv[1], (R)Vv[2]};
  const R Vd[3] = {(R)Vv[3], (R)Vv[4], (R)Vv[5]};
  const R N[3] = {(R)Nn[0], (R)Nn[1], (R)Nn[2]};
  const R Nd[3] = {(R)Nn[3], (R)Nn[4], (R)Nn[5]};
  const R ds = (R)X[3];
  const R P = (R)PR[0], Pdot = (R)PR[1];
  auto dot = [](const R* a, const R* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  const R Mr = dot(V, rhat) / a0;                                    // core.cu:93
  const R M = std::sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) / a0; // core.cu:94
  const R Mdot_r = dot(Vd, rhat) / a0;                               // core.cu:95
  const R D1 = (R)1.0 / ((R)4.0 * (R)M_PI * a0);                     // core.cu:53
  const R om = (R)1.0 - Mr;
  const R T1 = D1 / (r * (om * om));                                  // core.cu:97-100
  const R T2 = T1 / om;
  const R T3 = T1 / r;
  const R T4 = T3 / om;
  const R Vn = dot(V, N);                                             // core.cu:102
  const R Ma_n = rhoRef * Vn;                                         // core.cu:121
  const R TF2 = Ma_n * Mdot_r * T2 * a0 * ds;                         // core.cu:123
  const R TN = a0 * Ma_n * (Mr - M * M) * T4 * a0 * ds;               // core.cu:124
  *pThick = TN + TF2;                                                 // core.cu:128
  R Li[3], Lid[3];
  for (int i = 0; i < 3; i++) Li[i] = P * N[i];                       // core.cu:130
  for (int i = 0; i < 3; i++) Lid[i] = Pdot * N[i] + P * Nd[i];       // core.cu:131
  const R Lidot_r = dot(Lid, rhat);
  const R Li_r = dot(Li, rhat);
  const R Li_m = dot(Li, V) / a0;
  const R LF1 = Lidot_r * T1 * ds + Li_r * Mdot_r * T2 * ds;          // core.cu:138
  const R LN = (Li_r - Li_m) * T3 * a0 * ds + Li_r * (Mr - M * M) * T4 * a0 * ds; // core.cu:141
  *pLoad = LN + LF1;                                                  // core.cu:154
  return r_d > 0.0 && om > (R)0 && std::isfinite(*pThick) && std::isfinite(*pLoad);
}

// Tent binning of one eval (core.cu:166-231) in bin units: add(bin, w) for every rising bin,
// top down, then every falling bin, bottom up (the reference's order).  Returns false for a
// non-finite observer time or a tent longer than the period.
template <typename Add>
inline bool bin_eval(const TimeConst& k, double tp, double tc, double tn, int64_t* nd, int64_t* nu,
                     int64_t* ic_o, Add add) {
  if (!std::isfinite(tp) || !std::isfinite(tc) || !std::isfinite(tn)) return false;
  const int64_t nT = k.nT;
  const double fc = std::floor(tc);
  const int64_t ic = (int64_t)fc;
  auto phys = [nT](int64_t b) {
    int64_t m = b % nT;
    return m < 0 ? m + nT : m;
  };
  int64_t cnt = 0;
  for (int64_t b = ic; (double)b > tp; b--) {  // rising side: (tau_prev, tau]
    add(phys(b), ((double)b - tp) / (tc - tp));
    if (++cnt > nT) return false;
  }
  *nd = cnt;
  cnt = 0;
  for (int64_t b = ic + 1; (double)b < tn; b++) {  // falling side: (tau, tau_next)
    add(phys(b), (tn - (double)b) / (tn - tc));
    if (++cnt > nT) return false;
  }
  *nu = cnt;
  *ic_o = ic;
  return true;
}

template <typename R, typename Out>
int tonal_impl(const acum_tonal_case* c, Out* pT, Out* pL, Out* pA, int nthreads) {
  if (!c || c->nTime < 2 || c->nObs < 1 || c->nSurf < 1) return ACUM_ERR_ARG;
  for (int64_t i = 0; i < 3 * c->nObs; i++)
    if (!std::isfinite(c->rObs[i])) return ACUM_ERR_ARG;
  const TimeConst k = make_time_const(c);
  const int64_t nT = c->nTime, nObs = c->nObs;
  int status = ACUM_OK;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
    std::vector<R> bT(nT), bL(nT), bA(nT);
    std::vector<double> rr(nT), tt(nT);
#pragma omp for schedule(dynamic, 1)
    for (int64_t o = 0; o < nObs; o++) {
      std::fill(bT.begin(), bT.end(), (R)0);
      std::fill(bL.begin(), bL.end(), (R)0);
      std::fill(bA.begin(), bA.end(), (R)0);
      const double ox = c->rObs[3 * o], oy = c->rObs[3 * o + 1], oz = c->rObs[3 * o + 2];
      const ObsConst oc = obs_const(k, ox, oy, oz);
      bool ok = true;
      for (int64_t s = 0; s < c->nSurf && ok; s++) {
        for (int64_t j = 0; j < c->E[s] && ok; j++) {
          const float* Xe = c->X + c->XPtr[s] + j * nT * 4;
          const float* Ne = c->N + c->NPtr[s] + j * nT * 6;
          const float* Ve = c->V + c->VPtr[s] + j * nT * 6;
          const float* Pe = c->PR + c->PRPtr[s] + j * nT * 2;
          // One retarded solve per (element, step); the reference solves three times
          // (core.cu:84,189,193) with identical operands, hence identical values.
          for (int64_t t = 0; t < nT; t++) {
            tt[t] = obs_tau_r(k, t, oc, Xe + 4 * t, &rr[t]);
            // a source on the observer: within kCoincident of it, |obs - src|^2 from the floats
            const double dx = ox - (double)Xe[4 * t], dy = oy - (double)Xe[4 * t + 1],
                         dz = oz - (double)Xe[4 * t + 2];
            if (std::fma(dz, dz, std::fma(dy, dy, dx * dx)) < kCoincident * kCoincident) ok = false;
          }
          for (int64_t t = 0; t < nT && ok; t++) {
            const int64_t mm = (t > 0) ? t - 1 : nT - 1;       // core.cu:173-179
            const int64_t mp = (t < nT - 1) ? t + 1 : 0;       // core.cu:180-186
            const double tc = tt[t];
            double tp = tt[mm];
            if (t == 0) tp = tp - (double)nT;  // previous step of the periodic history
            double tn = tt[mp];
            if (t == nT - 1) tn = tn + (double)nT;
            R pth, plo;
            ok = integrand<R>(Xe + 4 * t, Ne + 6 * t, Ve + 6 * t, Pe + 2 * t, ox, oy, oz, rr[t], k,
                              (R)c->rhoRef, &pth, &plo);
            ok = ok && tc > tp;  // observer time increases from step to step
            int64_t nd, nu, ic;
            ok = ok && bin_eval(k, tp, tc, tn, &nd, &nu, &ic, [&](int64_t b, double w) {
              const R wr = (R)w;
              bT[b] += pth * wr;
              bL[b] += plo * wr;
              bA[b] += (pth + plo) * wr;
            });
          }
        }
      }
      if (!ok) {
#pragma omp critical
        status = ACUM_ERR_NUMERIC;
      }
      for (int64_t t = 0; t < nT; t++) {
        pT[t * nObs + o] = (Out)bT[t];
        pL[t * nObs + o] = (Out)bL[t];
        pA[t * nObs + o] = (Out)bA[t];
      }
    }
  }
  return status;
}

}  // namespace

extern "C" {

int orc_tonal_f64(const acum_tonal_case* c, double* pT, double* pL, double* pA, int nthreads) {
  return tonal_impl<double, double>(c, pT, pL, pA, nthreads);
}

int orc_tonal_f32(const acum_tonal_case* c, float* pT, float* pL, float* pA, int nthreads) {
  return tonal_impl<float, float>(c, pT, pL, pA, nthreads);
}

// Tent of every eval of one observer, in case order (surface, element, step): the top rising
// bin as (ic mod nT, floor(ic / nT)) and the two side lengths.
int orc_tonal_bins(const acum_tonal_case* c, int64_t iobs, int32_t* ipts, int32_t* npi,
                   int32_t* ndown, int32_t* nup) {
  if (!c || iobs < 0 || iobs >= c->nObs) return ACUM_ERR_ARG;
  const TimeConst k = make_time_const(c);
  const int64_t nT = c->nTime;
  const ObsConst oc = obs_const(k, c->rObs[3 * iobs], c->rObs[3 * iobs + 1], c->rObs[3 * iobs + 2]);
  std::vector<double> rr(nT), tt(nT);
  int64_t idx = 0;
  for (int64_t s = 0; s < c->nSurf; s++) {
    for (int64_t j = 0; j < c->E[s]; j++) {
      const float* Xe = c->X + c->XPtr[s] + j * nT * 4;
      for (int64_t t = 0; t < nT; t++) tt[t] = obs_tau_r(k, t, oc, Xe + 4 * t, &rr[t]);
      for (int64_t t = 0; t < nT; t++, idx++) {
        const int64_t mm = (t > 0) ? t - 1 : nT - 1;
        const int64_t mp = (t < nT - 1) ? t + 1 : 0;
        const double tc = tt[t];
        double tp = tt[mm];
        if (t == 0) tp = tp - (double)nT;
        double tn = tt[mp];
        if (t == nT - 1) tn = tn + (double)nT;
        int64_t nd = -1, nu = -1, ic = 0;
        if (!bin_eval(k, tp, tc, tn, &nd, &nu, &ic, [](int64_t, double) {})) return ACUM_ERR_NUMERIC;
        int64_t per = ic / nT, ph = ic % nT;
        if (ph < 0) {
          ph += nT;
          per -= 1;
        }
        ipts[idx] = (int32_t)ph;
        npi[idx] = (int32_t)per;
        ndown[idx] = (int32_t)nd;
        nup[idx] = (int32_t)nu;
      }
    }
  }
  return ACUM_OK;
}

// Observer time (bins) of one (element position, step) in the canonical FP64 form (for tests).
double orc_obs_tau(const acum_tonal_case* c, const float* src, const float* obs, int64_t j) {
  const TimeConst k = make_time_const(c);
  double r;
  return obs_tau_r(k, j, obs_const(k, obs[0], obs[1], obs[2]), src, &r);
}

double orc_retarded_r(const acum_tonal_case* c, const float* src, const float* obs) {
  const TimeConst k = make_time_const(c);
  double r;
  obs_tau_r(k, 0, obs_const(k, obs[0], obs[1], obs[2]), src, &r);
  return r;
}

}  // extern "C"
