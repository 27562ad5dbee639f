// preprocess_imper (preprocess.cu:228-413) for the acoustic kernels, in FP64:
//   * length scaling of x, y, z and dS (preprocess.cu:275-278);
//   * CP -> gauge pressure P = CP * 0.5 rhoRef Mref^2 a0^2 (preprocess.cu:279);
//   * make_periodic(P), always (preprocess.cu:321, mathfunc.cu:165-206);
//   * optional low-pass: two rounds of 3-point median + 13-point Gaussian (preprocess.cu:323-333,
//     mathfunc.cu:208-421);
//   * dP/dt by periodic central differences, or the biased form when periodic = 0
//     (preprocess.cu:344-348, mathfunc.cu:132-159);
//   * dn/dt, V = dx/dt and dV/dt by periodic central differences (preprocess.cu:351-364).
// Every series is formed in double from the float inputs and rounded to float once at the end,
// so dP/dt, dn/dt, V and dV/dt carry no float cancellation error.  Elements are independent and
// run in parallel.
#include <cmath>
#include <vector>
#include "run.hpp"

using namespace acum_host;

namespace {

void central(const double* u, double* ud, double dt, int n) {
  ud[0] = (u[1] - u[n - 1]) / (2.0 * dt);
  ud[n - 1] = (u[0] - u[n - 2]) / (2.0 * dt);
  for (int i = 1; i < n - 1; i++) ud[i] = (u[i + 1] - u[i - 1]) / (2.0 * dt);
}

void biased(const double* u, double* ud, double dt, int n) {
  ud[0] = (u[1] - u[0]) / dt;
  ud[n - 1] = (u[n - 1] - u[n - 2]) / dt;
  for (int i = 1; i < n - 1; i++) ud[i] = (u[i + 1] - u[i - 1]) / (2.0 * dt);
  ud[0] = 3 * ud[1] - 3 * ud[2] + ud[3];
  ud[n - 1] = 3 * ud[n - 2] - 3 * ud[n - 3] + ud[n - 4];
}

// Subtracts the cubic that makes the series' end values and end slopes periodic
// (mathfunc.cu:165-206); f2 is scratch of n+1 values.
void periodic_fix(double* f, int n, double* f2) {
  for (int i = 0; i < n; i++) f2[i + 1] = f[i];
  f2[0] = f2[1] - (f2[2] - f2[1]);
  const double k2 = f2[n], k1 = f2[0];
  const double s2 = f2[n] - f2[n - 1], s1 = f2[1] - f2[0];
  const double xm = n;
  const double d = k1, c = s1 - s2;
  const double b = -2.0 * (s1 - s2) / xm - 3.0 * (k1 - k2) / xm / xm;
  const double a = -b / xm - (s1 - s2) / xm / xm - (k1 - k2) / xm / xm / xm;
  for (int i = 0; i < n + 1; i++) {
    const double x = i;
    const double cub = a * x * x * x + b * x * x + c * x + d;
    f2[i] = f2[i] - (cub - k2);
  }
  for (int i = 0; i < n; i++) f[i] = f2[i + 1];
}

void median3(double* f, int n, double* out) {
  for (int i = 0; i < n; i++) {
    const double a = (i == 0) ? f[n - 1] : f[i - 1];
    const double b = f[i];
    const double c = (i == n - 1) ? f[0] : f[i + 1];
    double hi = a > b ? a : b;
    hi = hi > c ? hi : c;
    double lo = a < b ? a : b;
    lo = lo < c ? lo : c;
    if (a < hi && a > lo) out[i] = a;       // strict tests: ties return the third value
    else if (b < hi && b > lo) out[i] = b;  // (mathfunc.cu:244-255)
    else out[i] = c;
  }
  for (int i = 0; i < n; i++) f[i] = out[i];
}

void gauss_weights(double* g, int nw) {
  const double sd = (nw - 1) / 6.0;
  const double a = 1.0 / (std::sqrt(2.0 * M_PI) * sd);
  const int m = (nw - 1) / 2;
  double sum = 0.0;
  for (int i = 0; i < nw; i++) {
    const double x = i - m;
    g[i] = a * std::exp(-0.5 * (x * x) / sd / sd);
    sum += g[i];
  }
  for (int i = 0; i < nw; i++) g[i] = g[i] / sum;
}

void gauss_smooth(double* f, const double* g, int n, int nw, double* out) {
  const int m = (nw - 1) / 2;
  for (int i = 0; i < n; i++) {
    double acc = 0.0;
    for (int j = 0; j < nw; j++) {
      int k = i + j - m;
      if (k < 0) k = n + k;
      else if (k > n - 1) k = k - n;
      acc += g[j] * f[k];
    }
    out[i] = acc;
  }
  for (int i = 0; i < n; i++) f[i] = out[i];
}

}  // namespace

namespace acum_host {
void gauss_weights13(double* g) { gauss_weights(g, 13); }
}  // namespace acum_host

extern "C" int acum_run_preprocess(acum_run* r) {
  if (!r) return fail(ACUM_ERR_ARG, "null run");
  if (r->preprocessed) return fail(ACUM_ERR_ARG, "run already preprocessed");
  const int n = (int)r->nTime;
  if (n < 4) return fail(ACUM_ERR_ARG, "nTime must be >= 4");
  const int nw = 13;
  double g[13];
  gauss_weights(g, nw);
  const double C = r->in.CFDscale, C2 = C * C;
  const double pscale = 0.5 * r->in.rhoRef * (r->in.Mref * r->in.Mref) * (r->in.a0 * r->in.a0);
  const double dt = r->in.dTau;
  const int periodic = r->in.periodic, lowpass = r->in.lowpass;
  for (int64_t s = 0; s < r->nSurf; s++) {
    const int64_t E = r->E[s];
#pragma omp parallel
    {
      std::vector<double> p(n), pd(n), q(n), qd(n), v(n), vd(n), tmp(n + 1);
#pragma omp for schedule(static)
      for (int64_t j = 0; j < E; j++) {
        float* Xe = r->X.data() + r->XPtr[s] + j * n * 4;
        float* Ne = r->N.data() + r->NPtr[s] + j * n * 6;
        float* Ve = r->V.data() + r->VPtr[s] + j * n * 6;
        float* Pe = r->PR.data() + r->PRPtr[s] + j * n * 2;
        for (int k = 0; k < n; k++) p[k] = (double)Pe[2 * k] * pscale;
        periodic_fix(p.data(), n, tmp.data());
        if (lowpass == 1) {
          for (int it = 0; it < 2; it++) {
            median3(p.data(), n, tmp.data());
            gauss_smooth(p.data(), g, n, nw, tmp.data());
          }
        }
        if (periodic == 1) central(p.data(), pd.data(), dt, n);
        else biased(p.data(), pd.data(), dt, n);
        for (int k = 0; k < n; k++) {
          Pe[2 * k] = (float)p[k];
          Pe[2 * k + 1] = (float)pd[k];
        }
        for (int d = 0; d < 3; d++) {
          for (int k = 0; k < n; k++) q[k] = (double)Ne[6 * k + d];
          central(q.data(), qd.data(), dt, n);
          for (int k = 0; k < n; k++) Ne[6 * k + 3 + d] = (float)qd[k];
          for (int k = 0; k < n; k++) q[k] = (double)Xe[4 * k + d] * C;
          central(q.data(), v.data(), dt, n);
          central(v.data(), vd.data(), dt, n);
          for (int k = 0; k < n; k++) {
            Xe[4 * k + d] = (float)q[k];
            Ve[6 * k + d] = (float)v[k];
            Ve[6 * k + 3 + d] = (float)vd[k];
          }
        }
        for (int k = 0; k < n; k++) Xe[4 * k + 3] = (float)((double)Xe[4 * k + 3] * C2);
      }
    }
  }
  r->preprocessed = true;
  return ACUM_OK;
}

extern "C" int acum_oaspl(int64_t nTime, int64_t nObs, const float* pA, double* oaspl) {
  if (nTime < 1 || nObs < 1 || !pA || !oaspl) return fail(ACUM_ERR_ARG, "bad argument");
  for (int64_t o = 0; o < nObs; o++) {
    double mean = 0.0;
    for (int64_t t = 0; t < nTime; t++) mean += pA[t * nObs + o];
    mean /= (double)nTime;
    double ms = 0.0;
    for (int64_t t = 0; t < nTime; t++) {
      const double d = pA[t * nObs + o] - mean;
      ms += d * d;
    }
    ms /= (double)nTime;
    oaspl[o] = 10.0 * std::log10(ms / (2e-5 * 2e-5));
  }
  return ACUM_OK;
}
