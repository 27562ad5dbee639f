// ORACLE -- test infrastructure only.  CPU restatement of the reference's host preprocessing:
// preprocess_imper (preprocess.cu:228-413) and the numerics helpers of mathfunc.cu:
// differentiate (132-140), differentiate_biased (142-159), make_periodic (165-206),
// medfilt1 (208-341), pre_gaussfilt1 (350-370), gaussfilt1 (375-421).
//
// The *_f32 functions repeat the reference's float arithmetic, including the places where a
// double literal promotes an expression to double before the result is stored to float, so a
// build of the reference's own mathfunc.cu must agree with them bit for bit.
// The *_f64 functions are the same steps in double; orc_preprocess_imper_f64 is the form the
// product uses (FP64 derivatives of the float inputs, results stored as float).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>
#include "../include/acum.h"

namespace {

// ---- float instance: same promotions as the reference --------------------------------------
void differentiate_f(const float* u, float* ud, float dt, int n) {
  ud[0] = (float)((double)(u[1] - u[n - 1]) / (2.0 * dt));
  ud[n - 1] = (float)((double)(u[0] - u[n - 2]) / (2.0 * dt));
  for (int i = 1; i < n - 1; i++) ud[i] = (float)((double)(u[i + 1] - u[i - 1]) / (2.0 * dt));
}

void differentiate_biased_f(const float* u, float* ud, float dt, int n) {
  ud[0] = (u[1] - u[0]) / dt;
  ud[n - 1] = (u[n - 1] - u[n - 2]) / dt;
  for (int i = 1; i < n - 1; i++) ud[i] = (float)((double)(u[i + 1] - u[i - 1]) / (2.0 * dt));
  ud[0] = 3 * ud[1] - 3 * ud[2] + ud[3];
  ud[n - 1] = 3 * ud[n - 2] - 3 * ud[n - 3] + ud[n - 4];
}

void make_periodic_f(float* fun, int n, int is) {
  std::vector<float> fun2(n + 1);
  for (int i = 0; i < n; i++) fun2[i + 1] = fun[i];
  fun2[0] = fun2[1] - (fun2[2] - fun2[1]);
  float k2 = fun2[n];
  float k1 = fun2[is];
  float s2 = fun2[n] - fun2[n - 1];
  float s1 = fun2[1] - fun2[0];
  int xmax = n;
  float dd = k1;
  float cc = s1 - s2;
  float bb = (float)(-2.0 * (double)(s1 - s2) / xmax - 3.0 * (double)(k1 - k2) / xmax / xmax);
  float aa = -bb / xmax - (s1 - s2) / xmax / xmax - (k1 - k2) / xmax / xmax / xmax;
  for (int i = 0; i < n + 1; i++) {
    float fc = aa * i * i * i + bb * i * i + cc * i + dd;
    fun2[i] = fun2[i] - (fc - k2);
  }
  for (int i = 0; i < n; i++) fun[i] = fun2[i + 1];
}

void medfilt1_f(float* f, int n) {
  std::vector<float> out(n);
  for (int i = 0; i < n; i++) {
    float a = (i == 0) ? f[n - 1] : f[i - 1];
    float b = f[i];
    float c = (i == n - 1) ? f[0] : f[i + 1];
    float fmax = a > b ? a : b;
    fmax = fmax > c ? fmax : c;
    float fmin = a < b ? a : b;
    fmin = fmin < c ? fmin : c;
    // strict tests: on ties the third value is returned (mathfunc.cu:244-255)
    if (a < fmax && a > fmin) out[i] = a;
    else if (b < fmax && b > fmin) out[i] = b;
    else out[i] = c;
  }
  for (int i = 0; i < n; i++) f[i] = out[i];
}

void pre_gaussfilt1_f(float* g, int nw) {
  float sd = (float)((nw - 1) / 6.0);
  float aa = (float)(1.0 / (std::sqrt(2.0 * M_PI) * sd));
  int m = (nw - 1) / 2;
  float sumg = 0.0f;
  for (int i = 0; i < nw; i++) {
    float xx = (float)(i - m);
    g[i] = (float)(aa * std::exp(-0.5 * (double)(xx * xx) / sd / sd));
    sumg += g[i];
  }
  for (int i = 0; i < nw; i++) g[i] = g[i] / sumg;
}

void gaussfilt1_f(float* f, const float* g, int n, int nw) {
  std::vector<float> out(n), sample(nw);
  int m = (nw - 1) / 2;
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < nw; j++) {
      int k = i + j - m;
      if (k < 0) k = n + k;
      else if (k > n - 1) k = k - n;
      sample[j] = f[k];
    }
    out[i] = 0.0f;
    for (int j = 0; j < nw; j++) out[i] += g[j] * sample[j];
  }
  for (int i = 0; i < n; i++) f[i] = out[i];
}

// ---- double instance ---------------------------------------------------------------------
void differentiate_d(const double* u, double* ud, double dt, int n) {
  ud[0] = (u[1] - u[n - 1]) / (2.0 * dt);
  ud[n - 1] = (u[0] - u[n - 2]) / (2.0 * dt);
  for (int i = 1; i < n - 1; i++) ud[i] = (u[i + 1] - u[i - 1]) / (2.0 * dt);
}

void differentiate_biased_d(const double* u, double* ud, double dt, int n) {
  ud[0] = (u[1] - u[0]) / dt;
  ud[n - 1] = (u[n - 1] - u[n - 2]) / dt;
  for (int i = 1; i < n - 1; i++) ud[i] = (u[i + 1] - u[i - 1]) / (2.0 * dt);
  ud[0] = 3 * ud[1] - 3 * ud[2] + ud[3];
  ud[n - 1] = 3 * ud[n - 2] - 3 * ud[n - 3] + ud[n - 4];
}

void make_periodic_d(double* fun, int n, int is) {
  std::vector<double> fun2(n + 1);
  for (int i = 0; i < n; i++) fun2[i + 1] = fun[i];
  fun2[0] = fun2[1] - (fun2[2] - fun2[1]);
  double k2 = fun2[n], k1 = fun2[is];
  double s2 = fun2[n] - fun2[n - 1], s1 = fun2[1] - fun2[0];
  double xmax = n;
  double dd = k1, cc = s1 - s2;
  double bb = -2.0 * (s1 - s2) / xmax - 3.0 * (k1 - k2) / xmax / xmax;
  double aa = -bb / xmax - (s1 - s2) / xmax / xmax - (k1 - k2) / xmax / xmax / xmax;
  for (int i = 0; i < n + 1; i++) {
    double x = i;
    double fc = aa * x * x * x + bb * x * x + cc * x + dd;
    fun2[i] = fun2[i] - (fc - k2);
  }
  for (int i = 0; i < n; i++) fun[i] = fun2[i + 1];
}

void medfilt1_d(double* f, int n) {
  std::vector<double> out(n);
  for (int i = 0; i < n; i++) {
    double a = (i == 0) ? f[n - 1] : f[i - 1];
    double b = f[i];
    double c = (i == n - 1) ? f[0] : f[i + 1];
    double fmax = a > b ? a : b;
    fmax = fmax > c ? fmax : c;
    double fmin = a < b ? a : b;
    fmin = fmin < c ? fmin : c;
    if (a < fmax && a > fmin) out[i] = a;
    else if (b < fmax && b > fmin) out[i] = b;
    else out[i] = c;
  }
  for (int i = 0; i < n; i++) f[i] = out[i];
}

void pre_gaussfilt1_d(double* g, int nw) {
  double sd = (nw - 1) / 6.0;
  double aa = 1.0 / (std::sqrt(2.0 * M_PI) * sd);
  int m = (nw - 1) / 2;
  double sumg = 0.0;
  for (int i = 0; i < nw; i++) {
    double xx = i - m;
    g[i] = aa * std::exp(-0.5 * (xx * xx) / sd / sd);
    sumg += g[i];
  }
  for (int i = 0; i < nw; i++) g[i] = g[i] / sumg;
}

void gaussfilt1_d(double* f, const double* g, int n, int nw) {
  std::vector<double> out(n);
  int m = (nw - 1) / 2;
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

extern "C" {

void orc_differentiate_f32(const float* u, float* ud, float dt, int n) { differentiate_f(u, ud, dt, n); }
void orc_differentiate_biased_f32(const float* u, float* ud, float dt, int n) { differentiate_biased_f(u, ud, dt, n); }
void orc_make_periodic_f32(float* f, int n, int is) { make_periodic_f(f, n, is); }
void orc_medfilt1_f32(float* f, int n) { medfilt1_f(f, n); }
void orc_pre_gaussfilt1_f32(float* g, int nw) { pre_gaussfilt1_f(g, nw); }
void orc_gaussfilt1_f32(float* f, const float* g, int n, int nw) { gaussfilt1_f(f, g, n, nw); }
void orc_differentiate_f64(const double* u, double* ud, double dt, int n) { differentiate_d(u, ud, dt, n); }
void orc_differentiate_biased_f64(const double* u, double* ud, double dt, int n) { differentiate_biased_d(u, ud, dt, n); }
void orc_make_periodic_f64(double* f, int n, int is) { make_periodic_d(f, n, is); }
void orc_medfilt1_f64(double* f, int n) { medfilt1_d(f, n); }
void orc_pre_gaussfilt1_f64(double* g, int nw) { pre_gaussfilt1_d(g, nw); }
void orc_gaussfilt1_f64(double* f, const double* g, int n, int nw) { gaussfilt1_d(f, g, n, nw); }

// preprocess_imper (preprocess.cu:228-413) on the reference AoS arrays, float arithmetic as in
// the reference.  X[(x,y,z,ds)], N[(n,nd)], V[(v,vd)], PR[(P,Pd)] per (surface, element, step);
// on entry X holds raw coordinates/areas, N raw normals and PR[0] the pressure coefficient.
void orc_preprocess_imper_f32(int nSurf, const int64_t* E, int nTime, float* X, float* N,
                              float* V, float* PR, const int64_t* XPtr, const int64_t* NPtr,
                              const int64_t* VPtr, const int64_t* PRPtr, float CFDscale,
                              float rhoRef, float Mref, float a0, float dTau, int periodic,
                              int lowpass) {
  const int n_win = 13;
  std::vector<float> gauss(n_win), v5(nTime), v5d(nTime), nx(nTime), ny(nTime), nz(nTime),
      ndx(nTime), ndy(nTime), ndz(nTime), x(nTime), y(nTime), z(nTime), vx(nTime), vy(nTime),
      vz(nTime), vdx(nTime), vdy(nTime), vdz(nTime);
  for (int s = 0; s < nSurf; s++) {
    for (int64_t j = 0; j < E[s] * nTime; j++) {  // preprocess.cu:269-281
      float* Xh = X + XPtr[s] + j * 4;
      float* Ph = PR + PRPtr[s] + j * 2;
      Xh[0] = Xh[0] * CFDscale;
      Xh[1] = Xh[1] * CFDscale;
      Xh[2] = Xh[2] * CFDscale;
      Xh[3] = Xh[3] * CFDscale * CFDscale;
      Ph[0] = (float)((double)Ph[0] * 0.5 * rhoRef * std::pow((double)Mref, 2) *
                      std::pow((double)a0, 2));
    }
    pre_gaussfilt1_f(gauss.data(), n_win);
    for (int64_t j = 0; j < E[s]; j++) {
      for (int k = 0; k < nTime; k++) {
        const float* Xh = X + XPtr[s] + (j * nTime + k) * 4;
        const float* Nh = N + NPtr[s] + (j * nTime + k) * 6;
        v5[k] = PR[PRPtr[s] + (j * nTime + k) * 2];
        nx[k] = Nh[0];
        ny[k] = Nh[1];
        nz[k] = Nh[2];
        x[k] = Xh[0];
        y[k] = Xh[1];
        z[k] = Xh[2];
      }
      make_periodic_f(v5.data(), nTime, 0);  // unconditional (preprocess.cu:321)
      if (lowpass == 1) {
        for (int it = 0; it < 2; it++) {
          medfilt1_f(v5.data(), nTime);
          gaussfilt1_f(v5.data(), gauss.data(), nTime, n_win);
        }
      }
      if (periodic == 1) differentiate_f(v5.data(), v5d.data(), dTau, nTime);
      else differentiate_biased_f(v5.data(), v5d.data(), dTau, nTime);
      differentiate_f(nx.data(), ndx.data(), dTau, nTime);
      differentiate_f(ny.data(), ndy.data(), dTau, nTime);
      differentiate_f(nz.data(), ndz.data(), dTau, nTime);
      differentiate_f(x.data(), vx.data(), dTau, nTime);
      differentiate_f(y.data(), vy.data(), dTau, nTime);
      differentiate_f(z.data(), vz.data(), dTau, nTime);
      differentiate_f(vx.data(), vdx.data(), dTau, nTime);
      differentiate_f(vy.data(), vdy.data(), dTau, nTime);
      differentiate_f(vz.data(), vdz.data(), dTau, nTime);
      for (int k = 0; k < nTime; k++) {
        float* Nh = N + NPtr[s] + (j * nTime + k) * 6;
        float* Vh = V + VPtr[s] + (j * nTime + k) * 6;
        float* Ph = PR + PRPtr[s] + (j * nTime + k) * 2;
        Ph[0] = v5[k];
        Ph[1] = v5d[k];
        Nh[3] = ndx[k];
        Nh[4] = ndy[k];
        Nh[5] = ndz[k];
        Vh[0] = vx[k];
        Vh[1] = vy[k];
        Vh[2] = vz[k];
        Vh[3] = vdx[k];
        Vh[4] = vdy[k];
        Vh[5] = vdz[k];
      }
    }
  }
}

// Same steps in FP64 from the float inputs; every derived series is rounded to float once.
void orc_preprocess_imper_f64(int nSurf, const int64_t* E, int nTime, float* X, float* N,
                              float* V, float* PR, const int64_t* XPtr, const int64_t* NPtr,
                              const int64_t* VPtr, const int64_t* PRPtr, double CFDscale,
                              double rhoRef, double Mref, double a0, double dTau, int periodic,
                              int lowpass) {
  const int n_win = 13;
  std::vector<double> gauss(n_win);
  pre_gaussfilt1_d(gauss.data(), n_win);
  const double pscale = 0.5 * rhoRef * (Mref * Mref) * (a0 * a0);
  const double c2 = CFDscale * CFDscale;
  for (int s = 0; s < nSurf; s++) {
    std::vector<double> p(nTime), pd(nTime), q(nTime), qd(nTime), v(nTime), vd(nTime);
    for (int64_t j = 0; j < E[s]; j++) {
      float* Xe = X + XPtr[s] + j * nTime * 4;
      float* Ne = N + NPtr[s] + j * nTime * 6;
      float* Ve = V + VPtr[s] + j * nTime * 6;
      float* Pe = PR + PRPtr[s] + j * nTime * 2;
      for (int k = 0; k < nTime; k++) p[k] = (double)Pe[2 * k] * pscale;
      make_periodic_d(p.data(), nTime, 0);
      if (lowpass == 1) {
        for (int it = 0; it < 2; it++) {
          medfilt1_d(p.data(), nTime);
          gaussfilt1_d(p.data(), gauss.data(), nTime, n_win);
        }
      }
      if (periodic == 1) differentiate_d(p.data(), pd.data(), dTau, nTime);
      else differentiate_biased_d(p.data(), pd.data(), dTau, nTime);
      for (int k = 0; k < nTime; k++) {
        Pe[2 * k] = (float)p[k];
        Pe[2 * k + 1] = (float)pd[k];
      }
      for (int d = 0; d < 3; d++) {
        for (int k = 0; k < nTime; k++) q[k] = (double)Ne[6 * k + d];
        differentiate_d(q.data(), qd.data(), dTau, nTime);
        for (int k = 0; k < nTime; k++) Ne[6 * k + 3 + d] = (float)qd[k];
        for (int k = 0; k < nTime; k++) q[k] = (double)Xe[4 * k + d] * CFDscale;
        differentiate_d(q.data(), v.data(), dTau, nTime);
        differentiate_d(v.data(), vd.data(), dTau, nTime);
        for (int k = 0; k < nTime; k++) {
          Xe[4 * k + d] = (float)q[k];
          Ve[6 * k + d] = (float)v[k];
          Ve[6 * k + 3 + d] = (float)vd[k];
        }
      }
      for (int k = 0; k < nTime; k++) Xe[4 * k + 3] = (float)((double)Xe[4 * k + 3] * c2);
    }
  }
}

}  // extern "C"
