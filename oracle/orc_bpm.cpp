// ORACLE -- test infrastructure only.  FP64 CPU restatement of the reference's broadband (BPM)
// path: findBBnoise (bb_core.cu:11-586), core (bb_core.cu:591-1043) and the shape functions
// a_min/a_max/b_min/b_max (bb_utils.cu:6-88), g5comp (bb_utils.cu:90-121) and interp_2d
// (bb_utils.cu:124-168).  Not bb_core_gpu.cu, whose arithmetic differs from the CPU model.
//
// Deviations from the reference, documented:
//   * the retarded-time quadratic uses aa = aMv (bb_core.cu:520 reads an uninitialised aa);
//     with vinf = 0 both give tau = |rp|/a0 and rp is unchanged;
//   * the cn/cc section accumulators of read_BB_inputs (io.cu:478-506) are not restated;
//   * 64-bit offsets.
// Accumulation follows the reference: per (observer, step, band, component) cell, the energy
// of every section is added in dB form, 10 log10(10^(a/10) + 10^(b/10)), surface by surface
// and section by section (bb_core.cu:850-859,905-906,967-968,1033-1034), then the time
// average and the overall level (bb_core.cu:565-583).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>
#include <omp.h>
#include "../include/acum.h"

namespace {

inline double p2(double x) { return x * x; }

double a_min(double a) {  // bb_utils.cu:6-25
  if (a < 0.204) return std::sqrt(67.552 - 886.788 * p2(a)) - 8.219;
  if (a > 0.244) return -142.795 * a * a * a + 103.656 * p2(a) - 57.757 * a + 6.006;
  return -32.665 * a + 3.981;
}
double a_max(double a) {  // bb_utils.cu:27-46
  if (a < 0.13) return std::sqrt(67.552 - 886.788 * p2(a)) - 8.219;
  if (a > 0.321) return -4.669 * a * a * a + 3.491 * p2(a) - 16.699 * a + 1.149;
  return -15.901 * a + 1.098;
}
double b_min(double b) {  // bb_utils.cu:48-67
  if (b < 0.13) return std::sqrt(16.888 - 886.788 * p2(b)) - 4.109;
  if (b > 0.145) return -817.810 * b * b * b + 355.21 * p2(b) - 135.024 * b + 10.619;
  return -83.607 * b + 8.138;
}
double b_max(double b) {  // bb_utils.cu:69-88
  if (b < 0.1) return std::sqrt(16.888 - 886.788 * p2(b)) - 4.109;
  if (b > 0.187) return -80.541 * b * b * b + 44.174 * p2(b) - 39.381 * b + 2.344;
  return -31.313 * b + 1.854;
}
double g5comp(double h, double eta) {  // bb_utils.cu:90-121
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
  double eta0 = -std::sqrt((m * m * std::pow(mu, 4)) / (6.25 + m * m * mu * mu));
  double k = 2.5 * std::sqrt(1.0 - p2(eta0 / mu)) - 2.5 - m * eta0;
  if (eta <= eta0) return m * eta + k;
  if (eta <= 0.0) return 2.5 * std::sqrt(1.0 - p2(eta / mu)) - 2.5;
  if (eta <= 0.03616) return std::sqrt(1.5625 - 1194.99 * p2(eta)) - 1.25;
  return -155.543 * eta + 4.375;
}

struct Table {
  const acum_bpm_case* c;
  int s;
  int64_t idx(int is, int im, int ir, int ia, int id) const {
    const int64_t nm = c->nmach_t[s], nr = c->nre_t[s], na = c->naoa_t[s];
    return c->dPtr[s] + is * nm * nr * na * 7 + im * nr * na * 7 + ir * na * 7 + ia * 7 + id;
  }
  // interp_2d (bb_utils.cu:124-168): trilinear over (section, Mach, Re) at one AoA index.
  double interp(double rR, int is1, int is2, double mach, int im1, int im2, double re, int ir1,
                int ir2, int id, int ia) const {
    const float* st = c->sect_t + c->sPtr[s];
    const float* mt = c->mach_t + c->mPtr[s];
    const float* rt = c->re_t + c->rPtr[s];
    const float* d = c->bbdata_t;
    double ws = (is1 == is2) ? 0.5 : (rR - st[is1]) / ((double)st[is2] - st[is1]);
    double wm = (im1 == im2) ? 0.5 : (mach - mt[im1]) / ((double)mt[im2] - mt[im1]);
    double wr = (ir1 == ir2) ? 0.5 : (re - rt[ir1]) / ((double)rt[ir2] - rt[ir1]);
    double F1 = (1 - ws) * d[idx(is1, im1, ir1, ia, id)] + ws * d[idx(is2, im1, ir1, ia, id)];
    double F2 = (1 - ws) * d[idx(is1, im2, ir1, ia, id)] + ws * d[idx(is2, im2, ir1, ia, id)];
    double F3 = (1 - ws) * d[idx(is1, im1, ir2, ia, id)] + ws * d[idx(is2, im1, ir2, ia, id)];
    double F4 = (1 - ws) * d[idx(is1, im2, ir2, ia, id)] + ws * d[idx(is2, im2, ir2, ia, id)];
    return (1 - wm) * (1 - wr) * F1 + wm * (1 - wr) * F2 + (1 - wm) * wr * F3 + wm * wr * F4;
  }
  // bracket search of bb_core.cu:208-234
  void bracket(double mach, double re, double rR, int* im1, int* im2, int* ir1, int* ir2,
               int* is1, int* is2) const {
    const int nm = c->nmach_t[s], nr = c->nre_t[s], ns = c->nsect_t[s];
    const float* st = c->sect_t + c->sPtr[s];
    const float* mt = c->mach_t + c->mPtr[s];
    const float* rt = c->re_t + c->rPtr[s];
    int a = 0;
    for (int j = 0; j < nm; j++) {
      if (mach < mt[j]) break;
      a++;
    }
    int b = 0;
    for (int j = 0; j < nr; j++) {
      if (re < rt[j]) break;
      b++;
    }
    int e = 0;
    for (int j = 0; j < ns; j++) {
      if (rR < st[j]) break;
      e++;
    }
    *im1 = a - 1 > 0 ? a - 1 : 0;
    *ir1 = b - 1 > 0 ? b - 1 : 0;
    *is1 = e - 1 > 0 ? e - 1 : 0;
    *im2 = a < nm - 1 ? a : nm - 1;
    *ir2 = b < nr - 1 ? b : nr - 1;
    *is2 = e < ns - 1 ? e : ns - 1;
  }
};

// Per (surface, section, step) state computed before the observer loop (bb_core.cu:60-440).
struct SectState {
  double xyz[3], xyzt[3];
  double u, alpha, cl_alpha;
  double bl[4];
};

inline double db_add(double a, double b) {  // 10 log10(10^(a/10) + 10^(b/10))
  return 10.0 * std::log10(std::pow(10.0, a / 10.0) + std::pow(10.0, b / 10.0));
}

// core(), bb_core.cu:591-1043, accumulating into out[band*7 + comp] (dB).
void core(double rpnorm, double theta, double phi, double c, double lsect, double alpha,
          double u, double nu, double a0, int trip, double del_s, double del_p, double dis_s,
          double dis_p, int isect, int nsect, const float* ff, int nf2, double R, double* out) {
  if (del_s < 0.0) del_s = 1.0e-10;
  if (del_p < 0.0) del_p = 1.0e-10;
  if (dis_p < 0.0) dis_p = 1.0e-10;
  if (dis_s < 0.0) dis_s = 1.0e-10;
  u = std::fabs(u);
  const double mach = u / a0;
  const double r_c = u * c / nu;
  const double m_c = 0.6 * u / a0;
  if (theta == 0) theta = 1e-10;
  if (phi == 0) phi = 1e-10;
  const double d_h = (2.0 * p2(std::sin(theta / 2.0)) * p2(std::sin(phi))) /
                     ((1.0 + mach * std::cos(theta)) * p2(1.0 + (mach - m_c) * std::cos(theta)));
  const double d_l = (p2(std::sin(theta)) * p2(std::sin(phi))) / std::pow(1.0 + mach * std::cos(theta), 4);

  std::vector<double> st_p(nf2), st_s(nf2), as(nf2), ap(nf2), anew(nf2);
  const double r_deltas_p = u * dis_p / nu;
  for (int i = 0; i < nf2; i++) {
    st_p[i] = ff[i] * dis_p / u;
    st_s[i] = ff[i] * dis_s / u;
  }
  const double st_1 = 0.02 * std::pow(mach, -0.6);
  double st_2;
  if (alpha < 1.33) st_2 = st_1;
  else if (alpha > 12.5) st_2 = st_1 * 4.72;
  else st_2 = st_1 * std::pow(10.0, 0.0054 * p2(alpha - 1.33));
  for (int i = 0; i < nf2; i++) {
    as[i] = std::fabs(std::log10(st_s[i] / st_1));
    ap[i] = std::fabs(std::log10(st_p[i] / st_1));
    anew[i] = std::fabs(std::log10(st_s[i] / st_2));
  }
  auto a0_of = [](double rc) {
    if (rc < 9.52e4) return 0.57;
    if (rc > 8.57e5) return 1.13;
    return -9.57e-13 * p2(rc - 8.57e5) + 1.13;
  };
  auto b0_of = [](double rc) {
    if (rc < 9.52e4) return 0.3;
    if (rc > 8.57e5) return 0.56;
    return -4.48e-13 * p2(rc - 8.57e5) + 0.56;
  };
  auto k1_of = [](double rc) {
    if (rc < 2.47e5) return -4.31 * std::log10(rc) + 156.3;
    if (rc > 8.0e5) return 128.5;
    return -9.0 * std::log10(rc) + 181.6;
  };
  double a_0 = a0_of(r_c);
  double a_r = (-20.0 - a_min(a_0)) / (a_max(a_0) - a_min(a_0));
  std::vector<double> aas(nf2), aap(nf2), bb(nf2);
  for (int i = 0; i < nf2; i++) {
    aas[i] = a_min(as[i]) + a_r * (a_max(as[i]) - a_min(as[i]));
    aap[i] = a_min(ap[i]) + a_r * (a_max(ap[i]) - a_min(ap[i]));
  }
  const double b_0 = b0_of(r_c);
  const double b_r = (-20.0 - b_min(b_0)) / (b_max(b_0) - b_min(b_0));
  for (int i = 0; i < nf2; i++) {
    const double b = std::fabs(std::log10(st_s[i] / st_2));
    bb[i] = b_min(b) + b_r * (b_max(b) - b_min(b));
  }
  double k_1 = k1_of(r_c);
  const double deltak_1 = (r_deltas_p <= 5000) ? alpha * (1.43 * std::log10(r_deltas_p) - 5.29) : 0.0;
  const double gamm = 27.094 * mach + 3.31;
  const double gamm_0 = 23.43 * mach + 4.651;
  const double beta = 72.65 * mach + 10.74;
  const double beta_0 = -34.19 * mach - 13.82;
  auto k2_of = [&](double k1) {
    if (alpha < gamm_0 - gamm) return k1 - 1000.0;
    if (alpha > gamm_0 + gamm) return k1 - 12.0;
    return k1 + std::sqrt(p2(beta) - p2(beta / gamm) * p2(alpha - gamm_0)) + beta_0;
  };
  double k_2 = k2_of(k_1);
  const double m5 = std::pow(mach, 5);
  std::vector<double> spl_p(nf2), spl_s(nf2), spl_a(nf2);
  for (int i = 0; i < nf2; i++) {
    spl_p[i] = 10.0 * std::log10((dis_p * m5 * lsect * d_h) / p2(rpnorm)) + aap[i] + (k_1 - 3.0) + deltak_1;
    spl_s[i] = 10.0 * std::log10((dis_s * m5 * lsect * d_h) / p2(rpnorm)) + aas[i] + (k_1 - 3.0);
  }
  if (alpha >= gamm_0 || alpha > 12.5) {  // bb_core.cu:781-826
    const double r_c_new = 3.0 * r_c;
    a_0 = a0_of(r_c_new);
    k_1 = k1_of(r_c_new);
    k_2 = k2_of(k_1);
    a_r = (-20.0 - a_min(a_0)) / (a_max(a_0) - a_min(a_0));
    for (int i = 0; i < nf2; i++) {
      const double an2 = a_min(anew[i]) + a_r * (a_max(anew[i]) - a_min(anew[i]));
      spl_a[i] = 10.0 * std::log10((dis_s * m5 * lsect * d_l) / p2(rpnorm)) + an2 + k_2;
    }
  } else {
    for (int i = 0; i < nf2; i++)
      spl_a[i] = 10.0 * std::log10((dis_s * m5 * lsect * d_h) / p2(rpnorm)) + bb[i] + k_2;
  }
  for (int i = 0; i < nf2; i++) {
    if (spl_a[i] < -100) spl_a[i] = -100.0;
    if (spl_s[i] < -100) spl_s[i] = -100.0;
    if (spl_p[i] < -100) spl_p[i] = -100.0;
  }
  for (int i = 0; i < nf2; i++) {
    double* o = out + i * 7;
    o[1] = db_add(o[1], spl_a[i]);
    o[2] = db_add(o[2], spl_s[i]);
    o[3] = db_add(o[3], spl_p[i]);
    o[0] = 10.0 * std::log10(std::pow(10.0, o[0] / 10.0) + std::pow(10.0, spl_a[i] / 10.0) +
                             std::pow(10.0, spl_p[i] / 10.0) + std::pow(10.0, spl_s[i] / 10.0));
  }
  if (trip == 0) {  // LBL-VS, bb_core.cu:865-915
    double stlprim;
    if (r_c < 1.3e5) stlprim = 0.18;
    else if (r_c > 1.3e5 && r_c <= 4e5) stlprim = 0.001756 * std::pow(r_c, 0.3931);
    else stlprim = 0.28;
    const double stpkprm = std::pow(10.0, -0.04 * alpha) * stlprim;
    const double r_c0 = (alpha <= 3.0) ? std::pow(10.0, 0.215 * alpha + 4.978)
                                       : std::pow(10.0, 0.120 * alpha + 5.263);
    const double d = r_c / r_c0;
    double g2;
    if (d <= 0.3237) g2 = 77.852 * std::log10(d) + 15.328;
    else if (d <= 0.5689) g2 = 65.188 * std::log10(d) + 9.125;
    else if (d <= 1.7579) g2 = -114.052 * p2(std::log10(d));
    else if (d <= 3.0889) g2 = -65.188 * std::log10(d) + 9.125;
    else g2 = -77.852 * std::log10(d) + 15.328;
    const double g3 = 171.04 - 3.03 * alpha;
    const double scal = 10.0 * std::log10(del_p * m5 * d_h * lsect / p2(rpnorm));
    for (int i = 0; i < nf2; i++) {
      const double e = (ff[i] * del_p / u) / stpkprm;
      const double le = std::log10(e);
      double g1;
      if (e <= 0.5974) g1 = 39.8 * le - 11.12;
      else if (e <= 0.8545) g1 = 98.409 * le + 2.0;
      else if (e <= 1.17) g1 = -5.076 + std::sqrt(2.484 - 506.25 * p2(le));
      else if (e <= 1.674) g1 = -98.409 * le + 2.0;
      else g1 = -39.80 * le - 11.12;
      const double spl_lam = g1 + g2 + g3 + scal;
      out[i * 7 + 4] = db_add(out[i * 7 + 4], spl_lam);
      out[i * 7] = db_add(out[i * 7], spl_lam);
    }
  }
  {  // TEB-VS, bb_core.cu:917-971
    const double h = 0.0005, psi = 14.0;
    const double dstravg = (dis_s + dis_p) / 2.0;
    const double hdstar = h / dstravg;
    const double dstarh = 1.0 / hdstar;
    const double aterm = 0.212 - 0.0045 * psi;
    const double stpeak = (hdstar >= 0.2) ? aterm / (1.0 + 0.235 * dstarh - 0.0132 * p2(dstarh))
                                          : 0.1 * hdstar + 0.095 - 0.00243 * psi;
    const double g4 = (hdstar <= 5.0) ? 17.5 * std::log10(hdstar) + 157.5 - 1.114 * psi
                                      : 169.7 - 1.114 * psi;
    const double hdstarp = 6.724 * p2(hdstar) - 4.019 * hdstar + 1.107;
    const double scal = 10.0 * std::log10(std::pow(mach, 5.5) * h * d_h * lsect / p2(rpnorm));
    for (int i = 0; i < nf2; i++) {
      const double eta = std::log10((ff[i] * h / u) / stpeak);
      const double g514 = g5comp(hdstar, eta);
      const double g50 = g5comp(hdstarp, eta);
      double g5 = g50 + 0.0714 * psi * (g514 - g50);
      if (g5 > 0) g5 = 0.0;
      const double f4 = g5comp(0.25, eta);
      if (g5 > f4) g5 = f4;
      const double spl_blunt = g4 + g5 + scal;
      out[i * 7 + 5] = db_add(out[i * 7 + 5], spl_blunt);
      out[i * 7] = db_add(out[i * 7], spl_blunt);
    }
  }
  if (isect == nsect - 1) {  // TVF, bb_core.cu:975-1039
    const double AR = R / c;
    static const double AR0[6] = {2.0, 2.67, 4.0, 6.0, 12.0, 24.0};
    static const double al0[6] = {0.54, 0.62, 0.71, 0.79, 0.89, 0.95};
    double alprat;
    if (AR < AR0[0]) alprat = al0[0];
    else if (AR >= AR0[5]) alprat = al0[5];  // reference reads AR0[6] at AR == 24 exactly
    else {
      int kk = 0;
      for (int i = 0; i < 6; i++) {
        if (AR < AR0[i]) break;
        kk++;
      }
      alprat = al0[kk - 1] * (AR0[kk] - AR) / (AR0[kk] - AR0[kk - 1]) +
               al0[kk] * (AR - AR0[kk - 1]) / (AR0[kk] - AR0[kk - 1]);
    }
    const double alptipp = alpha * alprat;
    const double tip_l = (std::fabs(alptipp) <= 2.0) ? (0.023 + 0.0169 * alptipp) * c
                                                      : (0.0378 + 0.0095 * alptipp) * c;
    const double mm = (1.0 + 0.036 * alptipp) * mach;
    const double um = mm * a0;
    double term = mach * mach * std::pow(mm, 3) * p2(tip_l) * d_h / p2(rpnorm);
    term = (term > 2.2204460492503131e-16) ? term : 2.2204460492503131e-16;  // MAX macro
    const double scal = 10.0 * std::log10(term);
    for (int i = 0; i < nf2; i++) {
      const double stpp = ff[i] * tip_l / um;
      const double spl_tip = 126.0 - 30.5 * p2(std::log10(stpp) + 0.3) + scal;
      out[i * 7 + 6] = db_add(out[i * 7 + 6], spl_tip);
      out[i * 7] = db_add(out[i * 7], spl_tip);
    }
  }
}

// Section kinematics, table lookups, effective angle of attack and boundary-layer thickness of
// one surface (bb_core.cu:60-440).  Returns false on a table without the needed brackets.
bool prepare_surface(const acum_bpm_case* c, int s, std::vector<SectState>& st) {
  const int ns = c->nsect[s], nT = c->nTime;
  const double span = c->span[s], omega = c->omega[s];
  const int ccw = c->ccw[s];
  const float* xyzc = c->xyzcR + c->xsPtr[s];
  const float* bbd = c->bbdata + c->dsPtr[s];
  const double nu = 0.0000181206 / 1.22500;
  const double thx = c->rotate[3 * s] * M_PI / 180, thy = c->rotate[3 * s + 1] * M_PI / 180,
               thz = c->rotate[3 * s + 2] * M_PI / 180;
  const double tr[3] = {c->translate[3 * s], c->translate[3 * s + 1], c->translate[3 * s + 2]};
  const int iq = (int)std::lround(ns * 0.75);  // bb_core.cu:88 (round half away from zero)
  if (iq >= ns) return false;
  st.assign((size_t)ns * nT, SectState());
  Table tab{c, s};
  const int na = c->naoa_t[s];
  const float* aoa = c->aoa_t + c->aPtr[s];
  std::vector<double> cl(na);
  for (int is = 0; is < ns; is++) {
    const double ch = xyzc[4 * is + 3] * span;
    for (int it = 0; it < nT; it++) {
      double x0[3], x1[3], xt0[3], xt1[3];
      x0[0] = xyzc[4 * is] * span;
      x0[1] = xyzc[4 * iq + 1] * span;
      x0[2] = xyzc[4 * iq + 2] * span;
      if (ccw == 0) x0[1] = -x0[1];
      double psi = c->dTau * it * omega + c->psi_offset[s] * M_PI / 180;
      if (ccw == 0) psi = -psi;
      x1[2] = x0[2];
      x1[1] = x0[0] * std::sin(psi) + x0[1] * std::cos(psi);
      x1[0] = x0[0] * std::cos(psi) - x0[1] * std::sin(psi);
      xt0[0] = x0[0];
      const double theta = bbd[it * ns * 7 + is * 7 + 6] * M_PI / 180;
      xt0[2] = x0[2] - ch * 0.75 * std::sin(theta);
      xt0[1] = (ccw == 1) ? x0[1] - ch * 0.75 * std::cos(theta) : x0[1] + ch * 0.75 * std::cos(theta);
      xt1[2] = xt0[2];
      xt1[1] = xt0[0] * std::sin(psi) + xt0[1] * std::cos(psi);
      xt1[0] = xt0[0] * std::cos(psi) - xt0[1] * std::sin(psi);
      // about x (bb_core.cu:123-129)
      xt0[0] = xt1[0];
      xt0[2] = xt1[1] * std::sin(thx) + xt1[2] * std::cos(thx);
      xt0[1] = xt1[1] * std::cos(thx) - xt1[2] * std::sin(thx);
      x0[0] = x1[0];
      x0[2] = x1[1] * std::sin(thx) + x1[2] * std::cos(thx);
      x0[1] = x1[1] * std::cos(thx) - x1[2] * std::sin(thx);
      // about y (bb_core.cu:133-139)
      xt1[1] = xt0[1];
      xt1[2] = -xt0[0] * std::sin(thy) + xt0[2] * std::cos(thy);
      xt1[0] = xt0[0] * std::cos(thy) + xt0[2] * std::sin(thy);
      x1[1] = x0[1];
      x1[2] = -x0[0] * std::sin(thy) + x0[2] * std::cos(thy);
      x1[0] = x0[0] * std::cos(thy) + x0[2] * std::sin(thy);
      // about z (bb_core.cu:145-151)
      xt0[2] = xt1[2];
      xt0[1] = xt1[0] * std::sin(thz) + xt1[1] * std::cos(thz);
      xt0[0] = xt1[0] * std::cos(thz) - xt1[1] * std::sin(thz);
      x0[2] = x1[2];
      x0[1] = x1[0] * std::sin(thz) + x1[1] * std::cos(thz);
      x0[0] = x1[0] * std::cos(thz) - x1[1] * std::sin(thz);
      SectState& q = st[(size_t)is * nT + it];
      for (int i = 0; i < 3; i++) {
        q.xyz[i] = x0[i] + tr[i];
        q.xyzt[i] = xt0[i] + tr[i];
      }
    }
    for (int it = 0; it < nT; it++) {
      SectState& q = st[(size_t)is * nT + it];
      const SectState& qm = st[(size_t)is * nT + (it == 0 ? nT - 1 : it - 1)];
      const SectState& qp = st[(size_t)is * nT + (it == nT - 1 ? 0 : it + 1)];
      double v[3], v1[3];
      for (int i = 0; i < 3; i++) v[i] = c->vInf[i] - (qp.xyz[i] - qm.xyz[i]) / (2 * c->dTau);
      for (int i = 0; i < 3; i++) v1[i] = q.xyzt[i] - q.xyz[i];
      const double l1 = std::sqrt(p2(v1[0]) + p2(v1[1]) + p2(v1[2]));
      for (int i = 0; i < 3; i++) v1[i] /= l1;
      q.u = v[0] * v1[0] + v[1] * v1[1] + v[2] * v1[2];
      const double mach = std::fabs(q.u / c->a0);
      const double re = std::fabs(q.u * ch / nu);
      int im1, im2, ir1, ir2, is1, is2;
      tab.bracket(mach, re, xyzc[4 * is], &im1, &im2, &ir1, &ir2, &is1, &is2);
      for (int i = 0; i < na; i++)
        cl[i] = tab.interp(xyzc[4 * is], is1, is2, mach, im1, im2, re, ir1, ir2, 0, i);
      int a1 = -1, a2 = -1, a3 = -1;
      for (int i = 0; i < na; i++) {  // bb_core.cu:254-275
        if (aoa[i] >= 0.0 && a1 < 0) a1 = i;
        if (aoa[i] >= 6.0 && a2 < 0) a2 = i;
        if (cl[i] > 0 && a3 < 0) a3 = i;
        if (a1 >= 0 && a2 >= 0 && a3 >= 0) break;
      }
      if (a1 < 1 || a2 < 0 || a3 < 1) return false;
      q.cl_alpha = (cl[a2] - cl[a1]) / ((double)aoa[a2] - aoa[a1]);
      double wa1 = (0 - aoa[a1 - 1]) / ((double)aoa[a1] - aoa[a1 - 1]);
      const double cl_alpha0 = cl[a1 - 1] * (1 - wa1) + cl[a1] * wa1;
      wa1 = (0 - cl[a3 - 1]) / (cl[a3] - cl[a3 - 1]);
      const double alpha_cl0 = aoa[a3 - 1] * (1 - wa1) + aoa[a3] * wa1;
      const double cn = bbd[it * ns * 7 + is * 7 + 4], ct = bbd[it * ns * 7 + is * 7 + 5];
      double alpha_eff = 0.0, err = 1.0;
      int iter = 0;
      while (err > 1e-3) {  // bb_core.cu:302-321
        const double clsec = cn * std::cos(alpha_eff * M_PI / 180) - ct * std::sin(alpha_eff * M_PI / 180);
        const double ae1 = (clsec - cl_alpha0) / q.cl_alpha;
        err = std::fabs(ae1 - alpha_eff);
        alpha_eff = ae1;
        iter++;
        if (iter > 50) break;
      }
      q.alpha = std::fabs(alpha_eff - alpha_cl0);
      if (c->t_del == 0) {
        for (int i = 0; i < 4; i++) q.bl[i] = bbd[it * ns * 7 + is * 7 + i];
      } else if (c->t_del == 1) {  // bb_core.cu:357-413
        for (int i = 0; i < na; i++) {
          if (aoa[i] >= alpha_eff) {
            a1 = i;
            break;
          }
        }
        if (a1 < 1) return false;
        wa1 = (alpha_eff - aoa[a1 - 1]) / ((double)aoa[a1] - aoa[a1 - 1]);
        double blt[8];
        for (int jj = 0; jj < 2; jj++)
          for (int i = 0; i < 4; i++)
            blt[jj * 4 + i] = tab.interp(xyzc[4 * is], is1, is2, mach, im1, im2, re, ir1, ir2, i + 3, jj + a1 - 1);
        for (int i = 0; i < 4; i++) q.bl[i] = ch * (blt[i] * (1 - wa1) + blt[4 + i] * wa1);
      } else {  // empirical, bb_core.cu:414-438
        const double lr = std::log10(re);
        const double delta_0 = ch * std::pow(10.0, 1.6569 - 0.9045 * lr + 0.0596 * lr * lr);
        const double deltas_0 = ch * std::pow(10.0, 3.0187 - 1.5397 * lr + 0.1059 * lr * lr);
        const double al = q.alpha;
        q.bl[1] = delta_0 * std::pow(10.0, -0.04175 * al + 0.00106 * al * al);
        q.bl[3] = deltas_0 * std::pow(10.0, -0.0432 * al + 0.00113 * al * al);
        if (al <= 7.5) q.bl[0] = delta_0 * std::pow(10.0, 0.03114 * al);
        else if (al > 12.5) q.bl[0] = delta_0 * 12.0 * std::pow(10.0, 0.0258 * al);
        else q.bl[0] = delta_0 * 0.0303 * std::pow(10.0, 0.2336 * al);
        if (al <= 7.5) q.bl[2] = deltas_0 * std::pow(10.0, 0.0679 * al);
        else if (al > 12.5) q.bl[2] = deltas_0 * 52.42 * std::pow(10.0, 0.0258 * al);
        else q.bl[2] = deltas_0 * 0.0162 * std::pow(10.0, 0.3066 * al);
      }
    }
  }
  return true;
}

}  // namespace

extern "C" {

double orc_a_min(double a) { return a_min(a); }
double orc_a_max(double a) { return a_max(a); }
double orc_b_min(double b) { return b_min(b); }
double orc_b_max(double b) { return b_max(b); }
double orc_g5comp(double h, double eta) { return g5comp(h, eta); }

// Broadband levels of observers [obs0, obs0+nobs) (all when nobs <= 0).
// BBspl[(o-obs0)*nf2*7 + band*7 + comp], BBoaspl[o-obs0], dB.
int orc_bpm_f64(const acum_bpm_case* c, int64_t obs0, int64_t nobs, double* BBspl,
                double* BBoaspl, int nthreads) {
  if (!c || c->nSurf_BB < 1 || c->nTime < 3 || c->nf2 < 1) return ACUM_ERR_ARG;
  if (nobs <= 0) {
    obs0 = 0;
    nobs = c->nObs;
  }
  const int nS = c->nSurf_BB, nT = c->nTime, nf2 = c->nf2;
  std::vector<std::vector<SectState>> pre(nS);
  for (int s = 0; s < nS; s++)
    if (!prepare_surface(c, s, pre[s])) return ACUM_ERR_ARG;
  const double nu = 0.0000181206 / 1.22500;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
    std::vector<double> out((size_t)nT * nf2 * 7);
#pragma omp for schedule(dynamic, 1)
    for (int64_t oi = 0; oi < nobs; oi++) {
      const int64_t o = obs0 + oi;
      std::fill(out.begin(), out.end(), -100.0);
      const double ro[3] = {c->rObs[3 * o], c->rObs[3 * o + 1], c->rObs[3 * o + 2]};
      for (int s = 0; s < nS; s++) {
        const int ns = c->nsect[s];
        const float* xyzc = c->xyzcR + c->xsPtr[s];
        const double span = c->span[s];
        for (int is = 0; is < ns; is++) {
          double lsect;  // bb_core.cu:457-459
          if (is == 0) lsect = ((double)xyzc[4 * (is + 1)] - xyzc[4 * is]) / 2.0 * span;
          else if (is == ns - 1) lsect = ((double)xyzc[4 * is] - xyzc[4 * (is - 1)]) / 2.0 * span;
          else lsect = ((double)xyzc[4 * (is + 1)] - xyzc[4 * (is - 1)]) / 2.0 * span;
          const double ch = xyzc[4 * is + 3] * span;
          const int km = (is == ns - 1) ? is - 1 : is;
          const int kp = (is == ns - 1) ? is : is + 1;
          for (int it = 0; it < nT; it++) {
            const SectState& q = pre[s][(size_t)is * nT + it];
            const SectState& qm = pre[s][(size_t)km * nT + it];
            const SectState& qp = pre[s][(size_t)kp * nT + it];
            double tr[3][3];
            double tmp = 0;
            for (int i = 0; i < 3; i++) {
              tr[0][i] = q.xyzt[i] - q.xyz[i];
              tmp += p2(tr[0][i]);
            }
            for (int i = 0; i < 3; i++) tr[0][i] /= std::sqrt(tmp);
            tmp = 0;
            for (int i = 0; i < 3; i++) {
              tr[1][i] = qp.xyz[i] - qm.xyz[i];
              tmp += p2(tr[1][i]);
            }
            for (int i = 0; i < 3; i++) tr[1][i] /= std::sqrt(tmp);
            if (c->ccw[s] == 1) {
              tr[2][0] = tr[0][1] * tr[1][2] - tr[0][2] * tr[1][1];
              tr[2][1] = tr[0][2] * tr[1][0] - tr[0][0] * tr[1][2];
              tr[2][2] = tr[0][0] * tr[1][1] - tr[0][1] * tr[1][0];
              tmp = std::sqrt(p2(tr[2][0]) + p2(tr[2][1]) + p2(tr[2][2]));
              for (int i = 0; i < 3; i++) tr[2][i] /= tmp;
            } else {
              tr[2][0] = tr[1][1] * tr[0][2] - tr[1][2] * tr[0][1];
              tr[2][1] = tr[1][2] * tr[0][0] - tr[1][0] * tr[0][2];
              tr[2][2] = tr[1][0] * tr[0][1] - tr[1][1] * tr[0][0];
              tmp = std::sqrt(p2(tr[2][0]) + p2(tr[2][1]) + p2(tr[2][2]));
              for (int i = 0; i < 3; i++) tr[2][i] /= std::sqrt(tmp);  // bb_core.cu:512
            }
            double rp[3];
            for (int i = 0; i < 3; i++) rp[i] = ro[i] - q.xyzt[i];
            const double aMv = p2(c->a0) - p2(c->vInf[0]) - p2(c->vInf[1]) - p2(c->vInf[2]);
            const double bq = rp[0] * c->vInf[0] + rp[1] * c->vInf[1] + rp[2] * c->vInf[2];
            const double cq = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2];
            const double tau = (-bq + std::sqrt(p2(bq) + aMv * cq)) / aMv;
            for (int i = 0; i < 3; i++) rp[i] = ro[i] - q.xyzt[i] - c->vInf[i] * tau;
            const double rpn = std::sqrt(p2(rp[0]) + p2(rp[1]) + p2(rp[2]));
            double rp2[3];
            for (int i = 0; i < 3; i++) rp2[i] = rp[0] * tr[i][0] + rp[1] * tr[i][1] + rp[2] * tr[i][2];
            const double r_e = std::sqrt(p2(rp2[1]) + p2(rp2[2]));
            const double theta = std::acos(std::fabs(rp2[0] / rpn));
            const double phi = std::acos(std::fabs(rp2[1]) / r_e);
            core(rpn, theta, phi, ch, lsect, std::fabs(q.alpha), q.u, nu, c->a0, c->trip[s],
                 q.bl[0], q.bl[1], q.bl[2], q.bl[3], is, ns, c->ff, nf2, span,
                 out.data() + (size_t)it * nf2 * 7);
          }
        }
      }
      double oa = -100.0;  // bb_core.cu:565-583
      for (int i = 0; i < nf2; i++) {
        for (int j = 0; j < 7; j++) {
          double v = -100.0;
          for (int it = 0; it < nT; it++) v = db_add(v, out[(size_t)it * nf2 * 7 + i * 7 + j]);
          v = 10.0 * std::log10(std::pow(10.0, v / 10.0) / (double)nT);
          BBspl[oi * nf2 * 7 + i * 7 + j] = v;
        }
        oa = db_add(oa, BBspl[oi * nf2 * 7 + i * 7]);
      }
      BBoaspl[oi] = oa;
    }
  }
  return ACUM_OK;
}

}  // extern "C"

extern "C" {
// One call of core() (bb_core.cu:591) adding into out[nf2*7] (dB), for unit tests.
void orc_bpm_core(double rpnorm, double theta, double phi, double c, double lsect, double alpha,
                  double u, double a0, int trip, double del_s, double del_p, double dis_s,
                  double dis_p, int isect, int nsect, const float* ff, int nf2, double R,
                  double* out) {
  core(rpnorm, theta, phi, c, lsect, alpha, u, 0.0000181206 / 1.22500, a0, trip, del_s, del_p,
       dis_s, dis_p, isect, nsect, ff, nf2, R, out);
}
}

extern "C" {
// Arguments of every core() call for observer `o`, in the reference's order (surface, section,
// step): rpnorm, theta, phi, c, lsect, alpha, u, del_s, del_p, dis_s, dis_p, isect, nsect, span.
int orc_bpm_core_args(const acum_bpm_case* c, int64_t o, double* out) {
  const int nS = c->nSurf_BB, nT = c->nTime;
  int64_t k = 0;
  for (int s = 0; s < nS; s++) {
    std::vector<SectState> st;
    if (!prepare_surface(c, s, st)) return ACUM_ERR_ARG;
    const int ns = c->nsect[s];
    const float* xyzc = c->xyzcR + c->xsPtr[s];
    const double span = c->span[s];
    for (int is = 0; is < ns; is++) {
      double lsect;
      if (is == 0) lsect = ((double)xyzc[4 * (is + 1)] - xyzc[4 * is]) / 2.0 * span;
      else if (is == ns - 1) lsect = ((double)xyzc[4 * is] - xyzc[4 * (is - 1)]) / 2.0 * span;
      else lsect = ((double)xyzc[4 * (is + 1)] - xyzc[4 * (is - 1)]) / 2.0 * span;
      const int km = (is == ns - 1) ? is - 1 : is, kp = (is == ns - 1) ? is : is + 1;
      for (int it = 0; it < nT; it++) {
        const SectState& q = st[(size_t)is * nT + it];
        const SectState& qm = st[(size_t)km * nT + it];
        const SectState& qp = st[(size_t)kp * nT + it];
        double tr[3][3], tmp = 0;
        for (int i = 0; i < 3; i++) { tr[0][i] = q.xyzt[i] - q.xyz[i]; tmp += p2(tr[0][i]); }
        for (int i = 0; i < 3; i++) tr[0][i] /= std::sqrt(tmp);
        tmp = 0;
        for (int i = 0; i < 3; i++) { tr[1][i] = qp.xyz[i] - qm.xyz[i]; tmp += p2(tr[1][i]); }
        for (int i = 0; i < 3; i++) tr[1][i] /= std::sqrt(tmp);
        tr[2][0] = tr[0][1] * tr[1][2] - tr[0][2] * tr[1][1];
        tr[2][1] = tr[0][2] * tr[1][0] - tr[0][0] * tr[1][2];
        tr[2][2] = tr[0][0] * tr[1][1] - tr[0][1] * tr[1][0];
        tmp = std::sqrt(p2(tr[2][0]) + p2(tr[2][1]) + p2(tr[2][2]));
        for (int i = 0; i < 3; i++) tr[2][i] /= tmp;
        double rp[3];
        for (int i = 0; i < 3; i++) rp[i] = c->rObs[3 * o + i] - q.xyzt[i];
        const double rpn = std::sqrt(p2(rp[0]) + p2(rp[1]) + p2(rp[2]));
        double rp2[3];
        for (int i = 0; i < 3; i++) rp2[i] = rp[0] * tr[i][0] + rp[1] * tr[i][1] + rp[2] * tr[i][2];
        const double r_e = std::sqrt(p2(rp2[1]) + p2(rp2[2]));
        double* a = out + k * 14;
        a[0] = rpn; a[1] = std::acos(std::fabs(rp2[0] / rpn)); a[2] = std::acos(std::fabs(rp2[1]) / r_e);
        a[3] = xyzc[4 * is + 3] * span; a[4] = lsect; a[5] = q.alpha; a[6] = q.u;
        for (int i = 0; i < 4; i++) a[7 + i] = q.bl[i];
        a[11] = is; a[12] = ns; a[13] = span;
        k++;
      }
    }
  }
  return ACUM_OK;
}
}
