// Host FP64 preparation of the BPM sources (see bpm_prep.hpp).
// Follows findBBnoise's per-surface preprocessing: blade section kinematics with the
// quarter-chord y, z taken from section round(0.75 nsect) (bb_core.cu:79-164), relative flow by
// periodic central differences (bb_core.cu:166-206), Mach/Re brackets and trilinear cl lookup
// (bb_core.cu:208-250, bb_utils.cu:124-168), cl slope, cl(0) and alpha(cl=0) (bb_core.cu:252-287),
// the fixed-point effective angle (bb_core.cu:291-321) and the boundary-layer thickness models
// t_del = 0, 1, 2 (bb_core.cu:354-438); then the section frames of the observer loop
// (bb_core.cu:473-513), which do not depend on the observer.
#include "bpm_prep.hpp"

#include <cmath>
#include <cstdint>
#include <vector>

namespace acum {
namespace {

inline double sq(double x) { return x * x; }

struct TableView {
  const acum_bpm_case* c;
  int s;
  int nm, nr, ns, na;
  const float *sect, *mach, *re, *aoa, *data;
  TableView(const acum_bpm_case* cc, int ss) : c(cc), s(ss) {
    nm = c->nmach_t[s];
    nr = c->nre_t[s];
    ns = c->nsect_t[s];
    na = c->naoa_t[s];
    sect = c->sect_t + c->sPtr[s];
    mach = c->mach_t + c->mPtr[s];
    re = c->re_t + c->rPtr[s];
    aoa = c->aoa_t + c->aPtr[s];
    data = c->bbdata_t + c->dPtr[s];
  }
  double at(int is, int im, int ir, int ia, int id) const {
    return data[(((int64_t)is * nm + im) * nr + ir) * na * 7 + (int64_t)ia * 7 + id];
  }
  // the first index whose axis value exceeds x, then the clamped bracket (bb_core.cu:208-234)
  static void bracket(const float* ax, int n, double x, int* lo, int* hi) {
    int k = 0;
    while (k < n && !(x < ax[k])) k++;
    *lo = k - 1 > 0 ? k - 1 : 0;
    *hi = k < n - 1 ? k : n - 1;
  }
  double lerp3(double rR, int s1, int s2, double M, int m1, int m2, double Re, int r1, int r2,
               int id, int ia) const {
    const double ws = (s1 == s2) ? 0.5 : (rR - sect[s1]) / ((double)sect[s2] - sect[s1]);
    const double wm = (m1 == m2) ? 0.5 : (M - mach[m1]) / ((double)mach[m2] - mach[m1]);
    const double wr = (r1 == r2) ? 0.5 : (Re - re[r1]) / ((double)re[r2] - re[r1]);
    auto sec = [&](int im, int ir) { return (1 - ws) * at(s1, im, ir, ia, id) + ws * at(s2, im, ir, ia, id); };
    return (1 - wm) * (1 - wr) * sec(m1, r1) + wm * (1 - wr) * sec(m2, r1) + (1 - wm) * wr * sec(m1, r2) +
           wm * wr * sec(m2, r2);
  }
};

}  // namespace

int bpm_prepare(const acum_bpm_case* c, std::vector<BpmSource>* out, std::string* err) {
  const int nS = c->nSurf_BB, nT = c->nTime;
  int64_t tot = 0;
  for (int s = 0; s < nS; s++) tot += c->nsect[s];
  out->assign((size_t)tot * nT, BpmSource());
  const double nu = 0.0000181206 / 1.22500;
  const double d2r = M_PI / 180.0;
  int64_t soff = 0;
  for (int s = 0; s < nS; s++) {
    const int ns = c->nsect[s];
    const double span = c->span[s], omega = c->omega[s];
    const int ccw = c->ccw[s];
    const float* xc = c->xyzcR + c->xsPtr[s];
    const float* bd = c->bbdata + c->dsPtr[s];
    const int iq = (int)std::lround(ns * 0.75);
    if (ns < 2 || iq >= ns) {
      *err = "broadband surface needs at least 3 sections";
      return ACUM_ERR_ARG;
    }
    const double thx = c->rotate[3 * s] * d2r, thy = c->rotate[3 * s + 1] * d2r, thz = c->rotate[3 * s + 2] * d2r;
    const double trn[3] = {c->translate[3 * s], c->translate[3 * s + 1], c->translate[3 * s + 2]};
    std::vector<double> xyz((size_t)ns * nT * 3), xyzt((size_t)ns * nT * 3);
    TableView tab(c, s);
    int status = ACUM_OK;
#pragma omp parallel for schedule(static)
    for (int is = 0; is < ns; is++) {
      const double ch = xc[4 * is + 3] * span;
      for (int it = 0; it < nT; it++) {
        double x0[3] = {xc[4 * is] * span, xc[4 * iq + 1] * span, xc[4 * iq + 2] * span};
        if (ccw == 0) x0[1] = -x0[1];
        double psi = c->dTau * it * omega + c->psi_offset[s] * d2r;
        if (ccw == 0) psi = -psi;
        const double sp = std::sin(psi), cp = std::cos(psi);
        const double th = bd[(int64_t)it * ns * 7 + is * 7 + 6] * d2r;
        double xt0[3] = {x0[0], 0.0, x0[2] - ch * 0.75 * std::sin(th)};
        xt0[1] = (ccw == 1) ? x0[1] - ch * 0.75 * std::cos(th) : x0[1] + ch * 0.75 * std::cos(th);
        // azimuth, then the fixed rotations about x, y, z, then the translation
        auto place = [&](const double* a, double* o) {
          double p[3] = {a[0] * cp - a[1] * sp, a[0] * sp + a[1] * cp, a[2]};
          double q[3] = {p[0], p[1] * std::cos(thx) - p[2] * std::sin(thx), p[1] * std::sin(thx) + p[2] * std::cos(thx)};
          double r[3] = {q[0] * std::cos(thy) + q[2] * std::sin(thy), q[1], -q[0] * std::sin(thy) + q[2] * std::cos(thy)};
          o[0] = r[0] * std::cos(thz) - r[1] * std::sin(thz) + trn[0];
          o[1] = r[0] * std::sin(thz) + r[1] * std::cos(thz) + trn[1];
          o[2] = r[2] + trn[2];
        };
        place(x0, &xyz[((size_t)is * nT + it) * 3]);
        place(xt0, &xyzt[((size_t)is * nT + it) * 3]);
      }
      std::vector<double> cl(tab.na);
      for (int it = 0; it < nT; it++) {
        BpmSource& q = (*out)[(size_t)(soff + is) * nT + it];
        const double* P = &xyz[((size_t)is * nT + it) * 3];
        const double* Pm = &xyz[((size_t)is * nT + (it == 0 ? nT - 1 : it - 1)) * 3];
        const double* Pp = &xyz[((size_t)is * nT + (it == nT - 1 ? 0 : it + 1)) * 3];
        const double* Pt = &xyzt[((size_t)is * nT + it) * 3];
        double v[3], ch1[3];
        for (int i = 0; i < 3; i++) v[i] = c->vInf[i] - (Pp[i] - Pm[i]) / (2 * c->dTau);
        for (int i = 0; i < 3; i++) ch1[i] = Pt[i] - P[i];
        const double l1 = std::sqrt(sq(ch1[0]) + sq(ch1[1]) + sq(ch1[2]));
        q.u = (v[0] * ch1[0] + v[1] * ch1[1] + v[2] * ch1[2]) / l1;
        const double M = std::fabs(q.u / c->a0), Re = std::fabs(q.u * ch / nu), rR = xc[4 * is];
        int m1, m2, r1, r2, s1, s2;
        TableView::bracket(tab.mach, tab.nm, M, &m1, &m2);
        TableView::bracket(tab.re, tab.nr, Re, &r1, &r2);
        TableView::bracket(tab.sect, tab.ns, rR, &s1, &s2);
        for (int i = 0; i < tab.na; i++) cl[i] = tab.lerp3(rR, s1, s2, M, m1, m2, Re, r1, r2, 0, i);
        int a1 = -1, a2 = -1, a3 = -1;
        for (int i = 0; i < tab.na && (a1 < 0 || a2 < 0 || a3 < 0); i++) {
          if (a1 < 0 && tab.aoa[i] >= 0.0) a1 = i;
          if (a2 < 0 && tab.aoa[i] >= 6.0) a2 = i;
          if (a3 < 0 && cl[i] > 0) a3 = i;
        }
        if (a1 < 1 || a2 < 0 || a3 < 1) {
#pragma omp critical
          status = ACUM_ERR_ARG;
          break;
        }
        const double cla = (cl[a2] - cl[a1]) / ((double)tab.aoa[a2] - tab.aoa[a1]);
        double w = (0 - tab.aoa[a1 - 1]) / ((double)tab.aoa[a1] - tab.aoa[a1 - 1]);
        const double cl0 = cl[a1 - 1] * (1 - w) + cl[a1] * w;
        w = (0 - cl[a3 - 1]) / (cl[a3] - cl[a3 - 1]);
        const double acl0 = tab.aoa[a3 - 1] * (1 - w) + tab.aoa[a3] * w;
        const double cn = bd[(int64_t)it * ns * 7 + is * 7 + 4], ct = bd[(int64_t)it * ns * 7 + is * 7 + 5];
        double ae = 0.0, e = 1.0;
        for (int iter = 0; e > 1e-3;) {
          const double clsec = cn * std::cos(ae * d2r) - ct * std::sin(ae * d2r);
          const double ae1 = (clsec - cl0) / cla;
          e = std::fabs(ae1 - ae);
          ae = ae1;
          if (++iter > 50) break;
        }
        q.alpha = std::fabs(ae - acl0);
        double bl[4];
        if (c->t_del == 0) {
          for (int i = 0; i < 4; i++) bl[i] = bd[(int64_t)it * ns * 7 + is * 7 + i];
        } else if (c->t_del == 1) {
          for (int i = 0; i < tab.na; i++)
            if (tab.aoa[i] >= ae) {
              a1 = i;
              break;
            }
          w = (ae - tab.aoa[a1 - 1]) / ((double)tab.aoa[a1] - tab.aoa[a1 - 1]);
          for (int i = 0; i < 4; i++) {
            const double lo = tab.lerp3(rR, s1, s2, M, m1, m2, Re, r1, r2, i + 3, a1 - 1);
            const double hi = tab.lerp3(rR, s1, s2, M, m1, m2, Re, r1, r2, i + 3, a1);
            bl[i] = ch * (lo * (1 - w) + hi * w);
          }
        } else {
          const double lr = std::log10(Re), al = q.alpha;
          const double d0 = ch * std::pow(10.0, 1.6569 - 0.9045 * lr + 0.0596 * lr * lr);
          const double ds0 = ch * std::pow(10.0, 3.0187 - 1.5397 * lr + 0.1059 * lr * lr);
          bl[1] = d0 * std::pow(10.0, -0.04175 * al + 0.00106 * al * al);
          bl[3] = ds0 * std::pow(10.0, -0.0432 * al + 0.00113 * al * al);
          if (al <= 7.5) bl[0] = d0 * std::pow(10.0, 0.03114 * al);
          else if (al > 12.5) bl[0] = d0 * 12.0 * std::pow(10.0, 0.0258 * al);
          else bl[0] = d0 * 0.0303 * std::pow(10.0, 0.2336 * al);
          if (al <= 7.5) bl[2] = ds0 * std::pow(10.0, 0.0679 * al);
          else if (al > 12.5) bl[2] = ds0 * 52.42 * std::pow(10.0, 0.0258 * al);
          else bl[2] = ds0 * 0.0162 * std::pow(10.0, 0.3066 * al);
        }
        for (int i = 0; i < 4; i++) q.bl[i] = bl[i] < 0.0 ? 1.0e-10 : bl[i];
        q.c = ch;
        q.R = span;
        q.trip = c->trip[s];
        q.tip = (is == ns - 1) ? 1 : 0;
        for (int i = 0; i < 3; i++) q.xyzt[i] = Pt[i];
      }
    }
    if (status) {
      *err = "broadband table lacks a negative AoA, an AoA >= 6 deg or a cl sign change";
      return status;
    }
    // section length and frame (bb_core.cu:457-513)
#pragma omp parallel for schedule(static)
    for (int is = 0; is < ns; is++) {
      double ls;
      if (is == 0) ls = ((double)xc[4 * (is + 1)] - xc[4 * is]) / 2.0 * span;
      else if (is == ns - 1) ls = ((double)xc[4 * is] - xc[4 * (is - 1)]) / 2.0 * span;
      else ls = ((double)xc[4 * (is + 1)] - xc[4 * (is - 1)]) / 2.0 * span;
      const int km = (is == ns - 1) ? is - 1 : is, kp = (is == ns - 1) ? is : is + 1;
      for (int it = 0; it < nT; it++) {
        BpmSource& q = (*out)[(size_t)(soff + is) * nT + it];
        q.lsect = ls;
        const double* P = &xyz[((size_t)is * nT + it) * 3];
        const double* Pt = &xyzt[((size_t)is * nT + it) * 3];
        const double* Pkm = &xyz[((size_t)km * nT + it) * 3];
        const double* Pkp = &xyz[((size_t)kp * nT + it) * 3];
        double t0[3], t1[3], t2[3];
        for (int i = 0; i < 3; i++) {
          t0[i] = Pt[i] - P[i];
          t1[i] = Pkp[i] - Pkm[i];
        }
        const double n0 = std::sqrt(sq(t0[0]) + sq(t0[1]) + sq(t0[2]));
        const double n1 = std::sqrt(sq(t1[0]) + sq(t1[1]) + sq(t1[2]));
        for (int i = 0; i < 3; i++) {
          t0[i] /= n0;
          t1[i] /= n1;
        }
        if (ccw == 1) {
          t2[0] = t0[1] * t1[2] - t0[2] * t1[1];
          t2[1] = t0[2] * t1[0] - t0[0] * t1[2];
          t2[2] = t0[0] * t1[1] - t0[1] * t1[0];
          const double n2 = std::sqrt(sq(t2[0]) + sq(t2[1]) + sq(t2[2]));
          for (int i = 0; i < 3; i++) t2[i] /= n2;
        } else {
          t2[0] = t1[1] * t0[2] - t1[2] * t0[1];
          t2[1] = t1[2] * t0[0] - t1[0] * t0[2];
          t2[2] = t1[0] * t0[1] - t1[1] * t0[0];
          const double n2 = std::sqrt(sq(t2[0]) + sq(t2[1]) + sq(t2[2]));
          for (int i = 0; i < 3; i++) t2[i] /= std::sqrt(n2);  // as bb_core.cu:512
        }
        for (int i = 0; i < 3; i++) {
          q.tr[i] = t0[i];
          q.tr[3 + i] = t1[i];
          q.tr[6 + i] = t2[i];
        }
      }
    }
    soff += ns;
  }
  return ACUM_OK;
}

}  // namespace acum
