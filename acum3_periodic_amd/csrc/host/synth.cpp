// Synthetic rotor cases in the reference's file contract.
//
// Geometry: per blade, nSpan spanwise strips from rootCut*R to R and nChord cosine-spaced
// chordwise panels per side of a NACA0012 section (half thickness 0.6(0.2969 sqrt x - 0.1260 x -
// 0.3516 x^2 + 0.2843 x^3 - 0.1036 x^4) c), quarter chord on the blade axis, leading edge facing
// the direction of rotation.  Panel area, outward normal and centroid follow the converter's
// conventions (read_surface.py:106-153): area (s1 s2 + s3 s4)/2 from the four edge lengths,
// normal = normalised cross product of the diagonals, centroid = mean of the four nodes.
// Elements are ordered strip by strip, upper side leading->trailing edge then lower side.
// Blade b sits at azimuth 360 b/nB; step k rotates the rotor by k dPsi about z (a coaxial
// second rotor sits coaxSep above and turns the other way).
//
// Pressure coefficient (per element and step, independent N(0, cpNoise) noise from a per-element
// xoshiro256** stream seeded by SplitMix64(seed, surface, element)):
//   model 1: -0.6 (r/R)^2 (1 + 0.1 sin psi)
//   model 2: -+ 0.175 sqrt((1 - x/c)/(x/c)) (r/R)^2 (1 + 0.2 cos psi), suction side negative
//   model 3: model 2 plus four Gaussian pulses in azimuth (width 2 deg, amplitude 0.5 (r/R)^2)
// Observers: arc in the x-z plane, hemisphere below the disk (nAz x nEl) or Fibonacci sphere.
// Scalars are rounded to float, as reading them from the files would do.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>
#include "run.hpp"

using namespace acum_host;

namespace {

uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Rng {  // xoshiro256**
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) s[i] = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {  // Box-Muller
    double u1 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  int uniform_int(int lo, int hi) { return lo + (int)(uniform() * (hi - lo + 1)); }
};

uint64_t stream_seed(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t x = seed ^ (0x632be59bd9b4e019ULL * (a + 1)) ^ (0x9e3779b97f4a7c15ULL * (b + 7));
  splitmix(x);
  return x;
}

double naca_half(double x) {
  return 0.6 * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x -
                0.1036 * x * x * x * x);
}

struct Panel {
  double c[3], n[3], area, rR, xc;
  int side;  // +1 upper, -1 lower
};

void blade_panels(const acum_synth_params* p, std::vector<Panel>* out) {
  const int nc = p->nChord, ns = p->nSpan;
  const double R = p->R, ch = p->chord, r0 = p->rootCut * R;
  std::vector<double> xs(nc + 1);
  for (int i = 0; i <= nc; i++) xs[i] = 0.5 * (1.0 - std::cos(M_PI * i / nc));
  auto node = [&](int s, int i, int side, double* q) {
    q[0] = r0 + (R - r0) * s / ns;
    q[1] = 0.25 * ch - xs[i] * ch;
    q[2] = side * naca_half(xs[i]) * ch;
  };
  out->clear();
  for (int s = 0; s < ns; s++) {
    for (int side = 1; side >= -1; side -= 2) {
      for (int i = 0; i < nc; i++) {
        double P[4][3];
        if (side == 1) {
          node(s, i, 1, P[0]);
          node(s, i + 1, 1, P[1]);
          node(s + 1, i + 1, 1, P[2]);
          node(s + 1, i, 1, P[3]);
        } else {
          node(s, i, -1, P[0]);
          node(s + 1, i, -1, P[1]);
          node(s + 1, i + 1, -1, P[2]);
          node(s, i + 1, -1, P[3]);
        }
        auto dist = [](const double* a, const double* b) {
          return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                           (a[2] - b[2]) * (a[2] - b[2]));
        };
        Panel pn;
        const double s1 = dist(P[1], P[0]), s2 = dist(P[2], P[1]), s3 = dist(P[3], P[2]),
                     s4 = dist(P[0], P[3]);
        pn.area = (s1 * s2 + s3 * s4) / 2;
        double v1[3], v2[3];
        for (int d = 0; d < 3; d++) {
          v1[d] = P[2][d] - P[0][d];
          v2[d] = P[3][d] - P[1][d];
        }
        double n[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                       v1[0] * v2[1] - v1[1] * v2[0]};
        const double nm = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int d = 0; d < 3; d++) {
          pn.n[d] = n[d] / nm;
          pn.c[d] = 0.25 * (P[0][d] + P[1][d] + P[2][d] + P[3][d]);
        }
        pn.rR = pn.c[0] / R;
        pn.xc = 0.5 * (xs[i] + xs[i + 1]);
        pn.side = side;
        out->push_back(pn);
      }
    }
  }
}

double wrap_pi(double a) {
  a = std::fmod(a + M_PI, 2.0 * M_PI);
  if (a < 0) a += 2.0 * M_PI;
  return a - M_PI;
}

const std::vector<float>& iso_bands() {
  static const std::vector<float> t = {10.0f,   12.5f,   16.0f,   20.0f,   25.0f,   31.5f,
                                       40.0f,   50.0f,   63.0f,   80.0f,   100.0f,  125.0f,
                                       160.0f,  200.0f,  250.0f,  315.0f,  400.0f,  500.0f,
                                       630.0f,  800.0f,  1000.0f, 1250.0f, 1600.0f, 2000.0f,
                                       2500.0f, 3150.0f, 4000.0f, 5000.0f, 6300.0f, 8000.0f,
                                       10000.0f, 12500.0f, 16000.0f, 20000.0f};
  return t;
}

// Empirical (untripped) BPM boundary-layer thicknesses over chord at angle a (deg) and Re.
void bl_over_chord(double a, double re, double* out4) {
  const double lr = std::log10(re);
  const double d0 = std::pow(10.0, 1.6569 - 0.9045 * lr + 0.0596 * lr * lr);
  const double ds0 = std::pow(10.0, 3.0187 - 1.5397 * lr + 0.1059 * lr * lr);
  a = std::fabs(a);
  double dsuc, dssuc;
  if (a <= 7.5) {
    dsuc = d0 * std::pow(10.0, 0.03114 * a);
    dssuc = ds0 * std::pow(10.0, 0.0679 * a);
  } else if (a > 12.5) {
    dsuc = d0 * 12.0 * std::pow(10.0, 0.0258 * a);
    dssuc = ds0 * 52.42 * std::pow(10.0, 0.0258 * a);
  } else {
    dsuc = d0 * 0.0303 * std::pow(10.0, 0.2336 * a);
    dssuc = ds0 * 0.0162 * std::pow(10.0, 0.3066 * a);
  }
  out4[0] = dsuc;
  out4[1] = d0 * std::pow(10.0, -0.04175 * a + 0.00106 * a * a);
  out4[2] = dssuc;
  out4[3] = ds0 * std::pow(10.0, -0.0432 * a + 0.00113 * a * a);
}

int check(const acum_synth_params* p) {
  if (!p) return fail(ACUM_ERR_ARG, "null params");
  if (p->nBlades < 1 || p->nChord < 1 || p->nSpan < 1 || p->nTime < 4)
    return fail(ACUM_ERR_ARG, "synth: blades, panels and steps must be positive (nTime >= 4)");
  if (p->R <= 0 || p->chord <= 0 || p->oM == 0 || p->a0 <= 0 || p->obsRadius <= 0)
    return fail(ACUM_ERR_ARG, "synth: bad geometry or speeds");
  if (p->Minf < 0 || p->Minf >= 1) return fail(ACUM_ERR_ARG, "synth: Minf must be in [0,1)");
  if (p->bb && (p->nSect < 3 || (p->nBands != 34 && p->nBands != 42)))
    return fail(ACUM_ERR_ARG, "synth: broadband needs nSect >= 3 and nBands 34 or 42");
  return ACUM_OK;
}

}  // namespace

extern "C" {

int acum_synth_defaults(int32_t config, acum_synth_params* p) {
  if (!p) return fail(ACUM_ERR_ARG, "null params");
  std::memset(p, 0, sizeof(*p));
  p->seed = 31000 + (uint64_t)config;
  p->R = 1.0;
  p->chord = 0.08;
  p->rootCut = 0.2;
  p->a0 = 340.0;
  p->rhoRef = 1.225;
  p->pRef = 101325.0;
  p->twistRoot = 12.0;
  p->twistTip = 4.0;
  p->t_del = 2;
  p->trip = 1;
  p->nSect = 60;
  p->nBands = 34;
  switch (config) {
    case 1:  // two-blade hover rotor, CPU-runnable
      p->nBlades = 2; p->nChord = 25; p->nSpan = 20; p->nTime = 360; p->oM = 200.0;
      p->obs_layout = 0; p->nObs = 36; p->obsRadius = 10.0; p->obsElMin = -45.0; p->obsElMax = 45.0;
      p->cp_model = 1; p->cpNoise = 0.01;
      break;
    case 2:  // four-blade rotor in forward flight, plain vs mixed
      p->nBlades = 4; p->nChord = 100; p->nSpan = 100; p->nTime = 1024; p->oM = 250.0;
      p->Minf = 0.1; p->xy_angle = 0.0; p->xz_angle = -5.0;
      p->obs_layout = 1; p->nObsAz = 32; p->nObsEl = 32; p->obsRadius = 20.0;
      p->cp_model = 2; p->cpNoise = 0.005;
      break;
    case 3:  // broadband: 4 blades x 60 sections, tiny tonal surfaces
      p->nBlades = 4; p->nChord = 1; p->nSpan = 1; p->nTime = 360; p->oM = 250.0;
      p->obs_layout = 1; p->nObsAz = 32; p->nObsEl = 32; p->obsRadius = 20.0;
      p->cp_model = 1; p->cpNoise = 0.0; p->bb = 1;
      break;
    case 5:  // coaxial counter-rotating 2 x 4 blades, stress case
      p->nBlades = 4; p->coaxial = 1; p->coaxSep = 0.15; p->nChord = 125; p->nSpan = 200;
      p->nTime = 2048; p->oM = 250.0; p->Minf = 0.15; p->xz_angle = -5.0;
      p->obs_layout = 2; p->nObs = 4096; p->obsRadius = 30.0;
      p->cp_model = 3; p->cpNoise = 0.005;
      break;
    default:
      return fail(ACUM_ERR_ARG, "unknown config (1, 2, 3 or 5; config 4 is acum_synth_batch_case)");
  }
  return ACUM_OK;
}

int acum_synth_batch_case(uint64_t seed, int32_t icase, acum_synth_params* p) {
  if (!p || icase < 0) return fail(ACUM_ERR_ARG, "bad argument");
  acum_synth_defaults(1, p);
  Rng g(stream_seed(seed, 0xCA5E, (uint64_t)icase));
  p->seed = stream_seed(seed, 0x5EED, (uint64_t)icase);
  p->nBlades = g.uniform_int(2, 6);
  const int el = g.uniform_int(2000, 40000);  // elements per blade = 2 nChord nSpan
  int nc = std::max(5, (int)std::lround(std::sqrt(el / 2.0 / 2.0)));
  int ns = std::max(5, (int)std::lround(el / 2.0 / nc));
  p->nChord = nc;
  p->nSpan = ns;
  p->nTime = g.uniform_int(256, 2048);
  p->obs_layout = 2;
  p->nObs = g.uniform_int(16, 2048);
  const double rad = 5.0 + 45.0 * g.uniform();
  p->oM = 150.0 + 250.0 * g.uniform();
  p->Minf = 0.2 * g.uniform();
  // Radius: 1 m, reduced where needed so the advancing blade tip stays below Mach 0.85 (the
  // Farassat 1A integrand is singular at Mr = 1; at oM = 400 rad/s a 1 m blade is supersonic).
  p->R = std::min(1.0, (0.85 - p->Minf) * p->a0 / p->oM);
  p->chord = 0.08 * p->R;
  p->obsRadius = p->R * rad;
  p->xz_angle = -5.0;
  p->cp_model = 1;
  p->cpNoise = 0.01;
  return ACUM_OK;
}

int acum_synth(const acum_synth_params* p, acum_run** out) {
  if (!out) return fail(ACUM_ERR_ARG, "null output");
  *out = nullptr;
  int rc = check(p);
  if (rc) return rc;
  acum_run* r = new acum_run();
  const int nB = p->nBlades, nS = nB * (p->coaxial ? 2 : 1), nT = p->nTime;
  acum_inputs& in = r->in;
  in.a0 = (float)p->a0;
  in.rhoRef = (float)p->rhoRef;
  in.pRef = (float)p->pRef;
  in.oM = (float)p->oM;
  in.dPsi_deg = (float)(360.0 / nT);
  in.Minf = (float)p->Minf;
  in.Mref = (float)(std::fabs(p->oM) * p->R / p->a0);
  in.xy_angle = (float)p->xy_angle;
  in.xz_angle = (float)p->xz_angle;
  in.CFDscale = 1.0f;
  in.OBSscale = 1.0f;
  in.nSurf = nS;
  in.nTime = nT;
  in.periodic = 1;
  in.lowpass = 0;
  in.impermeable = 1;
  in.BB_noise = p->bb ? 1 : 0;
  derive_inputs(&in, &r->dPsi_f);
  // observers
  std::vector<double> obs;
  if (p->obs_layout == 0) {
    const int n = p->nObs > 0 ? p->nObs : 1;
    for (int k = 0; k < n; k++) {
      const double el = (p->obsElMin + (p->obsElMax - p->obsElMin) * (n > 1 ? (double)k / (n - 1) : 0.0)) * M_PI / 180;
      obs.insert(obs.end(), {p->obsRadius * std::cos(el), 0.0, p->obsRadius * std::sin(el)});
    }
  } else if (p->obs_layout == 1) {
    for (int i = 0; i < p->nObsAz; i++)
      for (int j = 0; j < p->nObsEl; j++) {
        const double az = 2.0 * M_PI * i / p->nObsAz;
        const double el = -0.5 * M_PI * (j + 0.5) / p->nObsEl;
        obs.insert(obs.end(), {p->obsRadius * std::cos(el) * std::cos(az),
                               p->obsRadius * std::cos(el) * std::sin(az), p->obsRadius * std::sin(el)});
      }
  } else {
    const int n = p->nObs;
    const double ga = M_PI * (3.0 - std::sqrt(5.0));
    for (int i = 0; i < n; i++) {
      const double z = 1.0 - 2.0 * (i + 0.5) / n, rr = std::sqrt(1.0 - z * z), ph = ga * i;
      obs.insert(obs.end(), {p->obsRadius * rr * std::cos(ph), p->obsRadius * rr * std::sin(ph),
                             p->obsRadius * z});
    }
  }
  r->nObs = (int64_t)obs.size() / 3;
  if (r->nObs < 1) {
    delete r;
    return fail(ACUM_ERR_ARG, "synth: no observers");
  }
  r->rObs.resize(obs.size());
  for (size_t i = 0; i < obs.size(); i++) r->rObs[i] = (float)obs[i];
  // surfaces
  std::vector<Panel> pan;
  blade_panels(p, &pan);
  const int64_t E = (int64_t)pan.size();
  r->nSurf = nS;
  r->nTime = nT;
  r->E.assign(nS, E);
  layout_surfaces(r);
  double pulse_az[4];
  {
    Rng g(stream_seed(p->seed, 0xB0B, 0));
    for (int m = 0; m < 4; m++) pulse_az[m] = 2.0 * M_PI * g.uniform();
  }
  const double dpsi = 2.0 * M_PI / nT;
  for (int s = 0; s < nS; s++) {
    const int rotor = s / nB, b = s % nB;
    const double sense = rotor == 0 ? 1.0 : -1.0;
    const double zoff = rotor == 0 ? 0.0 : p->coaxSep;
    const double off = 2.0 * M_PI * b / nB;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < E; j++) {
      const Panel& q = pan[j];
      Rng g(stream_seed(p->seed, (uint64_t)s, (uint64_t)j));
      const double cy = q.c[1] * sense, ny = q.n[1] * sense;  // mirror for the clockwise rotor
      for (int k = 0; k < nT; k++) {
        const double psi = sense * k * dpsi + off;  // blade azimuth
        const double cs = std::cos(psi), sn = std::sin(psi);
        float* Xh = &r->X[r->XPtr[s] + (j * nT + k) * 4];
        float* Nh = &r->N[r->NPtr[s] + (j * nT + k) * 6];
        float* Ph = &r->PR[r->PRPtr[s] + (j * nT + k) * 2];
        Xh[0] = (float)(q.c[0] * cs - cy * sn);
        Xh[1] = (float)(q.c[0] * sn + cy * cs);
        Xh[2] = (float)(q.c[2] + zoff);
        Xh[3] = (float)q.area;
        Nh[0] = (float)(q.n[0] * cs - ny * sn);
        Nh[1] = (float)(q.n[0] * sn + ny * cs);
        Nh[2] = (float)q.n[2];
        const double r2 = q.rR * q.rR;
        const double azm = sense * k * dpsi + off;  // azimuth in the rotor's own sense
        double cp;
        if (p->cp_model == 1) {
          cp = -0.6 * r2 * (1.0 + 0.1 * std::sin(azm));
        } else {
          const double x = std::max(q.xc, 0.002);
          cp = -q.side * 0.175 * std::sqrt((1.0 - x) / x) * r2 * (1.0 + 0.2 * std::cos(azm));
          if (p->cp_model == 3) {
            double pul = 0.0;
            const double w = 2.0 * M_PI / 180.0;
            for (int m = 0; m < 4; m++) {
              const double d = wrap_pi(sense * azm - pulse_az[m]) / w;
              pul += 0.5 * std::exp(-0.5 * d * d);
            }
            cp += -q.side * pul * r2;
          }
        }
        if (p->cpNoise > 0) cp += p->cpNoise * g.normal();
        Ph[0] = (float)cp;
      }
    }
  }
  // broadband inputs
  if (p->bb) {
    const int nBB = nB;  // blades of the first rotor
    r->nSurf_BB = nBB;
    r->t_del = p->t_del;
    r->span.assign(nS, 1.0f);
    r->omega.assign(nS, 1.0f);
    r->ccw.assign(nS, 1);
    r->psi_offset.assign(nS, 0.0f);
    r->trip.assign(nS, 0);
    r->rotate.assign(3 * nS, 0.0f);
    r->translate.assign(3 * nS, 0.0f);
    const int ns = p->nSect;
    r->nsect.assign(nBB, ns);
    r->xsPtr.assign(nBB, 0);
    r->dsPtr.assign(nBB, 0);
    for (int i = 1; i < nBB; i++) {
      r->xsPtr[i] = r->xsPtr[i - 1] + 4 * ns;
      r->dsPtr[i] = r->dsPtr[i - 1] + (int64_t)7 * ns * nT;
    }
    r->xyzcR.assign((size_t)4 * ns * nBB, 0.0f);
    r->bbdata.assign((size_t)7 * ns * nT * nBB, 0.0f);
    const double nu = 0.0000181206 / 1.22500;
    for (int i = 0; i < nBB; i++) {
      r->span[i] = (float)p->R;
      r->omega[i] = (float)p->oM;
      r->psi_offset[i] = (float)(360.0 * i / nBB);
      r->trip[i] = p->trip;
      Rng g(stream_seed(p->seed, 0xBB00 + i, 0));
      for (int k = 0; k < ns; k++) {
        float* x = &r->xyzcR[r->xsPtr[i] + 4 * k];
        x[0] = (float)(0.2 + 0.8 * k / (ns - 1));
        x[1] = 0.0f;
        x[2] = 0.0f;
        x[3] = (float)(p->chord / p->R);
      }
      for (int t = 0; t < nT; t++) {
        const double psi = t * dpsi + 2.0 * M_PI * i / nBB;
        for (int k = 0; k < ns; k++) {
          float* bd = &r->bbdata[r->dsPtr[i] + (int64_t)t * ns * 7 + k * 7];
          const double rR = r->xyzcR[r->xsPtr[i] + 4 * k];
          const double re = std::fabs(p->oM) * rR * p->R * p->chord / nu;
          double bl[4];
          bl_over_chord(4.0, re, bl);
          for (int w = 0; w < 4; w++) bd[w] = (float)(bl[w] * p->chord);
          // section coefficients 0.5 + 0.05 sin psi and 0.01 (plus noise), written normalised by
          // the tip speed as the file contract asks: the reader divides them by (r/R)^2
          const double cn = 0.5 + 0.05 * std::sin(psi) + 0.002 * g.normal();
          const double cc = 0.01 + 0.002 * g.normal();
          bd[4] = (float)(cn * rR * rR);
          bd[5] = (float)(cc * rR * rR);
          bd[6] = (float)(p->twistRoot + (p->twistTip - p->twistRoot) * (rR - 0.2) / 0.8);
        }
      }
    }
    // aerodynamic tables: 3 sections, 4 Mach, 3 Re, AoA -10..20 deg
    const float secs[3] = {0.2f, 0.6f, 1.0f}, machs[4] = {0.1f, 0.3f, 0.5f, 0.7f},
                res[3] = {1e5f, 1e6f, 1e7f};
    const int na = 31;
    r->nsect_t.assign(nBB, 3);
    r->nmach_t.assign(nBB, 4);
    r->naoa_t.assign(nBB, na);
    r->nre_t.assign(nBB, 3);
    r->aPtr.assign(nBB, 0);
    r->mPtr.assign(nBB, 0);
    r->rPtr.assign(nBB, 0);
    r->sPtr.assign(nBB, 0);
    r->dPtr.assign(nBB, 0);
    for (int i = 1; i < nBB; i++) {
      r->aPtr[i] = r->aPtr[i - 1] + na;
      r->mPtr[i] = r->mPtr[i - 1] + 4;
      r->rPtr[i] = r->rPtr[i - 1] + 3;
      r->sPtr[i] = r->sPtr[i - 1] + 3;
      r->dPtr[i] = r->dPtr[i - 1] + (int64_t)3 * 4 * 3 * na * 7;
    }
    r->sect_t.clear();
    r->mach_t.clear();
    r->re_t.clear();
    r->aoa_t.clear();
    r->bbdata_t.clear();
    for (int i = 0; i < nBB; i++) {
      r->sect_t.insert(r->sect_t.end(), secs, secs + 3);
      r->mach_t.insert(r->mach_t.end(), machs, machs + 4);
      r->re_t.insert(r->re_t.end(), res, res + 3);
      for (int m = 0; m < na; m++) r->aoa_t.push_back((float)(-10 + m));
      for (int j = 0; j < 3; j++)
        for (int k = 0; k < 4; k++)
          for (int l = 0; l < 3; l++)
            for (int m = 0; m < na; m++) {
              const double a = -10.0 + m;
              double bl[4];
              bl_over_chord(a, res[l], bl);
              const double cl = 2.0 * M_PI * std::sin(a * M_PI / 180.0) / std::sqrt(1.0 - machs[k] * machs[k]);
              const float v[7] = {(float)cl, 0.01f, 0.0f, (float)bl[0], (float)bl[1], (float)bl[2], (float)bl[3]};
              r->bbdata_t.insert(r->bbdata_t.end(), v, v + 7);
            }
    }
    r->ff = (p->nBands == 34) ? iso_bands() : reference_bands();
    r->nf2 = (int32_t)r->ff.size();
    // what read_BB_inputs does to cn and cc (io.cu:509-512)
    r->bbdata_raw = r->bbdata;
    for (int i = 0; i < nBB; i++)
      for (int t = 0; t < nT; t++)
        for (int k = 0; k < ns; k++) {
          float* bd = &r->bbdata[r->dsPtr[i] + (int64_t)t * ns * 7 + k * 7];
          const double rr = (double)r->xyzcR[r->xsPtr[i] + 4 * k];
          bd[4] = (float)((double)bd[4] / (rr * rr));
          bd[5] = (float)((double)bd[5] / (rr * rr));
        }
    r->has_bb = true;
  }
  *out = r;
  return ACUM_OK;
}

// Writes the run's inputs in the reference's file formats.  Floats are printed with 9
// significant digits, which round-trips every float exactly.
int acum_run_save(const acum_run* r, const char* dir) {
  if (!r || !dir) return fail(ACUM_ERR_ARG, "null argument");
  if (r->preprocessed) return fail(ACUM_ERR_ARG, "run already preprocessed; save the raw run");
  const std::string d = dir;
  auto path = [&](const std::string& n) { return d + "/" + n; };
  ::mkdir(d.c_str(), 0755);
  ::mkdir(path("surfaces").c_str(), 0755);
  FILE* f = std::fopen(path("input").c_str(), "w");
  if (!f) return fail(ACUM_ERR_IO, "cannot write " + path("input"));
  const acum_inputs& in = r->in;
  std::fprintf(f, "ACUM-3_INPUTS\n***Environmental\na0 %.9g\nrhoRef %.9g\npRef %.9g\n", in.a0, in.rhoRef, in.pRef);
  std::fprintf(f, "***Geometric\noM %.9g\nnSurf %d\nnTime %d\ndPsi %.9g\nMinf %.9g\nMref %.9g\n", in.oM,
               in.nSurf, in.nTime, in.dPsi_deg, in.Minf, in.Mref);
  std::fprintf(f, "xy_angle %.9g\nxz_angle %.9g\nCFDscale %.9g\nOBSscale %.9g\n", in.xy_angle,
               in.xz_angle, in.CFDscale, in.OBSscale);
  std::fprintf(f, "***Data\nimpermeable %d\nperiodic %d\nlowpass %d\nBB_noise %d\n", in.impermeable,
               in.periodic, in.lowpass, in.BB_noise);
  std::fclose(f);
  f = std::fopen(path("obs.dat").c_str(), "w");
  if (!f) return fail(ACUM_ERR_IO, "cannot write obs.dat");
  std::fprintf(f, "%d\n", (int)r->nObs);
  for (int64_t i = 0; i < r->nObs; i++)
    std::fprintf(f, "%.9g %.9g %.9g\n", r->rObs[3 * i], r->rObs[3 * i + 1], r->rObs[3 * i + 2]);
  std::fclose(f);
  const int64_t nT = r->nTime;
  for (int64_t s = 0; s < r->nSurf; s++) {
    f = std::fopen(path("surfaces/surface_" + std::to_string(s + 1) + ".dat").c_str(), "w");
    if (!f) return fail(ACUM_ERR_IO, "cannot write surface file");
    std::fprintf(f, "%d %d\n", (int)r->E[s], (int)nT);
    // one block of E lines per step, formatted in parallel; each value is the shortest decimal
    // that reads back as the same float
    const int64_t E = r->E[s];
    std::vector<std::string> blocks(64);
    for (int64_t k0 = 0; k0 < nT; k0 += 64) {
      const int64_t k1 = std::min<int64_t>(nT, k0 + 64);
#pragma omp parallel for schedule(dynamic, 1)
      for (int64_t k = k0; k < k1; k++) {
        std::string& b = blocks[(size_t)(k - k0)];
        b.clear();
        b.reserve((size_t)E * 8 * 16);
        char num[32];
        for (int64_t j = 0; j < E; j++) {
          const float* Xh = &r->X[r->XPtr[s] + (j * nT + k) * 4];
          const float* Nh = &r->N[r->NPtr[s] + (j * nT + k) * 6];
          const float* Ph = &r->PR[r->PRPtr[s] + (j * nT + k) * 2];
          const float v[8] = {Xh[0], Xh[1], Xh[2], Ph[0], Xh[3], Nh[0], Nh[1], Nh[2]};
          for (int q = 0; q < 8; q++) {
            const auto res = std::to_chars(num, num + sizeof(num), v[q]);
            b.append(num, (size_t)(res.ptr - num));
            b.push_back(q == 7 ? '\n' : ' ');
          }
        }
      }
      for (int64_t k = k0; k < k1; k++) std::fwrite(blocks[(size_t)(k - k0)].data(), 1, blocks[(size_t)(k - k0)].size(), f);
    }
    std::fclose(f);
  }
  if (r->has_bb) {
    ::mkdir(path("BB_inputs").c_str(), 0755);
    f = std::fopen(path("input_BB").c_str(), "w");
    if (!f) return fail(ACUM_ERR_IO, "cannot write input_BB");
    std::fprintf(f, "ACUM-3_Broadband_Inputs\nnSurf_BB %d\nt_del %d\n", r->nSurf_BB, r->t_del);
    for (int i = 0; i < r->nSurf_BB; i++) {
      std::fprintf(f, "***Surface-%d\nspan(m) %.9g\nomega(rad/s) %.9g\nccw %d\npsi_offset %.9g\ntrip %d\n",
                   i + 1, r->span[i], r->omega[i], r->ccw[i], r->psi_offset[i], r->trip[i]);
      std::fprintf(f, "Initial_Position\nrotate %.9g %.9g %.9g\ntranslate %.9g %.9g %.9g\n", r->rotate[3 * i],
                   r->rotate[3 * i + 1], r->rotate[3 * i + 2], r->translate[3 * i], r->translate[3 * i + 1],
                   r->translate[3 * i + 2]);
    }
    std::fclose(f);
    for (int i = 0; i < r->nSurf_BB; i++) {
      const int ns = r->nsect[i];
      f = std::fopen(path("BB_inputs/BB_surface_" + std::to_string(i + 1) + ".dat").c_str(), "w");
      if (!f) return fail(ACUM_ERR_IO, "cannot write BB surface file");
      std::fprintf(f, "%d %d\n", (int)nT, ns);
      for (int k = 0; k < ns; k++) {
        const float* x = &r->xyzcR[r->xsPtr[i] + 4 * k];
        std::fprintf(f, "%.9e %.9e %.9e %.9e\n", x[0], x[1], x[2], x[3]);
      }
      for (int64_t t = 0; t < nT; t++)
        for (int k = 0; k < ns; k++) {
          const float* bd = &r->bbdata_raw[r->dsPtr[i] + t * ns * 7 + k * 7];
          const float cn = bd[4], cc = bd[5];
          std::fprintf(f, "%.9e %.9e %.9e %.9e %.9e %.9e %.9e\n", bd[0], bd[1], bd[2], bd[3], cn, cc, bd[6]);
        }
      std::fclose(f);
      f = std::fopen(path("BB_inputs/BB_table_" + std::to_string(i + 1) + ".dat").c_str(), "w");
      if (!f) return fail(ACUM_ERR_IO, "cannot write BB table");
      std::fprintf(f, "VARIABLES=\"alpha\",\"mach\",\"cl\",\"cd\",\"cm\",\"bl_upper\",\"bl_lower\",\"disp_upper\",\"disp_lower\",\"Re\"\n");
      const int nst = r->nsect_t[i], na = r->naoa_t[i], nm = r->nmach_t[i], nr = r->nre_t[i];
      std::fprintf(f, "%d %d %d %d\n", nst, na, nm, nr);
      for (int j = 0; j < nst; j++) {
        std::fprintf(f, "%.9g\n", r->sect_t[r->sPtr[i] + j]);
        for (int k = 0; k < nm; k++)
          for (int l = 0; l < nr; l++)
            for (int m = 0; m < na; m++) {
              const float* v = &r->bbdata_t[r->dPtr[i] + (int64_t)j * nm * nr * na * 7 + (int64_t)k * nr * na * 7 + l * na * 7 + m * 7];
              std::fprintf(f, "%.9e %.9e %.9e %.9e %.9e %.9e %.9e %.9e %.9e %.9e\n", r->aoa_t[r->aPtr[i] + m],
                           r->mach_t[r->mPtr[i] + k], v[0], v[1], v[2], v[3], v[4], v[5], v[6], r->re_t[r->rPtr[i] + l]);
            }
      }
      std::fclose(f);
    }
    const std::vector<float>& rb = reference_bands();
    const bool ref = r->ff.size() == rb.size() && std::equal(rb.begin(), rb.end(), r->ff.begin());
    if (!ref) {
      f = std::fopen(path("BB_inputs/BB_bands.dat").c_str(), "w");
      if (!f) return fail(ACUM_ERR_IO, "cannot write BB_bands.dat");
      std::fprintf(f, "%d\n", (int)r->ff.size());
      for (float v : r->ff) std::fprintf(f, "%.9g\n", v);
      std::fclose(f);
    }
  }
  return ACUM_OK;
}

}  // extern "C"
