// File contract of the reference (io.cu): the `input` keyword file (io.cu:12-101), obs.dat
// (io.cu:106-135), surfaces/surface_k.dat (io.cu:206-333), input_BB and BB_inputs/*
// (io.cu:337-700), and the writers of pT/pL/pA.out (io.cu:137-203) and bb_oaspl.out /
// bb_spl.dat (io.cu:702-796).  Values are parsed into float exactly as fscanf("%f")/">>"
// would (correctly rounded), errors are returned instead of exit(1), offsets are 64-bit.
// An optional BB_inputs/BB_bands.dat ("n" then n frequencies) replaces the reference's fixed
// 42-band table; the reference ignores that file, so the directory stays readable by both.
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>
#include <omp.h>
#include "run.hpp"

namespace acum_host {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

const std::vector<float>& reference_bands() {
  static const std::vector<float> t = {
      1.0f,    1.25f,   1.6f,    2.0f,    2.5f,    3.15f,   4.0f,    5.0f,    6.3f,
      8.0f,    10.0f,   12.5f,   16.0f,   20.0f,   25.0f,   31.5f,   40.0f,   50.0f,
      63.0f,   80.0f,   100.0f,  125.0f,  160.0f,  200.0f,  250.0f,  315.0f,  400.0f,
      500.0f,  630.0f,  800.0f,  1000.0f, 1250.0f, 1600.0f, 2000.0f, 2500.0f, 3150.0f,
      4000.0f, 5000.0f, 6300.0f, 8000.0f, 10000.0f, 20000.0f};
  return t;
}

void derive_inputs(acum_inputs* in, float* dPsi_f) {
  // io.cu:73-83.  The parsed values are floats; the derived constants are kept in double.
  in->oM = std::fabs(in->oM);
  in->dTau = 2.0 * M_PI * in->dPsi_deg / in->oM / 360.0;
  in->dPsi = in->dPsi_deg * M_PI / 180.0;
  *dPsi_f = (float)((double)(float)in->dPsi_deg * M_PI / 180.0);
  const double d2r = M_PI / 180.0;
  in->vInf[0] = in->Minf * in->a0 * std::cos(in->xy_angle * d2r) * std::cos(in->xz_angle * d2r);
  in->vInf[1] = in->Minf * in->a0 * std::sin(in->xy_angle * d2r);
  in->vInf[2] = in->Minf * in->a0 * std::cos(in->xy_angle * d2r) * std::sin(in->xz_angle * d2r);
}

void layout_surfaces(acum_run* r) {
  const int64_t nS = r->nSurf, nT = r->nTime;
  r->XPtr.assign(nS, 0);
  r->NPtr.assign(nS, 0);
  r->VPtr.assign(nS, 0);
  r->PRPtr.assign(nS, 0);
  int64_t etot = 0;
  for (int64_t i = 0; i < nS; i++) {
    if (i > 0) {
      r->XPtr[i] = r->XPtr[i - 1] + 4 * r->E[i - 1] * nT;
      r->NPtr[i] = r->NPtr[i - 1] + 6 * r->E[i - 1] * nT;
      r->VPtr[i] = r->VPtr[i - 1] + 6 * r->E[i - 1] * nT;
      r->PRPtr[i] = r->PRPtr[i - 1] + 2 * r->E[i - 1] * nT;
    }
    etot += r->E[i];
  }
  fvec* arr[4] = {&r->X, &r->N, &r->V, &r->PR};
  const int64_t width[4] = {4, 6, 6, 2};
  for (int a = 0; a < 4; a++) {
    arr[a]->clear();
    arr[a]->resize((size_t)(etot * nT * width[a]));  // uninitialised
    float* p = arr[a]->data();
    const int64_t n = (int64_t)arr[a]->size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) p[i] = 0.0f;
  }
}

}  // namespace acum_host

using namespace acum_host;

namespace {

// Whole-file token reader.
struct Text {
  std::vector<char> buf;
  size_t pos = 0;
  bool load(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize((size_t)(n > 0 ? n : 0) + 1);
    size_t got = n > 0 ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
    buf[got] = '\0';
    buf.resize(got + 1);
    std::fclose(f);
    return true;
  }
  bool token(std::string* t) {
    while (pos + 1 < buf.size() && std::isspace((unsigned char)buf[pos])) pos++;
    if (pos + 1 >= buf.size()) return false;
    size_t s = pos;
    while (pos + 1 < buf.size() && !std::isspace((unsigned char)buf[pos])) pos++;
    t->assign(buf.data() + s, pos - s);
    return true;
  }
  bool f32(float* v) {
    while (pos + 1 < buf.size() && std::isspace((unsigned char)buf[pos])) pos++;
    if (pos + 1 >= buf.size()) return false;
    char* end = nullptr;
    errno = 0;
    *v = std::strtof(buf.data() + pos, &end);
    if (end == buf.data() + pos) return false;
    pos = (size_t)(end - buf.data());
    return true;
  }
  bool i32(int32_t* v) {
    while (pos + 1 < buf.size() && std::isspace((unsigned char)buf[pos])) pos++;
    if (pos + 1 >= buf.size()) return false;
    char* end = nullptr;
    long x = std::strtol(buf.data() + pos, &end, 10);
    if (end == buf.data() + pos) return false;
    pos = (size_t)(end - buf.data());
    *v = (int32_t)x;
    return true;
  }
  void skip_line() {
    while (pos + 1 < buf.size() && buf[pos] != '\n') pos++;
    if (pos + 1 < buf.size()) pos++;
  }
};

// ---- parallel number parsing for the large surface files -----------------------------------
// A decimal token (sign, digits, point, digits, exponent) with at most 19 significant digits and
// a power of ten within 1e+-22 is m * 10^e with m and 10^e exact doubles, so one IEEE multiply or
// divide gives the correctly rounded double (Clinger's fast path); rounding that double to float
// is the correctly rounded float unless the double sits exactly on a float rounding midpoint
// (the only case where the second rounding can differ).  Every other token goes to strtof, the
// function fscanf("%f") uses, so results are those of the reference's reader bit for bit.
inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }

bool fast_float(const char* p, const char* end, float* out) {
  static const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                    1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                    1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  bool neg = false;
  if (p < end && (*p == '-' || *p == '+')) neg = *p++ == '-';
  uint64_t m = 0;
  int sig = 0, e10 = 0;
  bool any = false;
  for (; p < end && *p >= '0' && *p <= '9'; p++) {
    any = true;
    if (m == 0 && *p == '0') continue;
    if (sig == 19) return false;
    m = m * 10 + (uint64_t)(*p - '0');
    sig++;
  }
  if (p < end && *p == '.') {
    for (p++; p < end && *p >= '0' && *p <= '9'; p++) {
      any = true;
      if (m == 0 && *p == '0') {
        e10--;
        continue;
      }
      if (sig == 19) return false;
      m = m * 10 + (uint64_t)(*p - '0');
      sig++;
      e10--;
    }
  }
  if (!any) return false;
  if (p < end && (*p == 'e' || *p == 'E')) {
    p++;
    bool eneg = false;
    if (p < end && (*p == '-' || *p == '+')) eneg = *p++ == '-';
    if (!(p < end && *p >= '0' && *p <= '9')) return false;
    int x = 0;
    for (; p < end && *p >= '0' && *p <= '9'; p++)
      if (x < 10000) x = x * 10 + (*p - '0');
    e10 += eneg ? -x : x;
  }
  if (p != end) return false;
  if (m == 0) {
    *out = neg ? -0.0f : 0.0f;
    return true;
  }
  if (m > (1ULL << 53) || e10 < -22 || e10 > 22) return false;
  double d = (double)m;
  d = e10 < 0 ? d / kPow10[-e10] : d * kPow10[e10];
  if (!(d >= 1.1754943508222875e-38 && d <= 3.4028234663852886e+38)) return false;  // float normal range
  uint64_t bits;
  std::memcpy(&bits, &d, sizeof(bits));
  if ((bits & ((1ULL << 29) - 1)) == (1ULL << 28)) return false;  // exactly on a float midpoint
  *out = (float)(neg ? -d : d);
  return true;
}

// Parses the first `count` whitespace-separated numbers of [b, e) (e points at a NUL) in parallel
// and hands number i to store(i, value).  Returns -1 on success, or the index of the first number
// that is missing or not a number.
template <typename Store>
int64_t parse_numbers(const char* b, const char* e, int64_t count, Store store) {
  const int T = std::max(1, omp_get_max_threads());
  const int64_t len = e - b;
  std::vector<const char*> cut(T + 1);
  cut[0] = b;
  cut[T] = e;
  for (int t = 1; t < T; t++) {  // chunk boundaries at whitespace
    const char* c = b + len * t / T;
    if (c < cut[t - 1]) c = cut[t - 1];
    while (c < e && !is_space(*c)) c++;
    cut[t] = c;
  }
  std::vector<int64_t> ntok(T + 1, 0);
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (int t = 0; t < T; t++) {
    int64_t n = 0;
    bool in = false;
    for (const char* c = cut[t]; c < cut[t + 1]; c++) {
      const bool sp = is_space(*c);
      if (!sp && !in) n++;
      in = !sp;
    }
    ntok[t + 1] = n;
  }
  for (int t = 0; t < T; t++) ntok[t + 1] += ntok[t];
  int64_t bad = ntok[T] < count ? ntok[T] : -1;
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (int t = 0; t < T; t++) {
    int64_t i = ntok[t];
    const char* c = cut[t];
    const char* stop = cut[t + 1];
    while (i < count) {
      while (c < stop && is_space(*c)) c++;
      if (c >= stop) break;
      const char* s0 = c;
      while (c < stop && !is_space(*c)) c++;
      float v;
      if (!fast_float(s0, c, &v)) {
        // strtof on a NUL-terminated copy (the text may be a file mapping without a NUL)
        char tok[128];
        const size_t tl = (size_t)(c - s0);
        char* endp = nullptr;
        if (tl < sizeof(tok)) {
          std::memcpy(tok, s0, tl);
          tok[tl] = '\0';
          v = std::strtof(tok, &endp);
        }
        if (tl >= sizeof(tok) || endp != tok + tl) {
#pragma omp critical
          if (bad < 0 || i < bad) bad = i;
          break;
        }
      }
      store(i, v);
      i++;
    }
  }
  return bad;
}

// Read-only mapping of a whole file.
struct MappedFile {
  const char* data = nullptr;
  size_t size = 0;
  bool open(const std::string& path) {
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (::fstat(fd, &st) != 0) {
      ::close(fd);
      return false;
    }
    size = (size_t)st.st_size;
    if (size > 0) {
      void* p = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
      if (p == MAP_FAILED) {
        ::close(fd);
        return false;
      }
      ::madvise(p, size, MADV_SEQUENTIAL);
      data = (const char*)p;
    }
    ::close(fd);
    return true;
  }
  ~MappedFile() {
    if (data) ::munmap((void*)data, size);
  }
};

std::string join(const std::string& dir, const std::string& name) {
  if (dir.empty()) return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}

bool exists(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0;
}

// read_inputs, io.cu:12-101.
int read_inputs(const std::string& path, acum_inputs* in, float* dPsi_f) {
  Text t;
  if (!t.load(path)) return fail(ACUM_ERR_IO, "No input file found: " + path);
  float a0 = 340.0f, rhoRef = 1.2f, pRef = 0.0f, dPsi = 1.0f, xy = 0.0f, xz = 0.0f, Mref = 0.08f;
  float oM = 0.0f, Minf = 0.0f, CFDscale = 0.0f, OBSscale = 0.0f;
  int32_t nTime = 360, periodic = 1, lowpass = 1, BB_noise = 0, impermeable = 0, nSurf = 0;
  std::string dat;
  int n = 0;
  while (t.token(&dat)) {
    n++;
    bool ok = true;
    if (dat == "a0") ok = t.f32(&a0);
    else if (dat == "impermeable") ok = t.i32(&impermeable);
    else if (dat == "rhoRef") ok = t.f32(&rhoRef);
    else if (dat == "pRef") ok = t.f32(&pRef);
    else if (dat == "oM") ok = t.f32(&oM);
    else if (dat == "nSurf") ok = t.i32(&nSurf);
    else if (dat == "nTime") ok = t.i32(&nTime);
    else if (dat == "dPsi") ok = t.f32(&dPsi);
    else if (dat == "Minf") ok = t.f32(&Minf);
    else if (dat == "Mref") ok = t.f32(&Mref);
    else if (dat == "xy_angle") ok = t.f32(&xy);
    else if (dat == "xz_angle") ok = t.f32(&xz);
    else if (dat == "CFDscale") ok = t.f32(&CFDscale);
    else if (dat == "OBSscale") ok = t.f32(&OBSscale);
    else if (dat == "periodic") ok = t.i32(&periodic);
    else if (dat == "lowpass") ok = t.i32(&lowpass);
    else if (dat == "BB_noise") ok = t.i32(&BB_noise);
    else if (dat == "ACUM-3_INPUTS" || dat == "***Environmental" || dat == "***Geometric" ||
             dat == "***Data")
      continue;
    else
      return fail(ACUM_ERR_ARG, "Input parameter not defined (" + std::to_string(n) + "): " + dat);
    if (!ok) return fail(ACUM_ERR_IO, "Malformed value for " + dat + " in " + path);
  }
  if (nSurf < 1) return fail(ACUM_ERR_ARG, "nSurf must be >= 1");
  if (nTime < 4) return fail(ACUM_ERR_ARG, "nTime must be >= 4");
  if (impermeable != 1)
    return fail(ACUM_ERR_ARG, "only impermeable 1 is supported (no kernel runs otherwise, main.cu:236)");
  in->a0 = a0;
  in->rhoRef = rhoRef;
  in->pRef = pRef;
  in->oM = oM;
  in->dPsi_deg = dPsi;
  in->Minf = Minf;
  in->Mref = Mref;
  in->xy_angle = xy;
  in->xz_angle = xz;
  in->CFDscale = CFDscale;
  in->OBSscale = OBSscale;
  in->nSurf = nSurf;
  in->nTime = nTime;
  in->periodic = periodic;
  in->lowpass = lowpass;
  in->impermeable = impermeable;
  in->BB_noise = BB_noise;
  if (oM == 0.0f) return fail(ACUM_ERR_ARG, "oM must be non-zero");
  derive_inputs(in, dPsi_f);
  return ACUM_OK;
}

// read_observers, io.cu:106-135 (the rotation there is the identity: cs = 1, ss = 0).
int read_observers(const std::string& path, acum_run* r) {
  Text t;
  if (!t.load(path)) return fail(ACUM_ERR_IO, "Unable to open observer file: " + path);
  int32_t n = 0;
  if (!t.i32(&n) || n < 1) return fail(ACUM_ERR_IO, "bad observer count in " + path);
  r->nObs = n;
  r->rObs.assign((size_t)n * 3, 0.0f);
  const float s = (float)r->in.OBSscale;
  const float cs = 1.0f, ss = 0.0f;
  for (int i = 0; i < n; i++) {
    float x, y, z;
    if (!t.f32(&x) || !t.f32(&y) || !t.f32(&z))
      return fail(ACUM_ERR_IO, "truncated observer file " + path);
    r->rObs[3 * i + 0] = x * s;
    r->rObs[3 * i + 1] = (y * cs - z * ss) * s;
    r->rObs[3 * i + 2] = (y * ss + z * cs) * s;
  }
  return ACUM_OK;
}

// Binary cache of a parsed surface file: surfaces/surface_k.acumc beside surface_k.dat.  A
// 64-byte header (magic, version, E, nt, and the size and modification time of the text file it
// was made from) and then, element-major, 8 floats per (element, step): x y z A nx ny nz CP --
// the run's own layout, so loading is a copy.  A cache whose header does not match the text
// file's current size and time (or is short or damaged) is ignored and the text is parsed.
constexpr char kCacheMagic[8] = {'A', 'C', 'U', 'M', 'S', 'R', 'F', '1'};
struct CacheHeader {
  char magic[8];
  uint32_t version, floats_per_record;
  int64_t E, nt, dat_size, dat_mtime_ns;
  int64_t reserved[2];
};
static_assert(sizeof(CacheHeader) == 64, "cache header is 64 bytes");

std::string cache_path(const std::string& dir, int64_t i) {
  return join(dir, "surfaces/surface_" + std::to_string(i + 1) + ".acumc");
}

bool dat_stamp(const std::string& path, int64_t* size, int64_t* mtime_ns) {
  struct stat st;
  if (::stat(path.c_str(), &st) != 0) return false;
  *size = (int64_t)st.st_size;
  *mtime_ns = (int64_t)st.st_mtim.tv_sec * 1000000000LL + st.st_mtim.tv_nsec;
  return true;
}

bool cache_disabled() {
  const char* v = std::getenv("ACUM_NO_SURFACE_CACHE");
  return v && v[0] && v[0] != '0';
}

// Maps surface i's cache when it matches the text file; the header is returned in *h.
bool open_cache(const std::string& dir, int64_t i, MappedFile* m, CacheHeader* h) {
  if (cache_disabled()) return false;
  const std::string dat = join(dir, "surfaces/surface_" + std::to_string(i + 1) + ".dat");
  int64_t size = 0, mt = 0;
  if (!dat_stamp(dat, &size, &mt)) return false;
  const std::string cp = cache_path(dir, i);
  if (!exists(cp) || !m->open(cp) || m->size < sizeof(CacheHeader)) return false;
  std::memcpy(h, m->data, sizeof(CacheHeader));
  if (std::memcmp(h->magic, kCacheMagic, 8) != 0 || h->version != 1 || h->floats_per_record != 8 ||
      h->E < 1 || h->nt < 1 || h->dat_size != size || h->dat_mtime_ns != mt)
    return false;
  const uint64_t need = sizeof(CacheHeader) + (uint64_t)h->E * (uint64_t)h->nt * 8 * sizeof(float);
  return m->size == need;
}

// read_surf, io.cu:206-333 (impermeable branch): nTime blocks of E lines x y z CP A nx ny nz.
// Each file is mapped, its numbers parsed in parallel into a time-major buffer, which is then
// transposed in tiles into the element-major arrays; a file with a matching binary cache is
// copied from the cache instead.
int read_surf(const std::string& dir, acum_run* r) {
  const int64_t nS = r->in.nSurf, nT = r->in.nTime;
  r->nSurf = nS;
  r->nTime = nT;
  r->E.assign(nS, 0);
  std::vector<MappedFile> files(nS), caches(nS);
  std::vector<size_t> body(nS, 0);
  std::vector<char> cached(nS, 0);
  for (int64_t i = 0; i < nS; i++) {
    const std::string p = join(dir, "surfaces/surface_" + std::to_string(i + 1) + ".dat");
    CacheHeader ch;
    if (open_cache(dir, i, &caches[i], &ch) && ch.nt == nT) {
      r->E[i] = ch.E;
      cached[i] = 1;
      continue;
    }
    if (!files[i].open(p)) return fail(ACUM_ERR_IO, "missing surface file " + p);
    // header: the first two integers "E nt" (fscanf "%d %d\n", io.cu:235)
    const char* d = files[i].data;
    const size_t n = files[i].size;
    size_t q = 0;
    int64_t hv[2] = {0, 0};
    for (int h = 0; h < 2; h++) {
      while (q < n && is_space(d[q])) q++;
      const size_t t0 = q;
      while (q < n && !is_space(d[q])) q++;
      const std::string tok(d + t0, q - t0);
      char* endp = nullptr;
      hv[h] = std::strtoll(tok.c_str(), &endp, 10);
      if (tok.empty() || *endp != '\0') return fail(ACUM_ERR_IO, "bad header in " + p);
    }
    const int64_t e = hv[0], nt = hv[1];
    if (e < 1) return fail(ACUM_ERR_IO, "bad header in " + p);
    if (nt != nT) std::fprintf(stderr, "\tInconsistent number of steps in %s\n", p.c_str());
    r->E[i] = e;
    body[i] = q;
  }
  layout_surfaces(r);
  int64_t Emax = 0;
  for (int64_t i = 0; i < nS; i++) Emax = std::max(Emax, r->E[i]);
  fvec tmp((size_t)(nT * Emax * 8));  // time-major numbers of one file, first touched by the parse
  for (int64_t i = 0; i < nS; i++) {
    const int64_t E = r->E[i];
    if (cached[i]) {
      const float* c = (const float*)(caches[i].data + sizeof(CacheHeader));
      float* X = r->X.data() + r->XPtr[i];
      float* N = r->N.data() + r->NPtr[i];
      float* P = r->PR.data() + r->PRPtr[i];
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < E * nT; q++) {
        const float* t = c + q * 8;
        X[q * 4 + 0] = t[0];
        X[q * 4 + 1] = t[1];
        X[q * 4 + 2] = t[2];
        X[q * 4 + 3] = t[3];
        N[q * 6 + 0] = t[4];
        N[q * 6 + 1] = t[5];
        N[q * 6 + 2] = t[6];
        P[q * 2] = t[7];
      }
      continue;
    }
    const int64_t count = nT * E * 8;
    const char* b = files[i].data + body[i];
    const char* e = files[i].data + files[i].size;
    const int64_t bad = parse_numbers(b, e, count, [&](int64_t n, float v) { tmp[(size_t)n] = v; });
    if (bad >= 0)
      return fail(ACUM_ERR_IO, "truncated or malformed surface file surface_" + std::to_string(i + 1) +
                                   ".dat at step " + std::to_string(bad / (8 * E)) + " element " +
                                   std::to_string((bad / 8) % E));
    // tmp[k][j][x y z CP A nx ny nz] -> X[j][k][x y z A], N[j][k][n], PR[j][k][CP]
    float* X = r->X.data() + r->XPtr[i];
    float* N = r->N.data() + r->NPtr[i];
    float* P = r->PR.data() + r->PRPtr[i];
    const int64_t TB = 64;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int64_t j0 = 0; j0 < E; j0 += TB)
      for (int64_t k0 = 0; k0 < nT; k0 += TB) {
        const int64_t j1 = std::min(E, j0 + TB), k1 = std::min(nT, k0 + TB);
        for (int64_t j = j0; j < j1; j++)
          for (int64_t k = k0; k < k1; k++) {
            const float* t = &tmp[(size_t)((k * E + j) * 8)];
            const int64_t q = j * nT + k;
            X[q * 4 + 0] = t[0];
            X[q * 4 + 1] = t[1];
            X[q * 4 + 2] = t[2];
            X[q * 4 + 3] = t[4];
            P[q * 2] = t[3];
            N[q * 6 + 0] = t[5];
            N[q * 6 + 1] = t[6];
            N[q * 6 + 2] = t[7];
          }
      }
  }
  return ACUM_OK;
}

// read_BB_inputs, io.cu:337-700.
int read_bb(const std::string& dir, acum_run* r) {
  Text t;
  const std::string path = join(dir, "input_BB");
  if (!t.load(path)) return fail(ACUM_ERR_IO, "No broadband input file found");
  const int nSurf = r->in.nSurf;
  int32_t nSurf_BB = nSurf, t_del = 0;
  r->span.assign(nSurf, 1.0f);
  r->omega.assign(nSurf, 1.0f);
  r->ccw.assign(nSurf, 1);
  r->psi_offset.assign(nSurf, 0.0f);
  r->trip.assign(nSurf, 0);
  r->rotate.assign(3 * nSurf, 0.0f);
  r->translate.assign(3 * nSurf, 0.0f);
  std::string dat;
  int n = 0, is = 0;
  while (t.token(&dat)) {
    n++;
    const std::string surf = "***Surface-" + std::to_string(is + 1);
    bool ok = true;
    if (dat == "ACUM-3_Broadband_Inputs") continue;
    if (dat == "nSurf_BB") {
      ok = t.i32(&nSurf_BB);
      if (ok && nSurf_BB > nSurf)
        return fail(ACUM_ERR_ARG, "Number of surfaces considered for broadband noise cannot be "
                                  "higher than the available surfaces");
      if (ok && nSurf_BB < 1) return fail(ACUM_ERR_ARG, "nSurf_BB must be >= 1");
    } else if (dat == "t_del") ok = t.i32(&t_del);
    else if (dat == surf) is += 1;
    else if (is >= 1 && is <= nSurf && dat == "span(m)") ok = t.f32(&r->span[is - 1]);
    else if (is >= 1 && is <= nSurf && dat == "omega(rad/s)") ok = t.f32(&r->omega[is - 1]);
    else if (is >= 1 && is <= nSurf && dat == "ccw") ok = t.i32(&r->ccw[is - 1]);
    else if (is >= 1 && is <= nSurf && dat == "psi_offset") ok = t.f32(&r->psi_offset[is - 1]);
    else if (is >= 1 && is <= nSurf && dat == "trip") ok = t.i32(&r->trip[is - 1]);
    else if (dat == "Initial_Position") continue;
    else if (is >= 1 && is <= nSurf && dat == "rotate") {
      for (int i = 0; i < 3 && ok; i++) ok = t.f32(&r->rotate[3 * (is - 1) + i]);
    } else if (is >= 1 && is <= nSurf && dat == "translate") {
      for (int i = 0; i < 3 && ok; i++) ok = t.f32(&r->translate[3 * (is - 1) + i]);
      if (ok && is == nSurf_BB) break;
    } else
      return fail(ACUM_ERR_ARG, "Input parameter not defined (" + std::to_string(n) + "): " + dat);
    if (!ok) return fail(ACUM_ERR_IO, "Malformed value for " + dat + " in " + path);
  }
  r->nSurf_BB = nSurf_BB;
  r->t_del = t_del;
  const int nB = nSurf_BB;
  const int64_t nT = r->in.nTime;
  // BB_surface files (io.cu:443-521)
  r->nsect.assign(nB, 0);
  r->xsPtr.assign(nB, 0);
  r->dsPtr.assign(nB, 0);
  std::vector<Text> sf(nB);
  int64_t tot = 0;
  for (int i = 0; i < nB; i++) {
    const std::string p = join(dir, "BB_inputs/BB_surface_" + std::to_string(i + 1) + ".dat");
    if (!sf[i].load(p)) return fail(ACUM_ERR_IO, "missing " + p);
    int32_t nt = 0, ns = 0;
    if (!sf[i].i32(&nt) || !sf[i].i32(&ns) || ns < 2) return fail(ACUM_ERR_IO, "bad header in " + p);
    if (nt != nT) std::fprintf(stderr, "\tInconsistent number of steps in %s\n", p.c_str());
    r->nsect[i] = ns;
    if (i > 0) {
      r->xsPtr[i] = r->xsPtr[i - 1] + (int64_t)r->nsect[i - 1] * 4;
      r->dsPtr[i] = r->dsPtr[i - 1] + (int64_t)r->nsect[i - 1] * 7 * nT;
    }
    tot += ns;
  }
  r->xyzcR.assign((size_t)tot * 4, 0.0f);
  r->bbdata.assign((size_t)(tot * nT * 7), 0.0f);
  r->bbdata_raw.assign((size_t)(tot * nT * 7), 0.0f);
  for (int i = 0; i < nB; i++) {
    Text& q = sf[i];
    const int ns = r->nsect[i];
    for (int j = 0; j < 4 * ns; j++)
      if (!q.f32(&r->xyzcR[r->xsPtr[i] + j]))
        return fail(ACUM_ERR_IO, "truncated BB_surface_" + std::to_string(i + 1) + ".dat");
    for (int64_t j = 0; j < nT; j++) {
      for (int k = 0; k < ns; k++) {
        float* b = &r->bbdata[r->dsPtr[i] + j * ns * 7 + k * 7];
        for (int w = 0; w < 7; w++)
          if (!q.f32(&b[w]))
            return fail(ACUM_ERR_IO, "truncated BB_surface_" + std::to_string(i + 1) + ".dat");
        for (int w = 0; w < 7; w++) r->bbdata_raw[r->dsPtr[i] + j * ns * 7 + k * 7 + w] = b[w];
        if (r->omega[i] != 0.0f) {  // io.cu:509-512
          const double rr = (double)r->xyzcR[r->xsPtr[i] + 4 * k];
          b[4] = (float)((double)b[4] / (rr * rr));
          b[5] = (float)((double)b[5] / (rr * rr));
        }
      }
    }
  }
  // aerodynamic tables (io.cu:523-687)
  r->nsect_t.assign(nB, 0);
  r->nmach_t.assign(nB, 0);
  r->naoa_t.assign(nB, 0);
  r->nre_t.assign(nB, 0);
  r->aPtr.assign(nB, 0);
  r->mPtr.assign(nB, 0);
  r->rPtr.assign(nB, 0);
  r->sPtr.assign(nB, 0);
  r->dPtr.assign(nB, 0);
  std::vector<Text> tf(nB);
  int64_t tm = 0, ta = 0, tr = 0, ts = 0, td = 0;
  for (int i = 0; i < nB; i++) {
    const std::string p = join(dir, "BB_inputs/BB_table_" + std::to_string(i + 1) + ".dat");
    if (!tf[i].load(p)) return fail(ACUM_ERR_IO, "missing " + p);
    tf[i].skip_line();
    int32_t a, b, c, d;
    if (!tf[i].i32(&a) || !tf[i].i32(&b) || !tf[i].i32(&c) || !tf[i].i32(&d) || a < 1 || b < 2 ||
        c < 1 || d < 1)
      return fail(ACUM_ERR_IO, "bad header in " + p);
    r->nsect_t[i] = a;
    r->naoa_t[i] = b;
    r->nmach_t[i] = c;
    r->nre_t[i] = d;
    if (i > 0) {
      r->aPtr[i] = r->aPtr[i - 1] + r->naoa_t[i - 1];
      r->sPtr[i] = r->sPtr[i - 1] + r->nsect_t[i - 1];
      r->mPtr[i] = r->mPtr[i - 1] + r->nmach_t[i - 1];
      r->rPtr[i] = r->rPtr[i - 1] + r->nre_t[i - 1];
      r->dPtr[i] = r->dPtr[i - 1] + (int64_t)r->nsect_t[i - 1] * r->naoa_t[i - 1] *
                                        r->nmach_t[i - 1] * r->nre_t[i - 1] * 7;
    }
    tm += c;
    ta += b;
    tr += d;
    ts += a;
    td += (int64_t)a * b * c * d * 7;
  }
  r->mach_t.assign(tm, 0.0f);
  r->sect_t.assign(ts, 0.0f);
  r->re_t.assign(tr, 0.0f);
  r->aoa_t.assign(ta, 0.0f);
  r->bbdata_t.assign(td, 0.0f);
  for (int i = 0; i < nB; i++) {
    Text& q = tf[i];
    const std::string nm = "BB_table_" + std::to_string(i + 1) + ".dat";
    const int ns = r->nsect_t[i], na = r->naoa_t[i], nmach = r->nmach_t[i], nr = r->nre_t[i];
    for (int j = 0; j < ns; j++) {
      if (!q.f32(&r->sect_t[r->sPtr[i] + j])) return fail(ACUM_ERR_IO, "truncated " + nm);
      for (int k = 0; k < nmach; k++)
        for (int l = 0; l < nr; l++)
          for (int m = 0; m < na; m++) {
            const int64_t lp = r->dPtr[i] + (int64_t)j * nmach * nr * na * 7 +
                               (int64_t)k * nr * na * 7 + (int64_t)l * na * 7 + m * 7;
            bool ok = q.f32(&r->aoa_t[r->aPtr[i] + m]) && q.f32(&r->mach_t[r->mPtr[i] + k]);
            for (int w = 0; w < 7 && ok; w++) ok = q.f32(&r->bbdata_t[lp + w]);
            ok = ok && q.f32(&r->re_t[r->rPtr[i] + l]);
            if (!ok) return fail(ACUM_ERR_IO, "truncated " + nm);
          }
    }
    // sanity checks of io.cu:628-686
    for (int j = 0; j < ns; j++) {
      const float v = r->sect_t[r->sPtr[i] + j];
      if (v < 0.0f || v > 1.0f) return fail(ACUM_ERR_ARG, "Check span section of " + nm);
      for (int k = 0; k < ns; k++)
        if (j != k && v == r->sect_t[r->sPtr[i] + k]) return fail(ACUM_ERR_ARG, "Check span sections of " + nm);
    }
    for (int j = 0; j < nmach; j++) {
      const float v = r->mach_t[r->mPtr[i] + j];
      if (v < 0.0f) return fail(ACUM_ERR_ARG, "Check Mach number of " + nm);
      for (int k = 0; k < nmach; k++)
        if (j != k && v == r->mach_t[r->mPtr[i] + k]) return fail(ACUM_ERR_ARG, "Check Mach numbers of " + nm);
    }
    for (int j = 0; j < nr; j++) {
      const float v = r->re_t[r->rPtr[i] + j];
      if (v < 0.0f) return fail(ACUM_ERR_ARG, "Check Re number of " + nm);
      for (int k = 0; k < nr; k++)
        if (j != k && v == r->re_t[r->rPtr[i] + k]) return fail(ACUM_ERR_ARG, "Check Re numbers of " + nm);
    }
    for (int j = 0; j < na; j++)
      for (int k = 0; k < na; k++)
        if (j != k && r->aoa_t[r->aPtr[i] + j] == r->aoa_t[r->aPtr[i] + k])
          return fail(ACUM_ERR_ARG, "Check AoAs of " + nm);
  }
  // bands
  const std::string bp = join(dir, "BB_inputs/BB_bands.dat");
  if (exists(bp)) {
    Text q;
    int32_t nb = 0;
    if (!q.load(bp) || !q.i32(&nb) || nb < 1) return fail(ACUM_ERR_IO, "bad " + bp);
    r->ff.assign(nb, 0.0f);
    for (int i = 0; i < nb; i++)
      if (!q.f32(&r->ff[i]) || !(r->ff[i] > 0.0f)) return fail(ACUM_ERR_IO, "bad " + bp);
  } else {
    r->ff = reference_bands();
  }
  r->nf2 = (int32_t)r->ff.size();
  r->has_bb = true;
  return ACUM_OK;
}

// Appends v as printf's "%<width>.<prec>E" / "%<width>.<prec>e" would write it.  std::to_chars
// in scientific form with a precision gives the same correctly rounded digits as printf; the
// exponent letter and the width are adjusted here; NaN and infinity go through snprintf.
void append_sci(std::string& s, double v, int prec, bool upper, int width) {
  char buf[48];
  int n;
  if (std::isfinite(v)) {
    const auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific, prec);
    n = (int)(r.ptr - buf);
    if (upper)
      for (int i = 0; i < n; i++)
        if (buf[i] == 'e') buf[i] = 'E';
  } else {
    n = std::snprintf(buf, sizeof(buf), upper ? "%.*E" : "%.*e", prec, v);
  }
  for (int i = n; i < width; i++) s.push_back(' ');
  s.append(buf, (size_t)n);
}

int write_time_history(const acum_run* r, const std::string& dir, const char* name, const float* p) {
  const std::string path = join(dir, name);
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return fail(ACUM_ERR_IO, "cannot write " + path);
  std::fprintf(f, "# ntime nobs\n");
  std::fprintf(f, "%d %d\n", (int)r->nTime, (int)r->nObs);
  // rows " %14.8E" (t dPsi, radians) then nObs x " %14.8E" (io.cu:143-157), formatted in parallel
  const int64_t nT = r->nTime, nObs = r->nObs;
  const int64_t block = 64;
  std::vector<std::string> rows((size_t)block);
  for (int64_t i0 = 0; i0 < nT; i0 += block) {
    const int64_t i1 = std::min(nT, i0 + block);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = i0; i < i1; i++) {
      std::string& s = rows[(size_t)(i - i0)];
      s.clear();
      s.reserve((size_t)(nObs + 1) * 16 + 1);
      s.push_back(' ');
      append_sci(s, (double)((float)i * r->dPsi_f), 8, true, 14);
      for (int64_t j = 0; j < nObs; j++) {
        s.push_back(' ');
        append_sci(s, (double)p[i * nObs + j], 8, true, 14);
      }
      s.push_back('\n');
    }
    for (int64_t i = i0; i < i1; i++) {
      const std::string& s = rows[(size_t)(i - i0)];
      if (std::fwrite(s.data(), 1, s.size(), f) != s.size()) {
        std::fclose(f);
        return fail(ACUM_ERR_IO, "write failed: " + path);
      }
    }
  }
  if (std::fclose(f) != 0) return fail(ACUM_ERR_IO, "write failed: " + path);
  return ACUM_OK;
}

}  // namespace

extern "C" {

const char* acum_host_error(void) { return g_err.c_str(); }

int acum_run_load(const char* dir, acum_run** out) {
  if (!out) return fail(ACUM_ERR_ARG, "null output");
  *out = nullptr;
  const std::string d = dir ? dir : "";
  acum_run* r = new acum_run();
  int rc = read_inputs(join(d, "input"), &r->in, &r->dPsi_f);
  if (rc == ACUM_OK && r->in.BB_noise == 1) rc = read_bb(d, r);
  if (rc == ACUM_OK) rc = read_observers(join(d, "obs.dat"), r);
  if (rc == ACUM_OK) rc = read_surf(d, r);
  if (rc != ACUM_OK) {
    delete r;
    return rc;
  }
  *out = r;
  return ACUM_OK;
}

void acum_run_free(acum_run* r) { delete r; }

int acum_run_inputs(const acum_run* r, acum_inputs* in) {
  if (!r || !in) return fail(ACUM_ERR_ARG, "null argument");
  *in = r->in;
  return ACUM_OK;
}

int acum_run_tonal_case(const acum_run* r, int32_t precision, acum_tonal_case* c) {
  if (!r || !c) return fail(ACUM_ERR_ARG, "null argument");
  std::memset(c, 0, sizeof(*c));
  c->nSurf = r->nSurf;
  c->nTime = r->nTime;
  c->nObs = r->nObs;
  c->E = r->E.data();
  c->X = r->X.data();
  c->N = r->N.data();
  c->V = r->V.data();
  c->PR = r->PR.data();
  c->XPtr = r->XPtr.data();
  c->NPtr = r->NPtr.data();
  c->VPtr = r->VPtr.data();
  c->PRPtr = r->PRPtr.data();
  c->rObs = r->rObs.data();
  c->dPsi = r->in.dPsi;
  c->dTau = r->in.dTau;
  c->a0 = r->in.a0;
  c->rhoRef = r->in.rhoRef;
  c->oM = r->in.oM;
  for (int i = 0; i < 3; i++) c->vInf[i] = r->in.vInf[i];
  c->precision = precision;
  if (!r->preprocessed) return fail(ACUM_ERR_ARG, "run not preprocessed (call acum_run_preprocess)");
  return ACUM_OK;
}

int acum_run_bpm_case(const acum_run* r, acum_bpm_case* c) {
  if (!r || !c) return fail(ACUM_ERR_ARG, "null argument");
  if (!r->has_bb) return fail(ACUM_ERR_ARG, "run has no broadband inputs (BB_noise 0)");
  std::memset(c, 0, sizeof(*c));
  c->nSurf_BB = r->nSurf_BB;
  c->nTime = (int32_t)r->nTime;
  c->nObs = (int32_t)r->nObs;
  c->nf2 = r->nf2;
  c->t_del = r->t_del;
  c->dTau = r->in.dTau;
  c->a0 = r->in.a0;
  for (int i = 0; i < 3; i++) c->vInf[i] = r->in.vInf[i];
  c->rObs = r->rObs.data();
  c->ff = r->ff.data();
  c->span = r->span.data();
  c->omega = r->omega.data();
  c->psi_offset = r->psi_offset.data();
  c->ccw = r->ccw.data();
  c->trip = r->trip.data();
  c->rotate = r->rotate.data();
  c->translate = r->translate.data();
  c->nsect = r->nsect.data();
  c->xsPtr = r->xsPtr.data();
  c->dsPtr = r->dsPtr.data();
  c->xyzcR = r->xyzcR.data();
  c->bbdata = r->bbdata.data();
  c->nsect_t = r->nsect_t.data();
  c->nmach_t = r->nmach_t.data();
  c->naoa_t = r->naoa_t.data();
  c->nre_t = r->nre_t.data();
  c->aPtr = r->aPtr.data();
  c->mPtr = r->mPtr.data();
  c->rPtr = r->rPtr.data();
  c->sPtr = r->sPtr.data();
  c->dPtr = r->dPtr.data();
  c->sect_t = r->sect_t.data();
  c->mach_t = r->mach_t.data();
  c->aoa_t = r->aoa_t.data();
  c->re_t = r->re_t.data();
  c->bbdata_t = r->bbdata_t.data();
  return ACUM_OK;
}

int acum_run_write_tonal(const acum_run* r, const char* dir, const float* pT, const float* pL,
                         const float* pA) {
  if (!r || !pT || !pL || !pA) return fail(ACUM_ERR_ARG, "null argument");
  const std::string d = dir ? dir : "";
  int rc = write_time_history(r, d, "pT.out", pT);
  if (rc == ACUM_OK) rc = write_time_history(r, d, "pL.out", pL);
  if (rc == ACUM_OK) rc = write_time_history(r, d, "pA.out", pA);
  return rc;
}

int acum_run_write_bb(const acum_run* r, const char* dir, const float* BBoaspl, const float* BBspl) {
  if (!r || !BBoaspl || !BBspl) return fail(ACUM_ERR_ARG, "null argument");
  const std::string d = dir ? dir : "";
  FILE* f = std::fopen(join(d, "bb_oaspl.out").c_str(), "w");
  if (!f) return fail(ACUM_ERR_IO, "cannot write bb_oaspl.out");
  std::fprintf(f, "nobs: %d\n", (int)r->nObs);
  for (int64_t i = 0; i < r->nObs; i++) std::fprintf(f, " %14.8E\n", (double)BBoaspl[i]);
  bool bad = std::ferror(f) != 0;  // read before the close
  if ((std::fclose(f) != 0) || bad) return fail(ACUM_ERR_IO, "write failed: bb_oaspl.out");
  f = std::fopen(join(d, "bb_spl.dat").c_str(), "w");
  if (!f) return fail(ACUM_ERR_IO, "cannot write bb_spl.dat");
  const int nf2 = r->nf2;
  std::fprintf(f, " # nfrequency nobs\n    %d      %d\n", nf2, (int)r->nObs);
  static const char* names[7] = {"TOTAL", "TBL_TE_P", "TBL_TE_S", "TBL_SS", "TEB_VS", "LBL-VS", "TVF"};
  static const int comp[7] = {0, 3, 2, 1, 5, 4, 6};  // io.cu:737-792
  for (int z = 0; z < 7; z++) {
    std::fprintf(f, " ZONE T=\"%s\"\n", names[z]);
    std::vector<std::string> rows((size_t)nf2);
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < nf2; j++) {  // "   %e" then nObs x "  %e" (io.cu:738-791)
      std::string& s = rows[(size_t)j];
      s.reserve((size_t)(r->nObs + 1) * 16 + 1);
      s.append("   ");
      append_sci(s, (double)r->ff[j], 6, false, 0);
      for (int64_t i = 0; i < r->nObs; i++) {
        s.append("  ");
        append_sci(s, (double)BBspl[i * nf2 * 7 + j * 7 + comp[z]], 6, false, 0);
      }
      s.push_back('\n');
    }
    for (const std::string& s : rows) std::fwrite(s.data(), 1, s.size(), f);
  }
  bad = std::ferror(f) != 0;
  if ((std::fclose(f) != 0) || bad) return fail(ACUM_ERR_IO, "write failed: bb_spl.dat");
  return ACUM_OK;
}

int acum_run_save_cache(const acum_run* r, const char* dir) {
  if (!r || !dir) return fail(ACUM_ERR_ARG, "null argument");
  if (r->preprocessed) return fail(ACUM_ERR_ARG, "run already preprocessed; cache the run as read");
  const std::string d = dir;
  const int64_t nT = r->nTime;
  for (int64_t i = 0; i < r->nSurf; i++) {
    const std::string dat = join(d, "surfaces/surface_" + std::to_string(i + 1) + ".dat");
    CacheHeader h{};
    std::memcpy(h.magic, kCacheMagic, 8);
    h.version = 1;
    h.floats_per_record = 8;
    h.E = r->E[i];
    h.nt = nT;
    if (!dat_stamp(dat, &h.dat_size, &h.dat_mtime_ns)) return fail(ACUM_ERR_IO, "missing surface file " + dat);
    const int64_t n = h.E * nT;
    std::vector<float> buf((size_t)n * 8);
    const float* X = r->X.data() + r->XPtr[i];
    const float* N = r->N.data() + r->NPtr[i];
    const float* P = r->PR.data() + r->PRPtr[i];
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; q++) {
      float* t = &buf[(size_t)q * 8];
      t[0] = X[q * 4 + 0];
      t[1] = X[q * 4 + 1];
      t[2] = X[q * 4 + 2];
      t[3] = X[q * 4 + 3];
      t[4] = N[q * 6 + 0];
      t[5] = N[q * 6 + 1];
      t[6] = N[q * 6 + 2];
      t[7] = P[q * 2];
    }
    // written beside the final name and renamed, so a reader never sees a partial cache
    const std::string fin = cache_path(d, i), tmp = fin + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return fail(ACUM_ERR_IO, "cannot write " + tmp);
    const bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
                    std::fwrite(buf.data(), sizeof(float), buf.size(), f) == buf.size();
    if (std::fclose(f) != 0 || !ok || std::rename(tmp.c_str(), fin.c_str()) != 0) {
      std::remove(tmp.c_str());
      return fail(ACUM_ERR_IO, "cannot write " + fin);
    }
  }
  return ACUM_OK;
}

int acum_run_surface_raw(const acum_run* r, int32_t isurf, float* out) {
  if (!r || !out || isurf < 0 || isurf >= r->nSurf) return fail(ACUM_ERR_ARG, "bad argument");
  if (r->preprocessed) return fail(ACUM_ERR_ARG, "run already preprocessed");
  const int64_t nT = r->nTime, E = r->E[isurf];
  for (int64_t k = 0; k < nT; k++)
    for (int64_t j = 0; j < E; j++) {
      const float* Xh = &r->X[r->XPtr[isurf] + (j * nT + k) * 4];
      const float* Nh = &r->N[r->NPtr[isurf] + (j * nT + k) * 6];
      const float* Ph = &r->PR[r->PRPtr[isurf] + (j * nT + k) * 2];
      float* o = out + (k * E + j) * 8;
      o[0] = Xh[0];
      o[1] = Xh[1];
      o[2] = Xh[2];
      o[3] = Ph[0];
      o[4] = Xh[3];
      o[5] = Nh[0];
      o[6] = Nh[1];
      o[7] = Nh[2];
    }
  return ACUM_OK;
}

}  // extern "C"
