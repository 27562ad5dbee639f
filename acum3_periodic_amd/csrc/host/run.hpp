// Host-side state of one run directory: the reference's global variables (variables.cu:1-64,
// variables.cuh:1-60) gathered into one object, with 64-bit offsets.
#pragma once
#include <cstdint>
#include <memory>
#include <utility>
#include <string>
#include <vector>
#include "../../../include/acum.h"

// Allocator that leaves new elements uninitialised, so the large element arrays can be filled
// (first-touched) in parallel instead of being zeroed by one thread.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <typename U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
using fvec = std::vector<float, NoInitAlloc<float>>;

struct acum_run {
  acum_inputs in{};
  float dPsi_f = 0.0f;  // radians, float as the reference stores it (io.cu:75); used by writers
  // observers (io.cu:106-135)
  int64_t nObs = 0;
  std::vector<float> rObs;
  // tonal surfaces, reference AoS layout (io.cu:206-333)
  int64_t nSurf = 0, nTime = 0;
  std::vector<int64_t> E, XPtr, NPtr, VPtr, PRPtr;
  fvec X, N, V, PR;  // filled in parallel (layout_surfaces zeroes them with all threads)
  bool preprocessed = false;
  // broadband (io.cu:337-700)
  bool has_bb = false;
  int32_t nSurf_BB = 0, t_del = 0, nf2 = 0;
  std::vector<float> span, omega, psi_offset, rotate, translate, xyzcR, bbdata;
  std::vector<float> bbdata_raw;  // bbdata as written in the files (cn, cc not yet divided)
  std::vector<int32_t> ccw, trip, nsect, nsect_t, nmach_t, naoa_t, nre_t;
  std::vector<int64_t> xsPtr, dsPtr, aPtr, mPtr, rPtr, sPtr, dPtr;
  std::vector<float> sect_t, mach_t, aoa_t, re_t, bbdata_t, ff;
};

namespace acum_host {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
// Derived constants of read_inputs (io.cu:70-83), computed in double from the parsed floats.
void derive_inputs(acum_inputs* in, float* dPsi_f);
// Reference 1/3-octave table (io.cu:693-699).
const std::vector<float>& reference_bands();
// Allocates the AoS arrays of a run from nSurf, nTime and E (io.cu:243-261).
void layout_surfaces(acum_run* r);
// Normalised 13-point Gaussian of the low-pass filter (mathfunc.cu:350-370), in FP64.
void gauss_weights13(double* g);
}  // namespace acum_host
