// Plain FP64 tonal path: the reference's kernel structure (findPressureTimeHistory,
// core.cu:18-259, launched at main.cu:194-236) computed in FP64 on gfx950.
//
// This is the measurement baseline for the restructured paths of tonal.hip, not a path anyone
// should run for speed.  It keeps what the restructuring removes:
//   * one thread per (source step, observer) pair and a stride over elements, with the
//     reference's block of 8 steps x 4 observers x 8 elements and its grid caps (512 blocks per
//     dimension, main.cu:194-233);
//   * three retarded-time solves per eval: the step's own and those of the steps before and
//     after it (core.cu:84, 189, 193), each with its square root and division;
//   * the integrand written term by term with its divisions (core.cu:86-154);
//   * one FP64 atomic add per (bin, output) into pT, pL and pA (core.cu:202-226), so
//     contributions to a bin land in whatever order the hardware serves them.
// Terms the impermeable surface makes identically zero are left out (U = V makes the momentum
// terms Mo of core.cu:143-150 vanish, and the dotted normal velocities are zero at
// core.cu:117-118), and the bin of an observer time is a floor in bin units instead of the
// binary search of core.cu:290-355: both favour this baseline.  Gauge and units follow the
// reference: observer time in bins is ((k dTau + r / a0) oM) / dPsi, bin b is stored at b mod nT.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "dev_types.hpp"

namespace acum {

struct PlainArgs {
  const float* X;    // [nE][nT][4]  x, y, z, dS
  const float* N;    // [nE][nT][6]  n, dn/dt
  const float* V;    // [nE][nT][6]  v, dv/dt
  const float* PR;   // [nE][nT][2]  P, dP/dt
  const float* obs;  // [nObs][3]
  int64_t nE;
  int32_t nT, nObs;
  double dTau, dPsi, a0, rhoRef, oM;
  double vx, vy, vz;
  double* pT;  // [nT][nObs] FP64 accumulators, zeroed before the launch
  double* pL;
  double* pA;
  uint32_t* flags;
};

// Distance a0 t that sound travels from src to obs through the medium moving at vInf, t the
// positive root of a t^2 + b t + c = 0 (acous_dist, core.cu:261-279).
__device__ double plain_distance(const PlainArgs& a, double sx, double sy, double sz, double ox,
                                 double oy, double oz) {
  const double dx = ox - sx, dy = oy - sy, dz = oz - sz;
  const double qa = a.vx * a.vx + a.vy * a.vy + a.vz * a.vz - a.a0 * a.a0;
  const double qb = -2.0 * (dx * a.vx + dy * a.vy + dz * a.vz);
  const double qc = dx * dx + dy * dy + dz * dz;
  const double disc = qb * qb - 4.0 * qa * qc;
  double t = (-qb - sqrt(disc)) / (2.0 * qa);
  if (t < 0.0) t = (-qb + sqrt(disc)) / (2.0 * qa);
  return t * a.a0;
}

// Observer time in bins of the source position at step j of element e.
__device__ double plain_time(const PlainArgs& a, int64_t e, int j, double ox, double oy, double oz) {
  const float* x = a.X + (e * a.nT + j) * 4;
  const double r = plain_distance(a, x[0], x[1], x[2], ox, oy, oz);
  return ((j * a.dTau + r / a.a0) * a.oM) / a.dPsi;
}

__device__ void plain_add(const PlainArgs& a, double b, int io, double th, double lo, double w) {
  int64_t ib = (int64_t)b % a.nT;
  if (ib < 0) ib += a.nT;
  const int64_t i = ib * a.nObs + io;
  unsafeAtomicAdd(&a.pT[i], th * w);
  unsafeAtomicAdd(&a.pL[i], lo * w);
  unsafeAtomicAdd(&a.pA[i], (th + lo) * w);
}

__global__ __launch_bounds__(256) void tonal_plain_kernel(PlainArgs a) {
  const double D1 = 1.0 / (4.0 * M_PI * a.a0);
  bool bad = false;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < a.nT; k += gridDim.x * blockDim.x) {
    for (int io = blockIdx.y * blockDim.y + threadIdx.y; io < a.nObs; io += gridDim.y * blockDim.y) {
      const double ox = a.obs[3 * io], oy = a.obs[3 * io + 1], oz = a.obs[3 * io + 2];
      for (int64_t e = blockIdx.z * blockDim.z + threadIdx.z; e < a.nE; e += (int64_t)gridDim.z * blockDim.z) {
        const int64_t q = e * a.nT + k;
        const float* x = a.X + q * 4;
        const float* n = a.N + q * 6;
        const float* v = a.V + q * 6;
        const double P = a.PR[2 * q], Pd = a.PR[2 * q + 1];
        const double sx = x[0], sy = x[1], sz = x[2], ds = x[3];
        const double Vx = v[0], Vy = v[1], Vz = v[2], Ax = v[3], Ay = v[4], Az = v[5];
        const double Nx = n[0], Ny = n[1], Nz = n[2], NDx = n[3], NDy = n[4], NDz = n[5];
        // radiation vector and Doppler factors (core.cu:84-95)
        const double r = plain_distance(a, sx, sy, sz, ox, oy, oz);
        const double rx = ox - sx - a.vx * r / a.a0;
        const double ry = oy - sy - a.vy * r / a.a0;
        const double rz = oz - sz - a.vz * r / a.a0;
        const double hx = rx / r, hy = ry / r, hz = rz / r;
        const double Mr = (Vx * hx + Vy * hy + Vz * hz) / a.a0;
        const double M = sqrt(Vx * Vx + Vy * Vy + Vz * Vz) / a.a0;
        const double Mdr = (Ax * hx + Ay * hy + Az * hz) / a.a0;
        const double om = 1.0 - Mr;
        const double T1 = D1 / (r * om * om);
        const double T2 = T1 / om;
        const double T3 = T1 / r;
        const double T4 = T3 / om;
        // thickness (core.cu:97-128): rho = 1, U = V, dotted normal velocities zero
        const double Vn = Vx * Nx + Vy * Ny + Vz * Nz;
        const double Man = a.rhoRef * Vn;
        const double pth = Man * Mdr * T2 * a.a0 * ds + a.a0 * Man * (Mr - M * M) * T4 * a.a0 * ds;
        // loading (core.cu:130-152)
        const double Lx = P * Nx, Ly = P * Ny, Lz = P * Nz;
        const double LDx = Pd * Nx + P * NDx, LDy = Pd * Ny + P * NDy, LDz = Pd * Nz + P * NDz;
        const double Ldr = LDx * hx + LDy * hy + LDz * hz;
        const double Lr = Lx * hx + Ly * hy + Lz * hz;
        const double Lm = (Lx * Vx + Ly * Vy + Lz * Vz) / a.a0;
        const double plo = Ldr * T1 * ds + Lr * Mdr * T2 * ds + (Lr - Lm) * T3 * a.a0 * ds +
                           Lr * (Mr - M * M) * T4 * a.a0 * ds;
        // observer times of this step and of its neighbours, wrapped over the period
        // (core.cu:156, 172-194)
        const double tc = ((k * a.dTau + r / a.a0) * a.oM) / a.dPsi;
        const int km = k > 0 ? k - 1 : a.nT - 1, kp = k + 1 < a.nT ? k + 1 : 0;
        const double tp = plain_time(a, e, km, ox, oy, oz) - (k > 0 ? 0.0 : (double)a.nT);
        const double tn = plain_time(a, e, kp, ox, oy, oz) + (k + 1 < a.nT ? 0.0 : (double)a.nT);
        if (!(r > 0.0) || !(om > 0.0) || !(tc > tp) || !(tn > tc) || !isfinite(pth) || !isfinite(plo) ||
            !isfinite(tn) || !isfinite(tp)) {
          bad = true;
          continue;
        }
        // tent (core.cu:196-231): rising bins in (tp, tc], falling bins in (tc, tn); at most a
        // period of bins on each side
        const double ic = floor(tc);
        int guard = 0;
        for (double b = ic; b > tp && guard < a.nT; b -= 1.0, guard++)
          plain_add(a, b, io, pth, plo, (b - tp) / (tc - tp));
        guard = 0;
        for (double b = ic + 1.0; b < tn && guard < a.nT; b += 1.0, guard++)
          plain_add(a, b, io, pth, plo, (tn - b) / (tn - tc));
      }
    }
  }
  if (bad) atomicOr(a.flags, kFlagNumeric);
}

__global__ void plain_to_float_kernel(const double* __restrict__ src, float* __restrict__ dst, int64_t n,
                                      uint32_t* flags) {
  bool bad = false;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float f = (float)src[i];
    bad |= !isfinite(f);
    dst[i] = f;
  }
  if (bad) atomicOr(flags, kFlagNumeric);
}

hipError_t launch_tonal_plain(hipStream_t s, const float* X, const float* N, const float* V, const float* PR,
                              const float* obs, int64_t nE, int nT, int nObs, double dTau, double dPsi,
                              double a0, double rhoRef, double oM, const double* vInf, double* acc,
                              uint32_t* flags) {
  PlainArgs a;
  a.X = X;
  a.N = N;
  a.V = V;
  a.PR = PR;
  a.obs = obs;
  a.nE = nE;
  a.nT = nT;
  a.nObs = nObs;
  a.dTau = dTau;
  a.dPsi = dPsi;
  a.a0 = a0;
  a.rhoRef = rhoRef;
  a.oM = oM;
  a.vx = vInf[0];
  a.vy = vInf[1];
  a.vz = vInf[2];
  const int64_t n = (int64_t)nT * nObs;
  a.pT = acc;
  a.pL = acc + n;
  a.pA = acc + 2 * n;
  a.flags = flags;
  // the reference's launch shape (main.cu:194-233): 8 steps x 4 observers x 8 elements per
  // block, at most 512 blocks along each dimension, 512 along the elements
  const int tx = nT < 8 ? nT : 8, ty = nObs < 4 ? nObs : 4;
  const int bx = (nT + tx - 1) / tx < 512 ? (nT + tx - 1) / tx : 512;
  const int by = (nObs + ty - 1) / ty < 512 ? (nObs + ty - 1) / ty : 512;
  hipLaunchKernelGGL(tonal_plain_kernel, dim3(bx, by, 512), dim3(tx, ty, 8), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_plain_to_float(hipStream_t s, const double* acc, float* const* out, int64_t n, uint32_t* flags) {
  const int grid = (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
  for (int i = 0; i < 3; i++)
    hipLaunchKernelGGL(plain_to_float_kernel, dim3(grid), dim3(256), 0, s, acc + i * n, out[i], n, flags);
  return hipGetLastError();
}

}  // namespace acum
