// common.hpp -- shared device helpers for the gfx950 MHD kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace setmhd {

constexpr double kCpi4 = 12.56637061435917295384;  // Storage_module.f90:8
constexpr double kGamma = 5.0 / 3.0;                // Storage_module.f90:9
constexpr double kDtReset = 1.0e15;                 // cuf_kernel.cuf:790
constexpr int kNpar = 16;                           // Storage_module.f90:13

// Fortran default-REAL literal promoted to REAL(8)
#define SETMHD_F4(x) ((double)(float)(x))

// device error word bits; the host maps them to SETMHD_ERR_*
enum : int { kErrRho = 1, kErrRegion = 2, kErrFaceRho = 4 };

// flags, same values as SETMHD_F_*
enum : int { kZeroBW = 1, kAxisym = 2, kSourceMK = 4, kHelioHacks = 8 };

// The three dt slots: step n uses slot n%3, min-reduces the face CFL
// candidates into slot (n+1)%3 and re-arms slot (n+2)%3 with 1e15.
struct DtRing {
  double* slot;  // [3] device
  double* t;     // [1] device: simulated time
};

// Atomic minimum of a non-negative double through its IEEE bit pattern
// (order of non-negative doubles = order of their bits as unsigned).
__device__ inline void atomic_min_pos(double* addr, double v) {
  atomicMin(reinterpret_cast<unsigned long long*>(addr),
            static_cast<unsigned long long>(__double_as_longlong(v)));
}

// wave64 reductions through cross-lane shuffles
__device__ inline double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---- fast reciprocal / square root (structured fast path) ----------------
// FP64: the hardware estimate (relative error e below 2^-22) refined by one
// third-order step from the series in e, 1/(1-e) = 1 + e + e^2 + ... and
// (1-e)^(-1/2) = 1 + e/2 + 3e^2/8 + ...: the neglected e^3 term is below
// 2^-66, so the result is within about 1 ulp, as after two Newton steps, with
// one instruction less for 1/x and two less for 1/sqrt and sqrt.  An FP64
// estimate issues at a quarter of the FMA rate, so each one saved counts four.
// FP32: the hardware instruction (1 ulp).
__device__ __forceinline__ double frcp(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  const double e = __builtin_fma(-x, r, 1.0);  // 1 - x r, exact before its rounding
  return __builtin_fma(r, __builtin_fma(e, e, e), r);
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// 1/a and 1/b from one reciprocal, r = 1/(a b): 1/a = b r, 1/b = a r (within
// about 2 ulp; a b must stay in range, as the HLLD's wave-speed differences
// do).  FP32 keeps two hardware reciprocals.
__device__ __forceinline__ void frcp2(double a, double b, double& ia, double& ib) {
  const double r = frcp(a * b);
  ia = b * r;
  ib = a * r;
}
__device__ __forceinline__ void frcp2(float a, float b, float& ia, float& ib) {
  ia = frcp(a);
  ib = frcp(b);
}

// 1/a, 1/b and 1/c from one reciprocal, r = 1/(a b c): 1/a = b c r,
// 1/b = a c r, 1/c = a b r (within about 3 ulp; a b c must stay in range, as
// the HLLD's wave-speed differences and switch denominators do).  FP32 keeps
// three hardware reciprocals.
__device__ __forceinline__ void frcp3(double a, double b, double c, double& ia, double& ib, double& ic) {
  const double ab = a * b;
  const double r = frcp(ab * c);
  const double rc = r * c;  // 1/(a b)
  ia = b * rc;
  ib = a * rc;
  ic = ab * r;
}
__device__ __forceinline__ void frcp3(float a, float b, float c, float& ia, float& ib, float& ic) {
  ia = frcp(a);
  ib = frcp(b);
  ic = frcp(c);
}

// sqrt of a positive normal number: g = x y ~ sqrt(x) from the estimate y of
// 1/sqrt(x), e = 1 - g y, sqrt(x) = g (1 + e/2 + 3e^2/8 + ...).
__device__ __forceinline__ double fsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double g = x * y;
  const double e = __builtin_fma(-g, y, 1.0);
  return __builtin_fma(g * e, __builtin_fma(e, 0.375, 0.5), g);
}
// FP32: the hardware instruction (1 ulp; sqrtf itself expands to a
// correctly rounded sequence of about a dozen instructions).
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// 1/sqrt of a positive normal number: e = 1 - x y^2, y (1 + e/2 + 3e^2/8).
__device__ __forceinline__ double frsq(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-x * y, y, 1.0);
  return __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
}
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

}  // namespace setmhd
