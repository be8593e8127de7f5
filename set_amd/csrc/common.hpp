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
// FP64: hardware estimate refined by two Newton steps (about 1 ulp);
// FP32: the hardware instruction (1 ulp).
__device__ __forceinline__ double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(e, r, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(e, r, r);
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// sqrt of a positive normal number: rsq estimate + Goldschmidt refinement.
__device__ __forceinline__ double fsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = 0.5 * y;
  double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}
// FP32: the hardware instruction (1 ulp; sqrtf itself expands to a
// correctly rounded sequence of about a dozen instructions).
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// 1/sqrt of a positive normal number: estimate + two Newton steps (FP64).
__device__ __forceinline__ double frsq(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double hx = 0.5 * x;
  double e = __builtin_fma(-hx * y, y, 0.5);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-hx * y, y, 0.5);
  return __builtin_fma(y, e, y);
}
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

}  // namespace setmhd
