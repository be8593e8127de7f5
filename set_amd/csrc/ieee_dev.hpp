// ieee_dev.hpp -- correctly rounded FP64 division and square root for the
// drop-in kernels, written as the instruction sequences hipcc emits for IEEE
// a / b and sqrt(x) on gfx950, minus the operand scaling those sequences
// apply at the ends of the exponent range:
//
//   a / b   v_div_scale x2, v_rcp_f64, two Newton steps on the reciprocal,
//           q = a y, r = fma(-b, q, a), fma(r, y, q), v_div_fixup.  Without
//           the two v_div_scale (identities unless an operand or the
//           quotient is near the ends of the range) the result has the same
//           bits as IEEE division whenever b is normal with |b| < 2^1022, a
//           is zero or |a| >= 2^-969, the exponents of a and b differ by
//           less than 768 and a / b is normal.  Zeros, infinities and NaNs
//           go through v_div_fixup as in the compiler's sequence.
//   sqrt(x) v_rsq_f64 and the Goldschmidt iteration with two corrections;
//           without the scaling of x < 2^-767 the result has the same bits
//           as IEEE sqrt for x >= 2^-767 and for +-0 and +inf.
//
// The solver's quantities (densities, pressures, speeds and fields of order
// 1e-10 to 1e10, geometry of order 1e-3 to 1e3) lie far inside these ranges.
// The reciprocal of the divisor is the only part of a / b that depends on b
// alone, so a divisor used several times is inverted once (recip_of) and
// each quotient (quot) costs a multiply, two FMAs and the fixup instead of
// the whole sequence.  The fast build (SETMHD_DROPIN_FAST) takes the same
// reciprocal, within about 1 ulp of 1/b, and multiplies by it.
#pragma once
#include "common.hpp"

namespace setmhd {

struct Recip {
  double d, y;  // the divisor and its refined reciprocal
};
__device__ __forceinline__ Recip recip_of(double d) {
  const double y0 = __builtin_amdgcn_rcp(d);
  const double e0 = __builtin_fma(-d, y0, 1.0);
  const double y1 = __builtin_fma(y0, e0, y0);
  const double e1 = __builtin_fma(-d, y1, 1.0);
  return Recip{d, __builtin_fma(y1, e1, y1)};
}
#if SETMHD_DROPIN_FAST
__device__ __forceinline__ double quot(double a, const Recip& r) { return a * r.y; }
#else
__device__ __forceinline__ double quot(double a, const Recip& r) {
  const double q = a * r.y;
  const double rem = __builtin_fma(-r.d, q, a);
  return __builtin_amdgcn_div_fixup(__builtin_fma(rem, r.y, q), r.d, a);
}
#endif
__device__ __forceinline__ double quot(double a, double b) { return quot(a, recip_of(b)); }

__device__ __forceinline__ double sqrt_rn(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  h = __builtin_fma(h, r, h);
  g = __builtin_fma(g, r, g);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  // +-0 and +inf are their own square roots (class mask: -0, +0, +inf)
  return __builtin_amdgcn_class(x, 0x260) ? x : g;
}

}  // namespace setmhd
