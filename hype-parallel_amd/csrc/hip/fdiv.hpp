// fdiv.hpp -- correctly rounded FP64 division with a shared reciprocal.
//
// For a / b the compiler emits the gfx9 IEEE division sequence
//
//   s_b = v_div_scale(b, b, a);  s_a = v_div_scale(a, b, a)   (range scaling)
//   y   = v_rcp_f64(s_b)
//   e   = fma(-s_b, y, 1);  y = fma(y, e, y)                   (Newton, twice)
//   e   = fma(-s_b, y, 1);  y = fma(y, e, y)
//   q   = s_a * y;  r = fma(-s_b, q, s_a)
//   d   = v_div_fmas(r, y, q)       = fma(r, y, q) unless scaled
//   res = v_div_fixup(d, b, a)      = d unless an operand or the result is
//                                     zero, infinite, NaN or out of range
//
// When |a| and |b| lie in [2^-380, 2^380] (or a is zero), the scaling steps
// return their operands unchanged: the exponent difference stays below 768,
// neither operand nor the quotient is denormal, the numerator's exponent is
// far above the near-underflow threshold and 1/b is normal.  In that case the
// fixup only matters for a zero numerator (signed zero).  Then the
// sequence without those steps yields the same bits.  y depends on b alone, so one
// refined reciprocal serves every quotient with that denominator: each
// further quotient costs a multiply and two FMAs instead of the whole
// sequence.
//
// Signed zeros: with the remainder taken as r' = fma(b, q, -a) = -r (exact)
// and the result fma(-r', y, q), a zero numerator yields the zero the fixup
// would return (sign of a times sign of b) for every b > 0.  All callers
// divide by positive quantities.
//
// Callers check the range condition (usually once for a group of
// divisions) and fall back to plain `/` when it fails.
#pragma once
#ifndef HYPE_KNS
#define HYPE_KNS hype  // kernel namespace; the contracted build uses its own
#endif
#include <hip/hip_runtime.h>

namespace HYPE_KNS {

constexpr double FDIV_HI = 0x1p380;
constexpr double FDIV_LO = 0x1p-380;

// refined reciprocal of b, as in the division sequence
__device__ __forceinline__ double recip_nr(double b) {
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  return fma(y, e, y);
}

// a / b for b > 0 in range, given y = recip_nr(b); bitwise equal to a / b
__device__ __forceinline__ double div_y(double a, double b, double y) {
  const double q = a * y;
  const double rn = fma(b, q, -a);
  return fma(-rn, y, q);
}

// 1 / b for b > 0 in range, given y = recip_nr(b); bitwise equal to 1.0 / b
__device__ __forceinline__ double inv_y(double b, double y) {
  const double rn = fma(b, y, -1.0);
  return fma(-rn, y, y);
}

// |x| in [2^-380, 2^380]; false for zero, NaN and infinities
__device__ __forceinline__ bool fdiv_in_range(double x) {
  const double a = fabs(x);
  return a >= FDIV_LO && a <= FDIV_HI;
}

}  // namespace HYPE_KNS
