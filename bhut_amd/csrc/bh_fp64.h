// bh_fp64.h -- FP64 square root and division in a middle range, bit for bit
// those of sqrt() and '/' (used by the plain walk; checked against them by
// fp64_check.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace bh {

// Correctly rounded FP64 square root and division for operands in a middle
// range, written out as the operation sequence hipcc emits for sqrt() and
// '/' on gfx950 without its range handling: the hardware estimate (v_rsq_f64,
// v_rcp_f64) refined by Newton-Raphson steps and a final fma correction
// (Markstein's method).  In the middle range that range handling is the
// identity (the square root's scale factor is 2^0 and its zero / infinity
// select is not taken; v_div_scale_f64 returns its operand unscaled with VCC
// clear, so v_div_fmas_f64 is a plain fma, and v_div_fixup_f64 returns its
// first operand), so the results are bit for bit those of sqrt() and '/'.
//   sqrt_mid(x):       x in [2^-700, 2^700]
//   div_mid(num, den): num in [2^-300, 2^300], den in [2^-350, 2^350]
//   recip_mid(den):    den in [2^-350, 2^350]  (1.0 / den)
__device__ __forceinline__ double sqrt_mid(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

__device__ __forceinline__ double recip_refined(double den) {
  double y = __builtin_amdgcn_rcp(den);
  double e = fma(-den, y, 1.0);
  y = fma(y, e, y);
  e = fma(-den, y, 1.0);
  return fma(y, e, y);
}

__device__ __forceinline__ double div_mid(double num, double den) {
  const double y = recip_refined(den);
  const double q = num * y;
  const double rem = fma(-den, q, num);
  return fma(rem, y, q);
}

__device__ __forceinline__ double recip_mid(double den) {
  const double y = recip_refined(den);  // the quotient estimate 1.0 * y is y
  const double rem = fma(-den, y, 1.0);
  return fma(rem, y, y);
}

}  // namespace bh
