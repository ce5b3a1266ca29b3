// Newton iterations of the reference BigFloat (BigFloat.cpp:445-567), written
// once for both float types of this package: BigFloat (host limbs) and
// DeviceFloat (limbs in device memory).  F provides mul(F, p), mul_keep and
// mul_reuse (mul, keeping the operand's transform between two products by
// it where the type can), mul(uint32_t),
// sub(F, p), F(uint32_t), shift_exp(d), top_double(&e) (the top three limbs
// as a double and the exponent of the lowest of them), get_sign() and
// F::seed(val, e, sign) (the two-limb float of the p == 0 branches).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <cmath>

namespace piculator {

// Precision of the recursive call: p / 2 + 1, with 1 -> 0 and 2 -> 1
// (BigFloat.cpp:467-470, 528-531).
inline size_t half_precision(size_t p) {
  if (p == 1) return 0;
  if (p == 2) return 1;
  return p / 2 + 1;
}

// 1 / a to precision p (BigFloat.cpp:445-511):  r1 = r0 - (r0 * a - 1) * r0.
template <class F>
F newton_rcp(const F& a, size_t p) {
  if (a.get_precision() == 0) throw "Divide by Zero";
  if (p == 0) {
    int64_t e = 0;
    const double val = 1. / a.top_double(&e);
    return F::seed(val, -e, a.get_sign());
  }
  const F r0 = newton_rcp(a, half_precision(p));
  // r0's transform from the first product serves the second
  const F t = a.mul_keep(r0, p).sub(F(1), p);
  return r0.sub(t.mul_reuse(r0, p), p);
}

// 1 / sqrt(x) to precision p (BigFloat.cpp:516-567):
// r1 = r0 - (r0^2 x - 1) / 2 * r0.
template <class F>
F newton_invsqrt(uint32_t x, size_t p) {
  if (x == 0) throw "Divide by Zero";
  if (p == 0) return F::seed(1. / std::sqrt((double)x), 0, true);
  const F r0 = newton_invsqrt<F>(x, half_precision(p));
  F t = r0.mul_keep(r0, p);  // r0's transform serves the last product
  // BigFloat.cpp:562 multiplies by BigFloat(x) with p = 0: no chop, the
  // exact product of |t| + 1 limbs less one leading zero limb, the sign and
  // exponent of t.  mul(uint32_t) (BigFloat.cpp:156-183) gives that same
  // value in one pass instead of a transform of |t| limbs.
  t = t.mul(x);
  t = t.sub(F(1), p);
  t = t.mul(2147483648u);
  t.shift_exp(-1);
  t = t.mul_reuse(r0, p);
  return r0.sub(t, p);
}

}  // namespace piculator
