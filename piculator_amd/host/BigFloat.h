// BigFloat with the reference's interface (reference BigFloat.h:8-46), whose
// big multiply runs on the MI355X through the picugpu C ABI.
//
// Everything except the limb product is host code with the reference's
// semantics: sign-magnitude, exponent in 32-bit limbs, precision p meaning
// "p + 2 limbs" (BigFloat.h:6), low limbs chopped without rounding.
// MulBatch additionally lets a caller hand many independent products to the
// GPU at once (one tree level of the binary splitting), producing exactly the
// BigFloats that the same sequence of BigFloat::mul calls would produce.
#pragma once
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#define YCL_BIGFLOAT_EXTRA_PRECISION 2

namespace piculator {

// Which engine BigFloat::mul uses.  GPU: the C ABI (pcg_mul); the default.
void set_mul_engine_flags(uint32_t flags);
uint32_t mul_engine_flags();
uint64_t big_mul_count();
void count_big_muls(uint64_t n);

class MulBatch;

class BigFloat {
 public:
  BigFloat(BigFloat&& x) noexcept;
  BigFloat(const BigFloat& x);
  BigFloat& operator=(BigFloat&& x) noexcept;
  BigFloat& operator=(const BigFloat& x);

  BigFloat();
  BigFloat(uint32_t x, bool sign = true);

  std::string to_string_hex(size_t digits = 0) const;
  size_t get_precision() const { return L; }
  int64_t get_exponent() const { return exp; }
  uint32_t word_at(int64_t mag) const;

  void negate();
  BigFloat mul(uint32_t x) const;
  BigFloat add(const BigFloat& x, size_t p = 0) const;
  BigFloat sub(const BigFloat& x, size_t p = 0) const;
  BigFloat mul(const BigFloat& x, size_t p = 0, int threads = 1) const;
  // The products of the Newton steps that multiply r0 twice (the device
  // type keeps r0's transform between them); plain products here.
  BigFloat mul_keep(const BigFloat& x, size_t p) const { return mul(x, p); }
  BigFloat mul_reuse(const BigFloat& x, size_t p) const { return mul(x, p); }
  BigFloat rcp(size_t p, int threads = 1) const;
  BigFloat div(const BigFloat& x, size_t p, int threads = 1) const;

  // Raw access for bindings and tests (not in the reference interface).
  bool get_sign() const { return sign; }
  const uint32_t* limbs() const { return T.get(); }
  static BigFloat from_parts(bool sign, int64_t exp, const uint32_t* limbs, size_t n);

  // Hooks of the Newton iterations (Newton.h): exponent shift, the top three
  // limbs as a double (e: exponent of the lowest of them), two-limb seed.
  void shift_exp(int64_t d) { exp += d; }
  double top_double(int64_t* e) const;
  static BigFloat seed(double val, int64_t e, bool sign);

 private:
  bool sign;     // true = positive or zero
  int64_t exp;   // exponent in limbs
  size_t L;      // number of limbs
  std::unique_ptr<uint32_t[]> T;

  int ucmp(const BigFloat& x) const;
  BigFloat uadd(const BigFloat& x, size_t p) const;
  BigFloat usub(const BigFloat& x, size_t p) const;

  friend BigFloat invsqrt(uint32_t x, size_t p, int threads);
  friend class MulBatch;
};

BigFloat invsqrt(uint32_t x, size_t p, int threads = 1);

// A set of independent products  dst_i = a_i.mul(b_i, p_i)  run together.
// The operands must stay alive until run() returns.
class MulBatch {
 public:
  void add(const BigFloat& a, const BigFloat& b, size_t p, BigFloat* dst);
  size_t size() const { return items_.size(); }
  // Executes all products (grouped per size class on the GPU); `plain`
  // runs one GPU launch sequence per product instead.
  void run(bool plain = false);

 private:
  struct Item {
    const uint32_t* at;
    size_t al;
    const uint32_t* bt;
    size_t bl;
    BigFloat* dst;
  };
  std::vector<Item> items_;
};

}  // namespace piculator
