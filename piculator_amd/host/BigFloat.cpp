// Host BigFloat whose big multiply runs on the GPU (see BigFloat.h).
// Semantics follow the reference BigFloat.cpp function by function; the
// citations give the reference lines each method reproduces.
#include "BigFloat.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "../../include/picugpu.h"
#include "Newton.h"

namespace piculator {

namespace {
uint32_t g_flags = PCG_ENGINE_AUTO;
std::atomic<uint64_t> g_mul_count{0};

// Error convention of the reference: throw string literals.
[[noreturn]] void throw_status(int rc) {
  if (rc == PCG_ERR_TOO_LARGE) throw "FFT length is too small.";
  throw pcg_last_error_string();
}

void gpu_product(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out) {
  int rc = pcg_mul(a, na, b, nb, out, g_flags);
  if (rc == PCG_ERR_NOT_INIT) {
    rc = pcg_init(0);
    if (rc == PCG_OK) rc = pcg_mul(a, na, b, nb, out, g_flags);
  }
  if (rc != PCG_OK) throw_status(rc);
}
}  // namespace

void set_mul_engine_flags(uint32_t flags) { g_flags = flags; }
uint32_t mul_engine_flags() { return g_flags; }
uint64_t big_mul_count() { return g_mul_count.load(); }
void count_big_muls(uint64_t n) { g_mul_count.fetch_add(n, std::memory_order_relaxed); }

// ------------------------------------------------ construction (BigFloat.cpp:14-79)
BigFloat::BigFloat(BigFloat&& x) noexcept : sign(x.sign), exp(x.exp), L(x.L), T(std::move(x.T)) {
  x.sign = true;
  x.exp = 0;
  x.L = 0;
}
BigFloat::BigFloat(const BigFloat& x) : sign(x.sign), exp(x.exp), L(x.L) {
  T.reset(new uint32_t[L ? L : 1]);
  if (L) std::memcpy(T.get(), x.T.get(), L * sizeof(uint32_t));
}
BigFloat& BigFloat::operator=(BigFloat&& x) noexcept {
  sign = x.sign;
  exp = x.exp;
  L = x.L;
  T = std::move(x.T);
  x.sign = true;
  x.exp = 0;
  x.L = 0;
  return *this;
}
BigFloat& BigFloat::operator=(const BigFloat& x) {
  if (this == &x) return *this;
  std::unique_ptr<uint32_t[]> t(new uint32_t[x.L ? x.L : 1]);
  if (x.L) std::memcpy(t.get(), x.T.get(), x.L * sizeof(uint32_t));
  sign = x.sign;
  exp = x.exp;
  L = x.L;
  T = std::move(t);
  return *this;
}
BigFloat::BigFloat() : sign(true), exp(0), L(0) {}
BigFloat::BigFloat(uint32_t x, bool sign_) : sign(true), exp(0), L(1) {
  if (x == 0) {
    L = 0;
    return;
  }
  sign = sign_;
  T.reset(new uint32_t[1]);
  T[0] = x;
}
BigFloat BigFloat::from_parts(bool s, int64_t e, const uint32_t* limbs, size_t n) {
  BigFloat z;
  if (n == 0) return z;
  z.sign = s;
  z.exp = e;
  z.L = n;
  z.T.reset(new uint32_t[n]);
  std::memcpy(z.T.get(), limbs, n * sizeof(uint32_t));
  return z;
}

// ------------------------------------------------ strings (BigFloat.cpp:95-106)
std::string BigFloat::to_string_hex(size_t digits) const {
  // top limb in decimal, '.', then 8 lowercase hex digits per lower limb
  const size_t nl = (digits + 7) / 8;
  if (L == 0 || nl > L - 1) throw "Limit Exceeded";
  std::string s = std::to_string(T[L - 1]) + ".";
  s.reserve(s.size() + 8 * nl);
  static const char* hx = "0123456789abcdef";
  char buf[8];
  for (size_t i = 0; i < nl; ++i) {
    uint32_t w = T[L - 2 - i];
    for (int d = 7; d >= 0; --d) {
      buf[d] = hx[w & 15];
      w >>= 4;
    }
    s.append(buf, 8);
  }
  return s.substr(0, 2 + digits);
}

// ------------------------------------------------ comparisons (BigFloat.cpp:109-147)
uint32_t BigFloat::word_at(int64_t mag) const {
  if (mag < exp) return 0;
  if (mag >= exp + (int64_t)L) return 0;
  return T[(size_t)(mag - exp)];
}
int BigFloat::ucmp(const BigFloat& x) const {
  const int64_t magA = exp + (int64_t)L, magB = x.exp + (int64_t)x.L;
  if (magA != magB) return magA > magB ? 1 : -1;
  for (int64_t mag = magA; mag >= exp || mag >= x.exp; --mag) {
    const uint32_t wa = word_at(mag), wb = x.word_at(mag);
    if (wa != wb) return wa > wb ? 1 : -1;
  }
  return 0;
}

// ------------------------------------------------ arithmetic (BigFloat.cpp:150-325)
void BigFloat::negate() {
  if (L != 0) sign = !sign;
}
BigFloat BigFloat::mul(uint32_t x) const {
  if (L == 0 || x == 0) return BigFloat();
  BigFloat z;
  z.sign = sign;
  z.exp = exp;
  z.L = L;
  z.T.reset(new uint32_t[L + 1]);
  uint64_t carry = 0;
  for (size_t c = 0; c < L; ++c) {
    carry += (uint64_t)T[c] * x;
    z.T[c] = (uint32_t)carry;
    carry >>= 32;
  }
  if (carry) z.T[z.L++] = (uint32_t)carry;
  return z;
}

namespace {
// Window [bot, top) of an addition at precision p (BigFloat.cpp:188-209,240-260).
void add_window(int64_t ea, size_t la, int64_t eb, size_t lb, size_t p, int64_t& bot, int64_t& top) {
  top = std::max(ea + (int64_t)la, eb + (int64_t)lb);
  bot = std::min(ea, eb);
  const int64_t len = top - bot;
  const int64_t keep = p == 0 ? len : (int64_t)(p + YCL_BIGFLOAT_EXTRA_PRECISION);
  if (len > keep) bot = top - keep;
}
}  // namespace

BigFloat BigFloat::uadd(const BigFloat& x, size_t p) const {
  int64_t bot, top;
  add_window(exp, L, x.exp, x.L, p, bot, top);
  BigFloat z;
  z.sign = sign;
  z.exp = bot;
  z.L = (uint32_t)(top - bot);
  z.T.reset(new uint32_t[z.L + 1]);
  uint64_t carry = 0;
  for (size_t c = 0; bot < top; ++bot, ++c) {
    carry += (uint64_t)word_at(bot) + x.word_at(bot);
    z.T[c] = (uint32_t)carry;
    carry >>= 32;
  }
  if (carry) z.T[z.L++] = 1;
  return z;
}
BigFloat BigFloat::usub(const BigFloat& x, size_t p) const {
  int64_t bot, top;
  add_window(exp, L, x.exp, x.L, p, bot, top);
  BigFloat z;
  z.sign = sign;
  z.exp = bot;
  z.L = (uint32_t)(top - bot);
  z.T.reset(new uint32_t[z.L ? z.L : 1]);
  uint64_t borrow = 0;
  for (size_t c = 0; bot < top; ++bot, ++c) {
    const uint64_t d = (uint64_t)word_at(bot) - x.word_at(bot) - borrow;
    z.T[c] = (uint32_t)d;
    borrow = (d >> 63) & 1;
  }
  while (z.L > 0 && z.T[z.L - 1] == 0) --z.L;
  if (z.L == 0) {
    z.exp = 0;
    z.sign = true;
    z.T.reset();
  }
  return z;
}
BigFloat BigFloat::add(const BigFloat& x, size_t p) const {
  if (sign == x.sign) return uadd(x, p);
  if (ucmp(x) > 0) return usub(x, p);
  return x.usub(*this, p);
}
BigFloat BigFloat::sub(const BigFloat& x, size_t p) const {
  if (sign != x.sign) return uadd(x, p);
  if (ucmp(x) > 0) return usub(x, p);
  BigFloat z = x.usub(*this, p);
  z.negate();
  return z;
}

// ------------------------------------------------ the big multiply (BigFloat.cpp:326-444)
namespace {
// Metadata of BigFloat::mul: operand windows after the precision chop.
struct MulShape {
  const uint32_t* at;
  size_t al;
  int64_t ae;
  const uint32_t* bt;
  size_t bl;
  int64_t be;
};
MulShape mul_shape(const uint32_t* at, size_t al, int64_t ae, const uint32_t* bt, size_t bl,
                   int64_t be, size_t p) {
  p = p == 0 ? al + bl : p + YCL_BIGFLOAT_EXTRA_PRECISION;
  if (al > p) {
    at += al - p;
    ae += (int64_t)(al - p);
    al = p;
  }
  if (bl > p) {
    bt += bl - p;
    be += (int64_t)(bl - p);
    bl = p;
  }
  return MulShape{at, al, ae, bt, bl, be};
}
}  // namespace

BigFloat BigFloat::mul(const BigFloat& x, size_t p, int threads) const {
  (void)threads;  // the GPU engine does not take a thread budget
  if (L == 0 || x.L == 0) return BigFloat();
  const MulShape s = mul_shape(T.get(), L, exp, x.T.get(), x.L, x.exp, p);
  BigFloat z;
  z.sign = sign == x.sign;
  z.exp = s.ae + s.be;
  z.L = s.al + s.bl;
  z.T.reset(new uint32_t[z.L]);
  gpu_product(s.at, s.al, s.bt, s.bl, z.T.get());
  g_mul_count.fetch_add(1, std::memory_order_relaxed);
  if (z.T[z.L - 1] == 0) --z.L;  // one leading-zero trim, BigFloat.cpp:411-412
  return z;
}

void MulBatch::add(const BigFloat& a, const BigFloat& b, size_t p, BigFloat* dst) {
  if (a.L == 0 || b.L == 0) {
    *dst = BigFloat();
    return;
  }
  const MulShape s = mul_shape(a.T.get(), a.L, a.exp, b.T.get(), b.L, b.exp, p);
  BigFloat z;
  z.sign = a.sign == b.sign;
  z.exp = s.ae + s.be;
  z.L = s.al + s.bl;
  z.T.reset(new uint32_t[z.L]);
  Item it{s.at, s.al, s.bt, s.bl, dst};
  *dst = std::move(z);  // limbs are filled by run()
  items_.push_back(it);
}

void MulBatch::run(bool plain) {
  if (items_.empty()) return;
  std::vector<pcg_mul_desc> d(items_.size());
  for (size_t i = 0; i < items_.size(); ++i)
    d[i] = pcg_mul_desc{items_[i].at, items_[i].al, items_[i].bt, items_[i].bl, items_[i].dst->T.get()};
  const uint32_t flags = g_flags | (plain ? PCG_BATCH_PLAIN : 0u);
  int rc = pcg_mul_batch(d.data(), d.size(), flags);
  if (rc == PCG_ERR_NOT_INIT) {
    rc = pcg_init(0);
    if (rc == PCG_OK) rc = pcg_mul_batch(d.data(), d.size(), flags);
  }
  if (rc != PCG_OK) throw_status(rc);
  g_mul_count.fetch_add(items_.size(), std::memory_order_relaxed);
  for (Item& it : items_) {
    BigFloat& z = *it.dst;
    if (z.T[z.L - 1] == 0) --z.L;
  }
  items_.clear();
}

// ------------------------------------------------ Newton iterations (BigFloat.cpp:445-567)
// Two-limb float seed (value 2^32 <= v < 2^64 scaled by exponent e), as the
// p == 0 branches of rcp and invsqrt build it.
BigFloat BigFloat::seed(double val, int64_t e, bool sign) {
  while (val < 4294967296.) {
    val *= 4294967296.;
    --e;
  }
  const uint64_t v = (uint64_t)val;
  const uint32_t limbs[2] = {(uint32_t)v, (uint32_t)(v >> 32)};
  return BigFloat::from_parts(sign, e, limbs, 2);
}

// Top three limbs as a double, BigFloat.cpp:451-464.
double BigFloat::top_double(int64_t* e) const {
  size_t al = L;
  int64_t ae = exp;
  const uint32_t* at = T.get();
  if (al > 3) {
    at += al - 3;
    ae += (int64_t)(al - 3);
    al = 3;
  }
  double val = at[0];
  if (al >= 2) val += at[1] * 4294967296.;
  if (al >= 3) val += at[2] * 4294967296. * 4294967296.;
  *e = ae;
  return val;
}

BigFloat BigFloat::rcp(size_t p, int threads) const { return newton_rcp(*this, p); }

BigFloat BigFloat::div(const BigFloat& x, size_t p, int threads) const {
  return this->mul(x.rcp(p, threads), p, threads);
}

BigFloat invsqrt(uint32_t x, size_t p, int threads) { return newton_invsqrt<BigFloat>(x, p); }

}  // namespace piculator
