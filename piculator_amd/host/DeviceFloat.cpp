// DeviceFloat over the pcg_df_* entry points (see DeviceFloat.h).
#include "DeviceFloat.h"

#include <cstring>
#include <vector>

#include "Newton.h"

namespace piculator {

namespace {
// Errors as the reference reports them: thrown string literals.
void check(int rc) {
  if (rc == PCG_OK) return;
  if (rc == PCG_ERR_TOO_LARGE) throw "FFT length is too small.";
  throw pcg_last_error_string();
}

// Runs f; on PCG_ERR_NOT_INIT initialises device 0 and runs it again.
template <class Fn>
void call(Fn f) {
  int rc = f();
  if (rc == PCG_ERR_NOT_INIT) {
    rc = pcg_init(0);
    if (rc == PCG_OK) rc = f();
  }
  check(rc);
}
}  // namespace

DeviceFloat::DeviceFloat() {
  std::memset(&v_, 0, sizeof v_);
  v_.sign = 1;
}

DeviceFloat::DeviceFloat(uint32_t x, bool sign) {
  std::memset(&v_, 0, sizeof v_);
  v_.sign = 1;
  if (x == 0) return;
  call([&] { return pcg_df_set(&v_, &x, 1, 0, sign ? 1 : 0); });
}

DeviceFloat::DeviceFloat(const BigFloat& x) {
  std::memset(&v_, 0, sizeof v_);
  v_.sign = 1;
  call([&] {
    return pcg_df_set(&v_, x.limbs(), x.get_precision(), x.get_exponent(), x.get_sign() ? 1 : 0);
  });
}

DeviceFloat::DeviceFloat(DeviceFloat&& x) noexcept : v_(x.v_) {
  std::memset(&x.v_, 0, sizeof x.v_);
  x.v_.sign = 1;
}

DeviceFloat::DeviceFloat(const DeviceFloat& x) {
  std::memset(&v_, 0, sizeof v_);
  v_.sign = 1;
  call([&] { return pcg_df_copy(&v_, &x.v_); });
}

DeviceFloat& DeviceFloat::operator=(DeviceFloat&& x) noexcept {
  if (this != &x) {
    if (v_.limbs) (void)pcg_df_free(&v_);
    v_ = x.v_;
    std::memset(&x.v_, 0, sizeof x.v_);
    x.v_.sign = 1;
  }
  return *this;
}

DeviceFloat& DeviceFloat::operator=(const DeviceFloat& x) {
  if (this != &x) call([&] { return pcg_df_copy(&v_, &x.v_); });
  return *this;
}

DeviceFloat::~DeviceFloat() {
  if (v_.limbs) (void)pcg_df_free(&v_);
}

BigFloat DeviceFloat::to_host() const {
  std::vector<uint32_t> t(v_.L ? v_.L : 1);
  if (v_.L) call([&] { return pcg_df_read(&v_, 0, v_.L, t.data()); });
  if (v_.L == 0) return BigFloat();
  return BigFloat::from_parts(v_.sign != 0, v_.exp, t.data(), v_.L);
}

void DeviceFloat::negate() {
  if (v_.L == 0) return;  // BigFloat.cpp:150-155
  v_.sign = !v_.sign;
}

DeviceFloat DeviceFloat::mul(uint32_t x) const {
  DeviceFloat z;
  call([&] { return pcg_df_mul_u32(&z.v_, &v_, x); });
  return z;
}

DeviceFloat DeviceFloat::add(const DeviceFloat& x, size_t p) const {
  DeviceFloat z;
  call([&] { return pcg_df_add(&z.v_, &v_, &x.v_, p); });
  return z;
}

DeviceFloat DeviceFloat::sub(const DeviceFloat& x, size_t p) const {
  DeviceFloat z;
  call([&] { return pcg_df_sub(&z.v_, &v_, &x.v_, p); });
  return z;
}

DeviceFloat DeviceFloat::mul(const DeviceFloat& x, size_t p) const {
  DeviceFloat z;
  call([&] { return pcg_df_mul(&z.v_, &v_, &x.v_, p, mul_engine_flags()); });
  if (v_.L && x.v_.L) count_big_muls(1);
  return z;
}

DeviceFloat DeviceFloat::mul_keep(const DeviceFloat& x, size_t p) const {
  DeviceFloat z;
  call([&] { return pcg_df_mul(&z.v_, &v_, &x.v_, p, mul_engine_flags() | PCG_KEEP_B); });
  if (v_.L && x.v_.L) count_big_muls(1);
  return z;
}

DeviceFloat DeviceFloat::mul_reuse(const DeviceFloat& x, size_t p) const {
  DeviceFloat z;
  call([&] { return pcg_df_mul(&z.v_, &v_, &x.v_, p, mul_engine_flags() | PCG_REUSE_B); });
  if (v_.L && x.v_.L) count_big_muls(1);
  return z;
}

std::pair<DeviceFloat, DeviceFloat> DeviceFloat::mul_pair(const DeviceFloat& x1, const DeviceFloat& x2,
                                                          size_t p) const {
  std::pair<DeviceFloat, DeviceFloat> z;
  call([&] { return pcg_df_mul_pair(&z.first.v_, &x1.v_, &z.second.v_, &x2.v_, &v_, p, mul_engine_flags()); });
  if (v_.L) count_big_muls((x1.v_.L ? 1 : 0) + (x2.v_.L ? 1 : 0));
  return z;
}

DeviceFloat DeviceFloat::rcp(size_t p) const { return newton_rcp(*this, p); }

DeviceFloat DeviceFloat::div(const DeviceFloat& x, size_t p) const { return mul(x.rcp(p), p); }

DeviceFloat invsqrt_device(uint32_t x, size_t p) { return newton_invsqrt<DeviceFloat>(x, p); }

double DeviceFloat::top_double(int64_t* e) const {
  const size_t al = v_.L > 3 ? 3 : (size_t)v_.L;
  uint32_t at[3] = {0, 0, 0};
  call([&] { return pcg_df_read(&v_, v_.L - al, al, at); });
  double val = at[0];
  if (al >= 2) val += at[1] * 4294967296.;
  if (al >= 3) val += at[2] * 4294967296. * 4294967296.;
  *e = v_.exp + (int64_t)(v_.L - al);
  return val;
}

DeviceFloat DeviceFloat::seed(double val, int64_t e, bool sign) {
  const BigFloat s = BigFloat::seed(val, e, sign);
  return DeviceFloat(s);
}

size_t DeviceFloat::to_hex_into(size_t digits, char* out, size_t cap) const {
  uint64_t n = 0;
  const int rc = pcg_df_to_hex(&v_, digits, out, cap, &n);
  if (rc == PCG_ERR_INVALID) throw "Limit Exceeded";
  check(rc);
  return (size_t)n;
}

std::string DeviceFloat::to_string_hex(size_t digits) const {
  std::string s(digits + 16, '\0');
  const size_t n = to_hex_into(digits, &s[0], s.size());
  s.resize(n);
  return s;
}

}  // namespace piculator
