// DeviceFloat: a BigFloat (reference BigFloat.h:8-46) whose limbs stay in
// MI355X memory.  Every operation has the reference's semantics (precision
// p meaning p + 2 limbs, low limbs cut without rounding, exponent in limbs)
// and runs on the GPU through the pcg_df_* entry points of the C ABI, so a
// chain of operations (the top of the binary splitting, the Newton division
// and square root of Pi()) never moves a value through host memory.
// Results are bit-identical to the same BigFloat operations.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <utility>

#include "../../include/picugpu.h"
#include "BigFloat.h"

namespace piculator {

class DeviceFloat {
 public:
  DeviceFloat();
  DeviceFloat(uint32_t x, bool sign = true);  // BigFloat(uint32_t, bool), BigFloat.cpp:64-79
  explicit DeviceFloat(const BigFloat& x);    // upload
  DeviceFloat(DeviceFloat&& x) noexcept;
  DeviceFloat(const DeviceFloat& x);
  DeviceFloat& operator=(DeviceFloat&& x) noexcept;
  DeviceFloat& operator=(const DeviceFloat& x);
  ~DeviceFloat();

  BigFloat to_host() const;  // download
  size_t get_precision() const { return (size_t)v_.L; }
  int64_t get_exponent() const { return v_.exp; }
  bool get_sign() const { return v_.sign != 0; }

  void negate();
  DeviceFloat mul(uint32_t x) const;
  DeviceFloat add(const DeviceFloat& x, size_t p = 0) const;
  DeviceFloat sub(const DeviceFloat& x, size_t p = 0) const;
  DeviceFloat mul(const DeviceFloat& x, size_t p = 0) const;
  // mul(x, p), keeping x's forward transform for a next product by x
  // (PCG_KEEP_B); mul_reuse: x is that operand, unchanged (PCG_REUSE_B)
  DeviceFloat mul_keep(const DeviceFloat& x, size_t p) const;
  DeviceFloat mul_reuse(const DeviceFloat& x, size_t p) const;
  // {x1.mul(*this, p), x2.mul(*this, p)}, this value's forward transform
  // shared by the two products (pcg_df_mul_pair)
  std::pair<DeviceFloat, DeviceFloat> mul_pair(const DeviceFloat& x1, const DeviceFloat& x2,
                                               size_t p = 0) const;
  DeviceFloat rcp(size_t p) const;
  DeviceFloat div(const DeviceFloat& x, size_t p) const;
  std::string to_string_hex(size_t digits) const;
  // to_string_hex written straight into out[0..cap); returns the length.
  size_t to_hex_into(size_t digits, char* out, size_t cap) const;

  // Newton hooks (Newton.h)
  void shift_exp(int64_t d) { v_.exp += d; }
  double top_double(int64_t* e) const;
  static DeviceFloat seed(double val, int64_t e, bool sign);

  pcg_dfloat& raw() { return v_; }
  const pcg_dfloat& raw() const { return v_; }

 private:
  pcg_dfloat v_;
};

DeviceFloat invsqrt_device(uint32_t x, size_t p);

}  // namespace piculator
