// Modular arithmetic on FP64 for the two-prime NTT (device side).
//
// Residues are exact integers held in doubles.  Both primes are below 2^47
// and have 2^31 | p - 1, so transforms up to 2^30 points exist.  Values are
// kept lazily reduced: stored values satisfy |x| <~ p, the outputs of one
// radix-8 stage satisfy |x| <= 8p < 2^50, far inside the 2^53 range where
// doubles hold integers exactly.
//
// mulmod (a * w mod p, result in (-p, p)):
//   h = fl(a*w); l = fma(a, w, -h)        -> a*w == h + l exactly (TwoProduct)
//   q = rint(a * (w/p))                    -> quotient estimate, off by < 1
//   r = fma(-q, p, h) + l                  -> a*w - q*p, exact (|.| < 2^53)
// Six full-rate FP64 VALU operations; the 64-bit integer alternatives need
// four quarter-rate 32x32->64 multiplies plus the reduction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace pcg {

constexpr uint64_t kP0 = 140713866035201ull;  // 0x7ffa80000001 = 131050 * 2^30 + 1
constexpr uint64_t kP1 = 140700981133313ull;  // 0x7ff780000001 = 131038 * 2^30 + 1
constexpr uint64_t kGen = 3;                  // primitive root of both primes
constexpr int kLogRootMax = 30;               // tables are built on a 2^30-th root

struct Prime {
  double p, pinv;           // p and fl(1/p)
  double2 w8f[3], w8i[3];   // omega_8^{1,2,3} and inverses, with (w, w/p) pairs
  double2 w16f[3], w16i[3]; // omega_16^{1,3,9} and inverses
  double2 ninv;             // N^{-1} for the current transform length
  const double2* twL;       // omega_4096^e, e < 4096
  const double2* twHi;      // omega_{2^30}^{e * 2^15}, e < 2^15
  const double2* twLo;      // omega_{2^30}^{e}, e < 2^15
};

__device__ __forceinline__ double mm_shoup(double a, double w, double wp, double p) {
  double h = a * w;
  double l = __builtin_fma(a, w, -h);
  double q = __builtin_rint(a * wp);
  double r = __builtin_fma(-q, p, h);
  return r + l;
}

__device__ __forceinline__ double mm_gen(double a, double b, double p, double pinv) {
  double h = a * b;
  double l = __builtin_fma(a, b, -h);
  double q = __builtin_rint(h * pinv);
  double r = __builtin_fma(-q, p, h);
  return r + l;
}

// |x| < 2^52  ->  |result| <= p/2 + small
__device__ __forceinline__ double reduce(double x, double p, double pinv) {
  double q = __builtin_rint(x * pinv);
  return __builtin_fma(-q, p, x);
}

// |x| < 2^52 -> [0, p).  x * pinv is within 2^-46 of x / p, so the nearest
// integer q leaves x - q p in (-p, p) (exact in the fma), and one
// conditional + p finishes.
__device__ __forceinline__ double canon(double x, double p, double pinv) {
  const double q = __builtin_rint(x * pinv);
  const double r = __builtin_fma(-q, p, x);
  return r < 0.0 ? r + p : r;
}

}  // namespace pcg
