// Device-resident BigFloat arithmetic: the limb loops of the reference's
// add/sub/compare/multiply-by-word (BigFloat.cpp:123-232, 233-287, 156-183)
// for values that stay in HBM.  Each operation is a sum per limb position
// followed by the carry scan of carry_kernels.hpp (block-local carries with
// carry-in 0, a scan of the block carry functions, a fix-up ripple), so
// every limb position is handled by its own thread and no loop runs over the
// length of a value.
//
// A value is (limbs, L, exp): limb i has weight 2^(32 (exp + i)).  Positions
// outside [exp, exp + L) read as zero (BigFloat's word_at).
#pragma once
#include "carry_kernels.hpp"

namespace pcg {

struct DfOperand {
  const uint32_t* t;
  int64_t exp;
  uint64_t L;
};

__device__ __forceinline__ uint32_t df_word(const DfOperand& a, int64_t pos) {
  const int64_t i = pos - a.exp;
  return (i >= 0 && (uint64_t)i < a.L) ? a.t[i] : 0u;
}

enum DfOp : int { kDfAdd = 0, kDfSub = 1, kDfMulU32 = 2 };

struct DfArgs {
  DfOperand a, b;   // add/sub: A +- B over positions [bot, bot + n); mul_u32: a * x
  int64_t bot;
  uint64_t n;       // limbs produced by the block pass
  uint32_t x;       // multiplier of kDfMulU32
  uint32_t* out;    // n limbs
  CarryFn* agg;     // per block function
};

// Limb sums of one block: A + B, A + ~B (+1 at the bottom, for A - B with
// |A| >= |B|: the carry out of the top is then 1 and is dropped), or
// lo(a_k x) + hi(a_{k-1} x).  Carries are resolved with carry-in 0; each
// block records its carry function for the scan.
template <int OP>
__global__ __launch_bounds__(kCarryThreads) void k_df_local(DfArgs a) {
  __shared__ CarryFn wtot[kCarryThreads / 64];
  const int tid = threadIdx.x;
  const uint64_t base = (uint64_t)blockIdx.x * kCarryBlock + (uint64_t)tid * kCarryLPT;
  uint64_t s[kCarryLPT];
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    const uint64_t k = base + q;
    uint64_t v = 0;
    if (k < a.n) {
      if (OP == kDfMulU32) {
        // n = L + 1: the top position receives only the high word of limb L - 1
        const uint64_t cur = k < a.a.L ? (uint64_t)a.a.t[k] * a.x : 0;
        const uint64_t prev = (k >= 1 && k - 1 < a.a.L) ? (uint64_t)a.a.t[k - 1] * a.x : 0;
        v = (cur & 0xffffffffull) + (prev >> 32);
      } else {
        const int64_t pos = a.bot + (int64_t)k;
        const uint32_t wa = df_word(a.a, pos), wb = df_word(a.b, pos);
        v = (uint64_t)wa + (OP == kDfSub ? (uint64_t)(uint32_t)~wb : (uint64_t)wb);
        if (OP == kDfSub && k == 0) v += 1;
      }
    }
    s[q] = v;
  }
  // Positions past the top pass the carry through unchanged (identity), so
  // the block function ends at the top limb: its value is the carry out.
  // sums below 2^33: carries of at most 2, 32-bit limb functions (the
  // scan's LDS scratch taken from wtot)
  LimbFn f = lf_identity();
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    if (base + q >= a.n) continue;
    f = lf_compose(f, lf_of_sum(s[q]));
  }
  const LimbFn pre = lf_block_exclusive(f, reinterpret_cast<LimbFn*>(wtot));
  uint64_t c = lf_eval0(pre);
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    const uint64_t k = base + q;
    const uint64_t t = (s[q] & 0xffffffffull) + c;
    if (k < a.n) a.out[k] = (uint32_t)t;
    c = (s[q] >> 32) + (t >> 32);
  }
  if (tid == kCarryThreads - 1) a.agg[blockIdx.x] = cf_of_limb(lf_compose(pre, f));
}

// Block-wide maximum, one atomicMax per workgroup (256 threads).
__device__ __forceinline__ void block_max_to(unsigned long long v, unsigned long long* slot) {
  __shared__ unsigned long long wmax[4];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long o = __shfl_xor(v, d, 64);
    v = o > v ? o : v;
  }
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wmax[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = wmax[w] > m ? wmax[w] : m;
    if (m) atomicMax(slot, m);
  }
}

// The searches below walk down from the top: thread j of the grid visits
// n - 1 - j, n - 1 - j - G, ... (G = grid threads) and stops at its first
// hit, which is the highest hit of its residue class; the block maximum of
// those goes to *slot (zeroed first) as index + 1 (0: no hit).  The work is
// the length of the run without hits at the top plus one pass of the grid.

// Highest position in [lo, lo + n) where the two operands differ.
__global__ __launch_bounds__(256) void k_df_cmp(DfOperand a, DfOperand b, int64_t lo, uint64_t n,
                                                unsigned long long* slot) {
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long hit = 0;
  for (uint64_t i = n - 1 - j; j < n; i -= G) {
    if (df_word(a, lo + (int64_t)i) != df_word(b, lo + (int64_t)i)) {
      hit = i + 1;
      break;
    }
    if (i < G) break;
  }
  block_max_to(hit, slot);
}

// Turns the difference position into the comparison result (-1, 0, 1).
__global__ void k_df_cmp_final(DfOperand a, DfOperand b, int64_t lo,
                               const unsigned long long* slot, int64_t* res) {
  const unsigned long long h = *slot;
  if (h == 0) {
    *res = 0;
    return;
  }
  const int64_t pos = lo + (int64_t)(h - 1);
  *res = df_word(a, pos) > df_word(b, pos) ? 1 : -1;
}

// Index + 1 of the highest nonzero limb of t[0, n) (0: all zero).
__global__ __launch_bounds__(256) void k_df_topnz(const uint32_t* t, uint64_t n,
                                                  unsigned long long* slot) {
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long hit = 0;
  for (uint64_t i = n - 1 - j; j < n; i -= G) {
    if (t[i] != 0) {
      hit = i + 1;
      break;
    }
    if (i < G) break;
  }
  block_max_to(hit, slot);
}

__global__ void k_df_put_u64(unsigned long long* p, unsigned long long v) { *p = v; }

// Writes limb value v at t[i] (the carry limb of an addition or of a
// multiply by a word).
__global__ void k_df_put_u32(uint32_t* t, uint64_t i, uint32_t v) { t[i] = v; }

// Hexadecimal text of the top `digits` hex digits below the point of a
// value whose integer part is the single top limb (BigFloat::to_string_hex,
// BigFloat.cpp:95-106): digit j (0-based) is nibble 7 - (j mod 8) of limb
// L - 2 - j / 8.
__global__ __launch_bounds__(256) void k_df_hex(const uint32_t* t, uint64_t L, uint64_t digits,
                                                char* out) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= digits) return;
  const uint64_t limb = j >> 3;
  uint32_t w = 0;
  if (limb + 2 <= L) w = t[L - 2 - limb];
  const uint32_t nib = (w >> (4 * (7 - (j & 7)))) & 15u;
  out[j] = (char)(nib < 10 ? '0' + nib : 'a' + nib - 10);
}

}  // namespace pcg
