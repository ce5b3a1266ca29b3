// Device-resident binary splitting of the Chudnovsky series (reference
// PI.h:64-112) for subtrees whose values stay inside the precision window.
//
// While every value of a subtree has at most p + 2 limbs, BigFloat never
// chops (BigFloat.cpp:201-209,342-366): all values are exact integers with
// exponent 0 and no leading zero limb, products are exact (one zero top limb
// trimmed, BigFloat.cpp:411-412), add/sub are exact with usub's full
// leading-zero strip (BigFloat.cpp:277-284).  The kernels below compute
// exactly those values, so results are bit-identical to the BigFloat path.
//
// Values live in one arena; metadata (offset, length, sign) in device memory.
#pragma once
#include "batch_kernels.hpp"
#include "carry_kernels.hpp"

namespace pcg {

struct BsrVal {
  uint64_t off;  // limb offset in the arena (fixed by the host plan)
  uint32_t L;    // limbs (0 = zero)
  int32_t sign;  // 1 positive or zero, 0 negative
};

// ------------------------------------------------------------------ leaves
// Fixed-capacity unsigned integer for the leaf terms (< 2^160).
struct Small {
  uint32_t t[6];
  int L;
};

__device__ __forceinline__ void small_set(Small& z, uint32_t x) {
  z.L = x ? 1 : 0;
  z.t[0] = x;
}
// z = z * x as BigFloat::mul(uint32_t) (BigFloat.cpp:156-183)
__device__ __forceinline__ void small_mul(Small& z, uint32_t x) {
  if (z.L == 0 || x == 0) {
    z.L = 0;
    return;
  }
  uint64_t carry = 0;
  for (int c = 0; c < z.L; ++c) {
    carry += (uint64_t)z.t[c] * x;
    z.t[c] = (uint32_t)carry;
    carry >>= 32;
  }
  if (carry) z.t[z.L++] = (uint32_t)carry;
}
// z = z + x (both non-negative) as BigFloat::uadd with p = 0
__device__ __forceinline__ void small_add(Small& z, uint32_t x) {
  uint64_t carry = x;
  for (int c = 0; c < z.L && carry; ++c) {
    carry += z.t[c];
    z.t[c] = (uint32_t)carry;
    carry >>= 32;
  }
  if (carry) z.t[z.L++] = (uint32_t)carry;
  if (z.L == 0 && x) {
    z.t[0] = x;
    z.L = 1;
  }
}

// Leaf b = a + rel[i] (PI.h:67-88), with the reference's uint32 arithmetic
// for 2b-1, 6b-5, 6b-1.  ids: value ids of (P, Q, R) for each leaf.
__global__ void k_bsr_leaves(const uint32_t* rel, uint32_t a, const uint32_t* ids, uint64_t n,
                             BsrVal* meta, uint32_t* arena) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = a + rel[i];
  Small P, Q, R;
  // P = (13591409 + 545140134 b)(2b-1)(6b-5)(6b-1), negative for odd b
  small_set(P, b);
  small_mul(P, 545140134u);
  small_add(P, 13591409u);
  small_mul(P, 2 * b - 1);
  small_mul(P, 6 * b - 5);
  small_mul(P, 6 * b - 1);
  // Q = b^3 * 26726400 * 409297880
  small_set(Q, b);
  small_mul(Q, b);
  small_mul(Q, b);
  small_mul(Q, 26726400u);
  small_mul(Q, 409297880u);
  // R = (2b-1)(6b-5)(6b-1)
  small_set(R, 2 * b - 1);
  small_mul(R, 6 * b - 5);
  small_mul(R, 6 * b - 1);
  const Small* v[3] = {&P, &Q, &R};
  for (int j = 0; j < 3; ++j) {
    BsrVal& m = meta[ids[3 * i + j]];
    for (int c = 0; c < v[j]->L; ++c) arena[m.off + c] = v[j]->t[c];
    m.L = (uint32_t)v[j]->L;
    m.sign = (j == 0 && (b & 1) && P.L) ? 0 : 1;
  }
}

// ---------------------------------------------------------------- products
// One product of a height: value ids of the operands and of the result.
struct BsrMul {
  uint32_t a, b, out;
};

// Fills the batch descriptors of a height's products from device metadata.
__global__ void k_bsr_descs(const BsrMul* muls, uint64_t n, const BsrVal* meta, uint32_t* arena,
                            BatchDesc* descs) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BsrMul m = muls[i];
  const BsrVal va = meta[m.a], vb = meta[m.b], vo = meta[m.out];
  BatchDesc d;
  d.a = arena + va.off;
  d.b = arena + vb.off;
  d.out = arena + vo.off;
  d.na = va.L;
  d.nb = vb.L;
  descs[i] = d;
}

// Length and sign of each product: L = na + nb less one zero top limb;
// zero operand -> zero (BigFloat.cpp:334-335,369-372,411-412).
__global__ void k_bsr_trim(const BsrMul* muls, uint64_t n, BsrVal* meta, const uint32_t* arena) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BsrMul m = muls[i];
  const BsrVal va = meta[m.a], vb = meta[m.b];
  BsrVal& vo = meta[m.out];
  if (va.L == 0 || vb.L == 0) {
    vo.L = 0;
    vo.sign = 1;
    return;
  }
  uint32_t L = va.L + vb.L;
  if (arena[vo.off + L - 1] == 0) --L;
  vo.L = L;
  vo.sign = va.sign == vb.sign ? 1 : 0;
}

// -------------------------------------------------------------- signed add
// P = x.add(y) (BigFloat.cpp:288-305 with uadd/usub/ucmp, 184-287): one
// workgroup per node, limbs in chunks of LPT * blockDim.x with a chunk carry.
struct BsrAdd {
  uint32_t x, y, out;
};

template <int NT>
__global__ __launch_bounds__(NT) void k_bsr_add(const BsrAdd* adds, BsrVal* meta, uint32_t* arena) {
  constexpr int LPT = 4;
  constexpr int CH = NT * LPT;
  __shared__ CarryFn wtot[NT / 64];
  __shared__ int sh_i;
  __shared__ uint64_t sh_c;
  const BsrAdd op = adds[blockIdx.x];
  const BsrVal mx = meta[op.x], my = meta[op.y];
  BsrVal& mo = meta[op.out];
  const int tid = threadIdx.x;
  const uint32_t* X = arena + mx.off;
  const uint32_t* Y = arena + my.off;
  uint32_t* O = arena + mo.off;
  // decide the operation: A +- B with |A| >= |B| for subtraction
  bool sub = mx.sign != my.sign;
  const uint32_t* A = X;
  const uint32_t* B = Y;
  uint32_t LA = mx.L, LB = my.L;
  int sign = mx.sign;
  if (sub) {
    int cmp;
    if (mx.L != my.L) {
      cmp = mx.L > my.L ? 1 : -1;
    } else {
      // ucmp: highest limb where the magnitudes differ
      if (tid == 0) sh_i = -1;
      __syncthreads();
      int top = -1;
      for (uint32_t k = tid; k < mx.L; k += NT)
        if (X[k] != Y[k]) top = (int)k;
      if (top >= 0) atomicMax(&sh_i, top);
      __syncthreads();
      const int i = sh_i;
      cmp = i < 0 ? 0 : (X[i] > Y[i] ? 1 : -1);
    }
    if (cmp == 0) {  // equal magnitudes: zero (usub strips everything)
      if (tid == 0) {
        mo.L = 0;
        mo.sign = 1;
      }
      return;
    }
    if (cmp < 0) {
      A = Y;
      B = X;
      LA = my.L;
      LB = mx.L;
      sign = my.sign;
    }
  } else if (LB > LA) {
    A = Y;
    B = X;
    LA = my.L;
    LB = mx.L;
  }
  // A + B, or A + ~B + 1 (mod 2^(32 LA)) for A - B.  The carry out of the
  // top limb lands at position LA: inside the last chunk when LA is not a
  // multiple of the chunk size (sh_fin), else it is the last chunk's carry.
  __shared__ uint32_t sh_fin;
  if (tid == 0) {
    sh_i = -1;
    sh_c = sub ? 1 : 0;
    sh_fin = 0;
  }
  __syncthreads();
  int topnz = -1;
  for (uint32_t base = 0; base < LA; base += CH) {
    uint64_t s[LPT];
    const uint64_t cin = sh_c;
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const uint32_t k = base + tid * LPT + q;
      uint64_t x = 0;
      if (k < LA) {
        const uint32_t bk = k < LB ? B[k] : 0u;
        x = (uint64_t)A[k] + (sub ? (uint32_t)~bk : bk);
      }
      if (tid == 0 && q == 0) x += cin;
      s[q] = x;
    }
    uint32_t D[LPT];
    uint64_t c;
    block_carry_sums<LPT>(s, wtot, D, c);
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const uint32_t k = base + tid * LPT + q;
      if (k < LA) {
        O[k] = D[q];
        if (D[q]) topnz = (int)k;
      } else if (k == LA) {
        sh_fin = D[q];
      }
    }
    __syncthreads();
    if (tid == NT - 1) sh_c = c;
    __syncthreads();
  }
  if (topnz >= 0) atomicMax(&sh_i, topnz);
  __syncthreads();
  if (tid == 0) {
    uint32_t L;
    if (sub) {
      L = (uint32_t)(sh_i + 1);  // strip every leading zero (usub)
    } else {
      const uint64_t carry = (LA % CH == 0) ? sh_c : sh_fin;
      L = LA;
      if (carry) O[L++] = 1;  // carry limb (uadd)
    }
    mo.L = L;
    mo.sign = L ? sign : 1;
  }
}

}  // namespace pcg
