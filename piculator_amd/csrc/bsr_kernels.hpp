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
// Leaves and the nodes of up to four leaves, one thread each, in fixed-size
// limb arrays (registers): a leaf's P < 2^158, Q < 2^150, R < 2^96 take 5,
// 5, 3 limbs, a two-leaf node's P, Q, R at most 11, 10, 6 and a node of
// three or four leaves' 22, 20, 12.  Such a node is computed from its leaves
// directly, so neither its leaves nor the products below it pass through
// memory.
// Values are normalised (no leading zero limb; zero has L = 0 and sign 1),
// as BigFloat leaves them: the product's one-limb trim and usub's strip
// (BigFloat.cpp:411-412, 277-284) give exactly the significant limbs.

// z[0..N] = z[0..N) * x (z[N] receives the carry)
template <int N>
__host__ __device__ __forceinline__ void fx_mul_u32(uint32_t* z, uint32_t x) {
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    c += (uint64_t)z[i] * x;
    z[i] = (uint32_t)c;
    c >>= 32;
  }
  z[N] = (uint32_t)c;
}

// z[0..NA+NB) = a[0..NA) * b[0..NB)
template <int NA, int NB>
__host__ __device__ __forceinline__ void fx_mul(const uint32_t* a, const uint32_t* b, uint32_t* z) {
#pragma unroll
  for (int k = 0; k < NA + NB; ++k) z[k] = 0;
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      c += (uint64_t)a[i] * b[j] + z[i + j];
      z[i + j] = (uint32_t)c;
      c >>= 32;
    }
    z[i + NB] = (uint32_t)c;
  }
}

// Number of significant limbs of z[0..N).
template <int N>
__host__ __device__ __forceinline__ uint32_t fx_len(const uint32_t* z) {
  uint32_t L = 0;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (z[i]) L = i + 1;
  return L;
}

// P, Q, R of the leaf b (PI.h:67-88) with the reference's uint32
// arithmetic for 2b-1, 6b-5, 6b-1; P is negative for odd b.
struct LeafVals {
  uint32_t P[5], Q[5], R[3];
  int32_t sP;  // sign of P: 1 positive, 0 negative
};

__host__ __device__ __forceinline__ LeafVals bsr_leaf(uint32_t b) {
  LeafVals v;
  const uint32_t f1 = 2 * b - 1, f2 = 6 * b - 5, f3 = 6 * b - 1;
  // P = (13591409 + 545140134 b) f1 f2 f3
  uint32_t p[6] = {b, 0, 0, 0, 0, 0};
  fx_mul_u32<1>(p, 545140134u);
  const uint64_t lo = (uint64_t)p[0] + 13591409u;
  p[0] = (uint32_t)lo;
  p[1] += (uint32_t)(lo >> 32);  // b 545140134 + 13591409 < 2^63: no carry out of p[1]
  fx_mul_u32<2>(p, f1);
  fx_mul_u32<3>(p, f2);
  fx_mul_u32<4>(p, f3);
  // Q = b^3 26726400 409297880
  uint32_t q[6] = {b, 0, 0, 0, 0, 0};
  fx_mul_u32<1>(q, b);
  fx_mul_u32<2>(q, b);
  fx_mul_u32<3>(q, 26726400u);
  fx_mul_u32<4>(q, 409297880u);
  // R = f1 f2 f3
  uint32_t r[4] = {f1, 0, 0, 0};
  fx_mul_u32<1>(r, f2);
  fx_mul_u32<2>(r, f3);
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    v.P[i] = p[i];
    v.Q[i] = q[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) v.R[i] = r[i];
  v.sP = (b & 1) ? 0 : 1;
  return v;
}

template <int N>
__host__ __device__ __forceinline__ void bsr_put(BsrVal* meta, uint32_t id, uint32_t* arena, const uint32_t* z, int32_t sign) {
  BsrVal& m = meta[id];
  const uint32_t L = fx_len<N>(z);
#pragma unroll
  for (int i = 0; i < N; ++i)
    if ((uint32_t)i < L) arena[m.off + i] = z[i];
  m.L = L;
  m.sign = L ? sign : 1;
}

// P, Q, R of a node in fixed-size arrays (zero above the value's limbs).
template <int NP, int NQ, int NR>
struct NodeVals {
  uint32_t P[NP], Q[NQ], R[NR];
  int32_t sP;  // sign of P: 1 positive (or zero), 0 negative
};

// The node combine of PI.h:95-112 on fixed-size values: x = P0 Q1,
// Q = Q0 Q1, y = P1 R0, R = R0 R1, P = x + y (x has P0's sign and y has
// P1's: Q and R are positive).  z.P needs NPL + NQR + 1 >= NPR + NRL + 1
// limbs.
template <int NPL, int NQL, int NRL, int NPR, int NQR, int NRR>
__host__ __device__ __forceinline__ void bsr_combine(const NodeVals<NPL, NQL, NRL>& l,
                                                     const NodeVals<NPR, NQR, NRR>& r,
                                                     NodeVals<NPL + NQR + 1, NQL + NQR, NRL + NRR>& z) {
  constexpr int NX = NPL + NQR, NY = NPR + NRL;
  static_assert(NY <= NX, "y is the shorter product");
  uint32_t x[NX], y[NX];
  fx_mul<NPL, NQR>(l.P, r.Q, x);
  fx_mul<NQL, NQR>(l.Q, r.Q, z.Q);
  fx_mul<NPR, NRL>(r.P, l.R, y);
#pragma unroll
  for (int k = NY; k < NX; ++k) y[k] = 0;
  fx_mul<NRL, NRR>(l.R, r.R, z.R);
  int32_t sign = l.sP;
  if (l.sP == r.sP) {
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      c += (uint64_t)x[k] + y[k];
      z.P[k] = (uint32_t)c;
      c >>= 32;
    }
    z.P[NX] = (uint32_t)c;
  } else {
    // |x| - |y| or |y| - |x|, the larger magnitude's sign
    int cmp = 0;
#pragma unroll
    for (int k = NX - 1; k >= 0; --k)
      if (cmp == 0 && x[k] != y[k]) cmp = x[k] > y[k] ? 1 : -1;
    int64_t br = 0;
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      const int64_t d = cmp >= 0 ? (int64_t)x[k] - y[k] + br : (int64_t)y[k] - x[k] + br;
      z.P[k] = (uint32_t)d;
      br = d < 0 ? -1 : 0;
    }
    z.P[NX] = 0;
    if (cmp < 0) sign = r.sP;
  }
  z.sP = sign;
}

using LeafNode = NodeVals<5, 5, 3>;
using PairNode = NodeVals<11, 10, 6>;
using QuadNode = NodeVals<22, 20, 12>;

__host__ __device__ __forceinline__ LeafNode bsr_leaf_node(uint32_t b) {
  const LeafVals v = bsr_leaf(b);
  LeafNode n;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    n.P[i] = v.P[i];
    n.Q[i] = v.Q[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) n.R[i] = v.R[i];
  n.sP = v.sP;
  return n;
}

// The node of the leaves b and b + 1.
__host__ __device__ __forceinline__ PairNode bsr_pair_node(uint32_t b) {
  PairNode z;
  bsr_combine(bsr_leaf_node(b), bsr_leaf_node(b + 1), z);
  return z;
}

template <int NP, int NQ, int NR>
__host__ __device__ __forceinline__ void bsr_put_node(BsrVal* meta, const uint32_t* id, uint32_t* arena,
                                                      const NodeVals<NP, NQ, NR>& v) {
  bsr_put<NP>(meta, id[0], arena, v.P, v.sP);
  bsr_put<NQ>(meta, id[1], arena, v.Q, 1);
  bsr_put<NR>(meta, id[2], arena, v.R, 1);
}

// One entry per leaf of the subtree (rel, ids as the plan lays them out):
// rel = b - a of the leaf, plus kBsrPairBit when this leaf is the left leaf
// of a two-leaf node computed here (one whose parent has 5 leaves, or the
// root), or kBsrQuadBit (and kBsrFourBit for 4 leaves) when it is the first
// leaf of a node of 3 or 4 leaves (height 3: a leaf and a two-leaf node, or
// two two-leaf nodes) computed here; ids = the value ids of P, Q, R of the
// leaf or node, or kBsrSkip for a leaf that another entry covers.
constexpr uint32_t kBsrPairBit = 0x80000000u;
constexpr uint32_t kBsrQuadBit = 0x40000000u;
constexpr uint32_t kBsrFourBit = 0x20000000u;
constexpr uint32_t kBsrRelMask = 0x1FFFFFFFu;
constexpr uint32_t kBsrSkip = 0xFFFFFFFFu;

__host__ __device__ inline void bsr_base_entry(const uint32_t* rel, uint32_t a, const uint32_t* ids,
                                               uint64_t i, BsrVal* meta, uint32_t* arena) {
  const uint32_t* id = ids + 3 * i;
  if (id[0] == kBsrSkip) return;
  const uint32_t rv = rel[i];
  const uint32_t b = a + (rv & kBsrRelMask);  // the entry's first leaf
  if (rv & kBsrQuadBit) {
    // [b - 1, b + 3) as (b, b+1) + (b+2, b+3), or [b - 1, b + 2) as
    // b + (b+1, b+2); a lone left leaf takes the pair-size arrays (zeros
    // above), so both shapes run the same code
    PairNode l, r;
    if (rv & kBsrFourBit) {
      l = bsr_pair_node(b);
      r = bsr_pair_node(b + 2);
    } else {
      const LeafNode v = bsr_leaf_node(b);
#pragma unroll
      for (int k = 0; k < 11; ++k) l.P[k] = k < 5 ? v.P[k] : 0u;
#pragma unroll
      for (int k = 0; k < 10; ++k) l.Q[k] = k < 5 ? v.Q[k] : 0u;
#pragma unroll
      for (int k = 0; k < 6; ++k) l.R[k] = k < 3 ? v.R[k] : 0u;
      l.sP = v.sP;
      r = bsr_pair_node(b + 1);
    }
    QuadNode z;
    bsr_combine(l, r, z);
    bsr_put_node(meta, id, arena, z);
  } else if (rv & kBsrPairBit) {
    bsr_put_node(meta, id, arena, bsr_pair_node(b));
  } else {
    bsr_put_node(meta, id, arena, bsr_leaf_node(b));
  }
}

// The subtree's first index a comes from device memory (set before each
// evaluation), so that a captured schedule serves every subtree of its plan.
// Launched with kBsrBaseThreads threads per block: the bound lets an entry's
// node arithmetic stay (almost) in registers, where the default 1024-thread
// bound capped it at 128 and spilled 82 registers.  Three waves per SIMD
// (168 registers, 8 of them spilled) measured faster than two without
// spills (173 registers).
constexpr int kBsrBaseThreads = 256;
__global__ __launch_bounds__(kBsrBaseThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_bsr_base(
    const uint32_t* rel, const uint32_t* pa, const uint32_t* ids, uint64_t n, BsrVal* meta, uint32_t* arena) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) bsr_base_entry(rel, *pa, ids, i, meta, arena);
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

// ------------------------------------------------ chunk-parallel large add
// The same addition for values above 1024 limbs, with every 1024-limb chunk
// of every node in its own workgroup: the comparison of equal-length
// magnitudes (k_bsr_add_cmp), block-local sums with carry-in 0 and the
// chunk carry functions (k_bsr_add_local), one wave per node scanning them
// and adding the carry-ins (k_bsr_add_scan; above 64 chunks k_bsr_add_scan_wide
// and the carry-in ripple k_bsr_add_fix), the top nonzero
// limb of a difference with the result length and sign (k_bsr_add_finish,
// which also clears the node's slot).  Decisions are recomputed from the metadata in each
// kernel (bsr_add_plan), so they agree everywhere.
struct BsrAddSlot {
  unsigned long long cmp;  // 1 + highest index where equal-length magnitudes differ
  unsigned long long top;  // 1 + highest nonzero limb of a difference
  uint64_t carry;          // carry out of the top limb of a sum
};

struct BsrAddPlan {
  const uint32_t* A;
  const uint32_t* B;
  uint32_t* O;
  uint32_t LA, LB;
  int sign;
  bool sub, zero;
};

__device__ __forceinline__ BsrAddPlan bsr_add_plan(const BsrAdd op, const BsrVal* meta, uint32_t* arena,
                                                   unsigned long long cmp_slot) {
  const BsrVal mx = meta[op.x], my = meta[op.y];
  BsrAddPlan p;
  p.O = arena + meta[op.out].off;
  const uint32_t* X = arena + mx.off;
  const uint32_t* Y = arena + my.off;
  p.A = X;
  p.B = Y;
  p.LA = mx.L;
  p.LB = my.L;
  p.sign = mx.sign;
  p.sub = mx.sign != my.sign;
  p.zero = false;
  bool swap = false;
  if (p.sub) {
    int cmp;
    if (mx.L != my.L) {
      cmp = mx.L > my.L ? 1 : -1;
    } else if (cmp_slot == 0) {
      cmp = 0;
    } else {
      const uint64_t i = cmp_slot - 1;
      cmp = X[i] > Y[i] ? 1 : -1;
    }
    if (cmp == 0) {
      p.zero = true;
      return p;
    }
    if (cmp < 0) {
      swap = true;
      p.sign = my.sign;
    }
  } else if (my.L > mx.L) {
    swap = true;
  }
  if (swap) {
    p.A = Y;
    p.B = X;
    p.LA = my.L;
    p.LB = mx.L;
  }
  return p;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long o = __shfl_xor(v, d, 64);
    v = o > v ? o : v;
  }
  return v;
}

__global__ __launch_bounds__(256) void k_bsr_add_cmp(const BsrAdd* adds, const BsrVal* meta,
                                                     const uint32_t* arena, BsrAddSlot* slots) {
  const BsrAdd op = adds[blockIdx.y];
  const BsrVal mx = meta[op.x], my = meta[op.y];
  if (mx.sign == my.sign || mx.L != my.L) return;
  const uint64_t base = (uint64_t)blockIdx.x * kCarryBlock;
  if (base >= mx.L) return;
  const uint32_t* X = arena + mx.off;
  const uint32_t* Y = arena + my.off;
  unsigned long long hit = 0;
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    const uint64_t k = base + (uint64_t)threadIdx.x * kCarryLPT + q;
    if (k < mx.L && X[k] != Y[k]) hit = k + 1;
  }
  hit = wave_max_u64(hit);
  if ((threadIdx.x & 63) == 0 && hit) atomicMax(&slots[blockIdx.y].cmp, hit);
}

__global__ __launch_bounds__(kCarryThreads) void k_bsr_add_local(const BsrAdd* adds, const BsrVal* meta,
                                                                 uint32_t* arena, const BsrAddSlot* slots,
                                                                 CarryFn* agg, uint32_t maxch) {
  __shared__ CarryFn wtot[kCarryThreads / 64];
  const uint32_t node = blockIdx.y;
  const BsrAddPlan pl = bsr_add_plan(adds[node], meta, arena, slots[node].cmp);
  const uint64_t base = (uint64_t)blockIdx.x * kCarryBlock;
  if (pl.zero || base >= pl.LA) return;  // uniform over the block
  const int tid = threadIdx.x;
  uint64_t s[kCarryLPT];
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    const uint64_t k = base + (uint64_t)tid * kCarryLPT + q;
    uint64_t x = 0;
    if (k < pl.LA) {
      const uint32_t bk = k < pl.LB ? pl.B[k] : 0u;
      x = (uint64_t)pl.A[k] + (pl.sub ? (uint32_t)~bk : bk);
      if (pl.sub && k == 0) x += 1;
    }
    s[q] = x;
  }
  // Positions past the top pass the carry through unchanged (identity), so
  // the block function ends at the top limb: its value is the carry out.
  // sums below 2^33: carries of at most 2, 32-bit limb functions (the
  // scan's LDS scratch taken from wtot)
  LimbFn f = lf_identity();
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    if (base + (uint64_t)tid * kCarryLPT + q >= pl.LA) continue;
    f = lf_compose(f, lf_of_sum(s[q]));
  }
  const LimbFn pre = lf_block_exclusive(f, reinterpret_cast<LimbFn*>(wtot));
  uint64_t c = lf_eval0(pre);
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    const uint64_t k = base + (uint64_t)tid * kCarryLPT + q;
    const uint64_t t = (s[q] & 0xffffffffull) + c;
    if (k < pl.LA) pl.O[k] = (uint32_t)t;
    c = (s[q] >> 32) + (t >> 32);
  }
  if (tid == kCarryThreads - 1) agg[(uint64_t)node * maxch + blockIdx.x] = cf_of_limb(lf_compose(pre, f));
}

// One wave per node: exclusive scan of its chunk functions; the total is
// the carry out of the top limb.
__global__ __launch_bounds__(64) void k_bsr_add_scan(const BsrAdd* adds, const BsrVal* meta, uint32_t* arena,
                                                     BsrAddSlot* slots, const CarryFn* agg, uint64_t* cin,
                                                     uint32_t maxch) {
  const uint32_t node = blockIdx.x;
  const BsrAddPlan pl = bsr_add_plan(adds[node], meta, arena, slots[node].cmp);
  if (pl.zero) return;
  const uint32_t nc = (uint32_t)((pl.LA + kCarryBlock - 1) / kCarryBlock);
  const int lane = threadIdx.x;
  CarryFn run;
  run.base = 0;
  run.T = 0;
  for (uint32_t c0 = 0; c0 < nc; c0 += 64) {
    const uint32_t idx = c0 + lane;
    CarryFn f;
    f.base = 0;
    f.T = 0;
    if (idx < nc) f = agg[(uint64_t)node * maxch + idx];
    CarryFn incl = f;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const CarryFn o = cf_shfl_up(incl, d);
      if (lane >= d) incl = cf_compose(o, incl);
    }
    CarryFn excl = cf_shfl_up(incl, 1);
    if (lane == 0) excl.T = 0;
    if (idx < nc) {
      // the chunk's carry-in, rippled into its limbs here (as k_bsr_add_fix)
      uint64_t c = cf_eval0(cf_compose(run, excl));
      uint64_t k = (uint64_t)idx * kCarryBlock;
      const uint64_t end = k + kCarryBlock < pl.LA ? k + kCarryBlock : pl.LA;
      while (c != 0 && k < end) {
        const uint64_t v = (uint64_t)pl.O[k] + (c & 0xffffffffull);
        pl.O[k] = (uint32_t)v;
        c = (c >> 32) + (v >> 32);
        ++k;
      }
    }
    CarryFn tot;
    tot.base = __shfl(incl.base, 63, 64);
    tot.T = __shfl(incl.T, 63, 64);
    run = cf_compose(run, tot);
  }
  if (lane == 0) slots[node].carry = cf_eval0(run);
}

// The same scan with one workgroup per node, for adds of more than 64
// chunks (cf_scan_runs).
__global__ __launch_bounds__(256) void k_bsr_add_scan_wide(const BsrAdd* adds, const BsrVal* meta,
                                                          uint32_t* arena, BsrAddSlot* slots,
                                                          const CarryFn* agg, uint64_t* cin, uint32_t maxch) {
  __shared__ CarryFn wtot[4], tot;
  const uint32_t node = blockIdx.x;
  const BsrAddPlan pl = bsr_add_plan(adds[node], meta, arena, slots[node].cmp);
  if (pl.zero) return;  // the same for the whole workgroup
  const uint32_t nc = (uint32_t)((pl.LA + kCarryBlock - 1) / kCarryBlock);
  const CarryFn t = cf_scan_runs(agg + (uint64_t)node * maxch, nc, cin + (uint64_t)node * maxch, wtot, &tot);
  if (threadIdx.x == 0) slots[node].carry = cf_eval0(t);
}

__global__ __launch_bounds__(256) void k_bsr_add_fix(const BsrAdd* adds, const BsrVal* meta, uint32_t* arena,
                                                     const BsrAddSlot* slots, const uint64_t* cin,
                                                     uint32_t maxch, uint64_t n) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t node = t / maxch, ch = t % maxch;
  if (node >= n) return;
  const BsrAddPlan pl = bsr_add_plan(adds[node], meta, arena, slots[node].cmp);
  uint64_t k = ch * kCarryBlock;
  if (pl.zero || k >= pl.LA) return;
  uint64_t c = cin[node * maxch + ch];
  uint64_t end = k + kCarryBlock;
  if (end > pl.LA) end = pl.LA;
  while (c != 0 && k < end) {
    const uint64_t v = (uint64_t)pl.O[k] + (c & 0xffffffffull);
    pl.O[k] = (uint32_t)v;
    c = (c >> 32) + (v >> 32);
    ++k;
  }
}

// One workgroup (kCarryThreads) per node, last: the length and sign of the
// result — a difference's top nonzero limb, scanning 1024-limb chunks from
// the top down (usually the top chunk holds it), or a sum's carry limb —
// then the node's slot is cleared for the next height (the slots start
// zeroed when the plan is laid out).
__global__ __launch_bounds__(kCarryThreads) void k_bsr_add_finish(const BsrAdd* adds, BsrVal* meta,
                                                                  uint32_t* arena, BsrAddSlot* slots) {
  __shared__ unsigned long long sh_hit;
  const uint32_t node = blockIdx.x;
  const BsrAdd op = adds[node];
  const BsrAddPlan pl = bsr_add_plan(op, meta, arena, slots[node].cmp);
  const uint64_t carry = slots[node].carry;
  uint64_t L = 0;
  if (pl.zero) {
    L = 0;
  } else if (pl.sub) {  // the same for the whole workgroup
    for (uint64_t end = pl.LA; end > 0;) {
      const uint64_t base = end > kCarryBlock ? end - kCarryBlock : 0;
      if (threadIdx.x == 0) sh_hit = 0;
      __syncthreads();
      unsigned long long hit = 0;
#pragma unroll
      for (int q = 0; q < kCarryLPT; ++q) {
        const uint64_t k = base + (uint64_t)threadIdx.x * kCarryLPT + q;
        if (k < end && pl.O[k] != 0) hit = k + 1;
      }
      hit = wave_max_u64(hit);
      if ((threadIdx.x & 63) == 0 && hit) atomicMax(&sh_hit, hit);
      __syncthreads();
      const unsigned long long h = sh_hit;
      __syncthreads();  // every thread has read sh_hit before the next reset
      if (h) {
        L = h;  // every leading zero stripped (usub)
        break;
      }
      end = base;
    }
  } else {
    L = pl.LA;
    if (carry) {
      if (threadIdx.x == 0) pl.O[L] = 1;  // carry limb (uadd)
      ++L;
    }
  }
  __syncthreads();  // every thread has read the slot
  if (threadIdx.x == 0) {
    BsrVal& mo = meta[op.out];
    mo.L = (uint32_t)L;
    mo.sign = L ? pl.sign : 1;
    slots[node].cmp = 0;
    slots[node].top = 0;
    slots[node].carry = 0;
  }
}

// ------------------------------------------------------------------ layout
// A subtree of d leaves splits at m = (a + b) / 2 (PI.h:90), so a node of s
// leaves has children of floor(s / 2) and s - floor(s / 2) leaves, and
// everything the plan records about a node (value bounds, arena slots,
// height, the size classes of its products) depends on s alone.  The sizes
// that occur are few (at most two per depth), so the plan is a small table
// of them, and each node's entries are computed from its preorder id t by
// walking down from the root:
//   - node ids are preorder: the left child of t is t + 1, the right one
//     t + 2 floor(s / 2); the arena holds the nodes in id order, each node's
//     five values (P Q R x y) in a row, so a node's offset is the arena of
//     everything before it in preorder;
//   - the products and adds of a height are grouped by class, each group in
//     id order; the nodes of one height are disjoint subtrees, so the nodes
//     before t in a group are those of the left siblings met on the walk,
//     counted per size with the table's subtree counts.
constexpr int kBsrPeers = 4;  // most distinct sizes of one height

struct BsrSizeInfo {
  uint64_t s;         // leaves
  uint64_t own;       // arena limbs of the node's five values (slot + 1 each)
  uint64_t arena;     // arena limbs of the whole subtree
  uint64_t pre[5];    // offset of value j from the node's first limb
  int32_t l, r;       // size index of the children, -1 for a leaf
  int32_t npeer;      // sizes of the same height (this one included)
  int32_t peer[kBsrPeers];
  uint64_t mbase[4];  // first index (in the plan's product list) of product j's group
  uint32_t mloc[4];   // products of the same node ahead of product j in its group
  uint32_t mmul[4][kBsrPeers];  // products a node of size peer[p] puts in product j's group
  uint64_t abase;     // first index of the node's add group (small or large)
  uint32_t amul[kBsrPeers];     // 1 if size peer[p] is in the same add group
};

struct BsrLayout {
  const BsrSizeInfo* info;
  const uint32_t* cnt;  // cnt[i * ns + q]: nodes of size q in a subtree of size i
  int32_t ns, root;
  uint64_t nodes, nmeta, nleaf, nmul, nadd;
  BsrVal* meta;
  uint32_t* leaf_rel;
  uint32_t* leaf_ids;
  BsrMul* muls;
  BsrAdd* adds;
  uint32_t* err;  // k_bsr_layout sets it when bsr_layout_node fails
};

// Writes every plan entry of node t (0 <= t < nodes); false if an index
// fell outside its list (nothing is written there).  Host and device.
__host__ __device__ inline bool bsr_layout_node(const BsrLayout& L, uint64_t t) {
  const BsrSizeInfo* I = L.info;
  // walk 1: the node's size, range start and arena offset
  int32_t si = L.root;
  uint64_t id = 0, a = 0, base = 0;
  uint64_t pid = 0, ps = 0, pa = 0;  // parent's id, size and first leaf
  uint64_t gid = 0, gs = 0, ga = 0;  // grandparent's (size 0: none)
  while (id != t) {
    const BsrSizeInfo& n = I[si];
    const uint64_t sl = I[n.l].s;
    gid = pid;
    gs = ps;
    ga = pa;
    pid = id;
    ps = n.s;
    pa = a;
    if (t < id + 2 * sl) {
      id += 1;
      base += n.own;
      si = n.l;
    } else {
      id += 2 * sl;
      base += n.own + I[n.l].arena;
      a += sl;
      si = n.r;
    }
  }
  const BsrSizeInfo& me = I[si];
  bool bad = false;
  for (int j = 0; j < 5; ++j) {
    const uint64_t v = 5 * t + j;
    if (v < L.nmeta) {
      L.meta[v].off = base + me.pre[j];
      L.meta[v].L = 0;
      L.meta[v].sign = 1;
    } else {
      bad = true;
    }
  }
  if (me.l < 0) {  // leaf [a, a + 1): an entry of k_bsr_base
    if (a < L.nleaf) {
      // a leaf's parent has 2 or 3 leaves, a two-leaf node's 3, 4 or 5
      uint32_t rel = (uint32_t)(a + 1), id0 = (uint32_t)(5 * t);
      if (ps == 3) {  // the lone left leaf of a three-leaf node: the node
        rel |= kBsrQuadBit;
        id0 = (uint32_t)(5 * pid);
      } else if (ps == 2 && (gs == 3 || gs == 4)) {  // under a node of height 3
        if (a == ga) {
          rel |= kBsrQuadBit | (gs == 4 ? kBsrFourBit : 0u);
          id0 = (uint32_t)(5 * gid);
        } else {
          id0 = kBsrSkip;
        }
      } else if (ps == 2) {  // a two-leaf node under a five-leaf one, or the root
        if (t == pid + 1) {
          rel |= kBsrPairBit;
          id0 = (uint32_t)(5 * pid);
        } else {
          id0 = kBsrSkip;
        }
      }
      L.leaf_rel[a] = rel;
      for (int j = 0; j < 3; ++j) L.leaf_ids[3 * a + j] = id0 == kBsrSkip ? kBsrSkip : id0 + j;
    } else {
      bad = true;
    }
  } else {
    // walk 2: nodes of each size of this height ahead of t in preorder
    uint64_t before[kBsrPeers] = {0, 0, 0, 0};
    int32_t sj = L.root;
    uint64_t jd = 0;
    while (jd != t) {
      const BsrSizeInfo& n = I[sj];
      const uint64_t sl = I[n.l].s;
      if (t < jd + 2 * sl) {
        jd += 1;
        sj = n.l;
      } else {
        for (int p = 0; p < me.npeer; ++p) before[p] += L.cnt[(uint64_t)n.l * L.ns + me.peer[p]];
        jd += 2 * sl;
        sj = n.r;
      }
    }
    const uint32_t v = (uint32_t)(5 * t), lf = (uint32_t)(5 * (t + 1)),
                   rt = (uint32_t)(5 * (t + 2 * I[me.l].s));
    // x = P0 Q1, Q = Q0 Q1, y = P1 R0, R = R0 R1 (PI.h:110-112), as two
    // pairs around the operands used twice
    const BsrMul m4[4] = {{rt + 1, lf + 0, v + 3}, {rt + 1, lf + 1, v + 1},
                          {lf + 2, rt + 0, v + 4}, {lf + 2, rt + 2, v + 2}};
    for (int j = 0; j < 4; ++j) {
      uint64_t k = me.mbase[j] + me.mloc[j];
      for (int p = 0; p < me.npeer; ++p) k += before[p] * me.mmul[j][p];
      if (k < L.nmul)
        L.muls[k] = m4[j];
      else
        bad = true;
    }
    uint64_t k = me.abase;
    for (int p = 0; p < me.npeer; ++p) k += before[p] * me.amul[p];
    if (k < L.nadd)
      L.adds[k] = BsrAdd{v + 3, v + 4, v + 0};  // P = x.add(y)
    else
      bad = true;
  }
  return !bad;
}

// The size table is read on every step of both walks: a copy in LDS when it
// has at most kBsrLdsSizes sizes (subtrees below 2^31 leaves have at most
// two per depth).
constexpr int kBsrLdsSizes = 64;

__global__ __launch_bounds__(256) void k_bsr_layout(BsrLayout L) {
  __shared__ BsrSizeInfo s_info[kBsrLdsSizes];
  __shared__ uint32_t s_cnt[kBsrLdsSizes * kBsrLdsSizes];
  if (L.ns <= kBsrLdsSizes) {
    for (int i = threadIdx.x; i < L.ns; i += blockDim.x) s_info[i] = L.info[i];
    for (int i = threadIdx.x; i < L.ns * L.ns; i += blockDim.x) s_cnt[i] = L.cnt[i];
    __syncthreads();
    L.info = s_info;
    L.cnt = s_cnt;
  }
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < L.nodes; t += stride)
    if (!bsr_layout_node(L, t)) atomicOr(L.err, 1u);
}

}  // namespace pcg
