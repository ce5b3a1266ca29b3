// Batched products: one workgroup per product, carries resolved inside the
// workgroup, so a whole size class of a tree level is one launch.
//
//  k_batch_school : zL = na + nb <= 1024 limbs.  Thread t forms the column
//                   sums of limbs 4t..4t+3 (exact 96-bit accumulation of
//                   32x32 products) from operands staged in LDS.
//  k_batch_ntt    : zL <= 4096 limbs.  Two-prime NTT of N = 2^ceil(log2 zL)
//                   points entirely in LDS (operands as two contiguous lines),
//                   primes one after the other, CRT, carry.
#pragma once
#include "carry_kernels.hpp"
#include "ntt_kernels.hpp"

#pragma clang fp contract(off)

namespace pcg {

struct BatchDesc {
  const uint32_t* a;
  const uint32_t* b;
  uint32_t* out;
  uint32_t na, nb;
};

constexpr int kSchoolMaxLimbs = 1024;  // zL bound of k_batch_school
constexpr int kSmallNttMaxLog = 12;    // N bound of k_batch_ntt (rows padded by one)

__global__ __launch_bounds__(256) void k_batch_school(const BatchDesc* descs) {
  __shared__ uint32_t sa[kSchoolMaxLimbs], sb[kSchoolMaxLimbs];
  __shared__ uint32_t sw1[kSchoolMaxLimbs], sw2[kSchoolMaxLimbs];
  __shared__ CarryFn wtot[4];
  const BatchDesc d = descs[blockIdx.x];
  const int tid = threadIdx.x;
  const int na = (int)d.na, nb = (int)d.nb, zL = na + nb;
  for (int i = tid; i < na; i += 256) sa[i] = d.a[i];
  for (int i = tid; i < nb; i += 256) sb[i] = d.b[i];
  __syncthreads();
  uint32_t w0[4], w1[4], w2[4], D[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = 4 * tid + q;
    int i0 = k - nb + 1;
    if (i0 < 0) i0 = 0;
    int i1 = k < na - 1 ? k : na - 1;
    uint64_t lo = 0;
    uint32_t hi = 0;
    for (int i = i0; i <= i1; ++i) {
      const uint64_t pr = (uint64_t)sa[i] * sb[k - i];
      lo += pr;
      hi += lo < pr;
    }
    w0[q] = (uint32_t)lo;
    w1[q] = (uint32_t)(lo >> 32);
    w2[q] = hi;
  }
  uint64_t c, spill;
  block_carry<4>(w0, w1, w2, sw1, sw2, wtot, D, c, spill);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = 4 * tid + q;
    if (k < zL) d.out[k] = D[q];
  }
}

// Block size N / 16 (N = 2^logN, 2^10 <= N <= 2^12); all products of one
// launch share logN.  prm points at the constants for this N.
template <int LOGN>
__global__ __launch_bounds__((1 << LOGN) / kEPT) void k_batch_ntt(const BatchDesc* descs,
                                                                  const Prime* prm, double2 crt) {
  using GE = Geo<LOGN, 0, true>;
  constexpr int logN = LOGN;
  // [0, 2N): the two operand lines; [2N, 3N): prime-0 results
  __shared__ double lds[3 << LOGN];
  __shared__ CarryFn wtot[4];
  const BatchDesc d = descs[blockIdx.x];
  const int tid = threadIdx.x, nth = blockDim.x;
  const int N = 1 << logN;
  const int na = (int)d.na, nb = (int)d.nb, zL = na + nb;
  for (int pi = 0; pi < 2; ++pi) {
    const Prime P = prm[pi];
    for (int i = tid; i < N; i += nth) {
      lds[i] = i < na ? (double)d.a[i] : 0.0;
      lds[N + i] = i < nb ? (double)d.b[i] : 0.0;
    }
    __syncthreads();
    line_dft<GE>(lds, P, false);
    line_dft<GE>(lds + N, P, false);
    for (int i = tid; i < N; i += nth) {
      const double x = mm_gen(lds[i], lds[N + i], P.p, P.pinv);
      lds[i] = mm_shoup(x, P.ninv.x, P.ninv.y, P.p);
    }
    __syncthreads();
    line_dft<GE>(lds, P, true);
    if (pi == 0) {
#pragma unroll
      for (int q = 0; q < kEPT; ++q) lds[2 * N + tid * kEPT + q] = lds[tid * kEPT + q];
      __syncthreads();
    }
  }
  const Prime P0 = prm[0], P1 = prm[1];
  uint32_t w0[kEPT], w1[kEPT], w2[kEPT], D[kEPT];
#pragma unroll
  for (int q = 0; q < kEPT; ++q) {
    const int k = tid * kEPT + q;
    w0[q] = w1[q] = w2[q] = 0;
    if (k < zL)
      crt_words(lds[2 * N + tid * kEPT + q], lds[tid * kEPT + q], P0, P1, crt, w0[q], w1[q], w2[q]);
  }
  __syncthreads();  // sw1/sw2 below reuse the operand lines
  uint32_t* sw1 = reinterpret_cast<uint32_t*>(lds);
  uint32_t* sw2 = sw1 + N;
  uint64_t c, spill;
  block_carry<kEPT>(w0, w1, w2, sw1, sw2, wtot, D, c, spill);
#pragma unroll
  for (int q = 0; q < kEPT; ++q) {
    const int k = tid * kEPT + q;
    if (k < zL) d.out[k] = D[q];
  }
}

}  // namespace pcg
