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
  constexpr int N = 1 << LOGN;
  constexpr int LP = GE::LP;  // padded line length
  // [0, LP): operand a's line; [LP, 2 LP): operand b's; [2 LP, 2 LP + N):
  // prime-0 results (natural order)
  __shared__ double lds[2 * LP + N];
  __shared__ CarryFn wtot[4];
  __shared__ double2 ltw[LOGN <= kLogLdsTw ? (1 << LOGN) : 1];
  const BatchDesc d = descs[blockIdx.x];
  const int tid = threadIdx.x, nth = blockDim.x;
  const int na = (int)d.na, nb = (int)d.nb, zL = na + nb;
  double* la = lds;
  double* lb = lds + LP;
  double* r0 = lds + 2 * LP;
  for (int pi = 0; pi < 2; ++pi) {
    const Prime P = prm[pi];
    const double2* tw = stage_twiddles<GE>(ltw, P);
    for (int i = tid; i < N; i += nth) {
      la[GE::idx(i, 0)] = i < na ? (double)d.a[i] : 0.0;
      lb[GE::idx(i, 0)] = i < nb ? (double)d.b[i] : 0.0;
    }
    __syncthreads();
    line_dft<GE>(la, tw, P, false);
    line_dft<GE>(lb, tw, P, false);
    for (int i = tid; i < N; i += nth) {
      const int li = GE::idx(i, 0);
      const double x = mm_gen(la[li], lb[li], P.p, P.pinv);
      la[li] = mm_shoup(x, P.ninv.x, P.ninv.y, P.p);
    }
    __syncthreads();
    line_dft<GE>(la, tw, P, true);
    if (pi == 0) {
#pragma unroll
      for (int q = 0; q < kEPT; ++q) r0[tid * kEPT + q] = la[GE::idx(tid * kEPT + q, 0)];
      __syncthreads();
    }
  }
  const Prime P0 = prm[0], P1 = prm[1];
  uint32_t w0[kEPT], w1[kEPT], w2[kEPT], D[kEPT];
#pragma unroll
  for (int q = 0; q < kEPT; ++q) {
    const int k = tid * kEPT + q;
    w0[q] = w1[q] = w2[q] = 0;
    if (k < zL) crt_words(r0[k], la[GE::idx(k, 0)], P0, P1, crt, w0[q], w1[q], w2[q]);
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

// CRT and carry of batched multi-pass products: one workgroup per product
// walks its limbs in chunks of 1024, folding each chunk's carry-in into the
// chunk's first limbs.  r: inverse-transform outputs of product z at
// r + z * zstride, prime q at offset q * N.
__global__ __launch_bounds__(256) void k_batch_carry(const BatchDesc* descs, const double* r,
                                                      uint64_t zstride, uint64_t N,
                                                      const Prime* prm, double2 crt) {
  __shared__ uint32_t sw1[kCarryBlock], sw2[kCarryBlock];
  __shared__ CarryFn wtot[4];
  __shared__ uint64_t next_cin;
  const BatchDesc d = descs[blockIdx.x];
  const double* rz = r + (uint64_t)blockIdx.x * zstride;
  const Prime P0 = prm[0], P1 = prm[1];
  const uint64_t zL = (uint64_t)d.na + d.nb;
  const int tid = threadIdx.x;
  uint64_t cin = 0;
  for (uint64_t base = 0; base < zL; base += kCarryBlock) {
    uint32_t w0[kCarryLPT], w1[kCarryLPT], w2[kCarryLPT], D[kCarryLPT];
#pragma unroll
    for (int q = 0; q < kCarryLPT; ++q) {
      const uint64_t k = base + tid * kCarryLPT + q;
      w0[q] = w1[q] = w2[q] = 0;
      if (k < zL) crt_words(rz[k], rz[N + k], P0, P1, crt, w0[q], w1[q], w2[q]);
    }
    uint64_t c, spill;
    block_carry<kCarryLPT>(w0, w1, w2, sw1, sw2, wtot, D, c, spill, cin);
#pragma unroll
    for (int q = 0; q < kCarryLPT; ++q) {
      const uint64_t k = base + tid * kCarryLPT + q;
      if (k < zL) d.out[k] = D[q];
    }
    if (tid == kCarryThreads - 1) next_cin = spill + c;
    __syncthreads();
    cin = next_cin;
    __syncthreads();
  }
}

}  // namespace pcg
