// Two-prime FP64 NTT convolution kernels for gfx950.
//
// A transform of N = 2^n points is split into passes; pass j performs DFTs of
// length L_j = 2^{l_j} over "lines" of stride s_j = N / (L_1 ... L_j)
// (decimation in frequency, no reordering):
//   forward pass j : line DFT with root omega, then multiply output k of the
//                    line whose column is c by omega_{L_j s_j}^{c k};
//   inverse pass j : multiply by omega_{L_j s_j}^{-c k}, line DFT with
//                    omega^{-1}; passes run in reverse order.
// The spectrum is left in digit-reversed order, which the pointwise product
// does not care about, so no transpose or bit-reversal pass exists.
// The last forward pass of operand b (stride 1, contiguous lines) is fused
// with the pointwise product, the N^{-1} scaling and the first inverse pass.
//
// Each workgroup stages 4096 (or N if smaller) points of one prime in LDS --
// several adjacent columns for strided passes so that global accesses are
// contiguous runs -- and runs Stockham autosort radix-8 stages (one radix-2
// or radix-4 stage when l_j is not a multiple of 3) with each thread holding
// 16 points in registers per stage.
#pragma once
#include "modfp64.hpp"

#pragma clang fp contract(off)

namespace pcg {

constexpr int kLogE = 12;  // points staged per workgroup (max)
constexpr int kE = 1 << kLogE;
constexpr int kEPT = 16;   // points per thread per stage

enum PassMode { kFwd = 0, kInv = 1, kFused = 2 };

struct PassArgs {
  const Prime* prm;        // device: the two primes' constants for this N
  const double* src;       // [2][N] or null (then limbs)
  const uint32_t* limbs;   // operand limbs (first forward pass)
  uint64_t nlimbs;
  double* dst;             // [2][N]
  const double* spec;      // fused: spectrum of operand a, [2][N]
  uint64_t N;
  int logL, logS, logLines, logG, logB;
};

// -------------------------------------------------------------- radix cores
template <int R>
__device__ __forceinline__ void dft_core(double* v, const Prime& P, bool inv);

template <>
__device__ __forceinline__ void dft_core<2>(double* v, const Prime& P, bool inv) {
  double a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <>
__device__ __forceinline__ void dft_core<4>(double* v, const Prime& P, bool inv) {
  const double2 w4 = inv ? P.w8i[1] : P.w8f[1];
  double e0 = v[0] + v[2], e1 = v[0] - v[2];
  double o0 = v[1] + v[3];
  double o1 = mm_shoup(v[1] - v[3], w4.x, w4.y, P.p);
  v[0] = e0 + o0;
  v[2] = e0 - o0;
  v[1] = e1 + o1;
  v[3] = e1 - o1;
}

template <>
__device__ __forceinline__ void dft_core<8>(double* v, const Prime& P, bool inv) {
  const double2* w = inv ? P.w8i : P.w8f;  // w[0] = omega_8, w[1] = omega_4, w[2] = omega_8^3
  const double p = P.p;
  // DFT4 of the even inputs
  double e0 = v[0] + v[4], e1 = v[0] - v[4];
  double o0 = v[2] + v[6];
  double o1 = mm_shoup(v[2] - v[6], w[1].x, w[1].y, p);
  double E0 = e0 + o0, E2 = e0 - o0, E1 = e1 + o1, E3 = e1 - o1;
  // DFT4 of the odd inputs
  double f0 = v[1] + v[5], f1 = v[1] - v[5];
  double g0 = v[3] + v[7];
  double g1 = mm_shoup(v[3] - v[7], w[1].x, w[1].y, p);
  double O0 = f0 + g0, O2 = f0 - g0, O1 = f1 + g1, O3 = f1 - g1;
  O1 = mm_shoup(O1, w[0].x, w[0].y, p);
  O2 = mm_shoup(O2, w[1].x, w[1].y, p);
  O3 = mm_shoup(O3, w[2].x, w[2].y, p);
  v[0] = E0 + O0;
  v[4] = E0 - O0;
  v[1] = E1 + O1;
  v[5] = E1 - O1;
  v[2] = E2 + O2;
  v[6] = E2 - O2;
  v[3] = E3 + O3;
  v[7] = E3 - O3;
}

// omega_{2^logM}^{+-e} (e < 2^logM <= 2^30) from the two-level table, as (w, w/p)
__device__ __forceinline__ double2 root_pow(uint64_t e, int logM, bool inv, const Prime& P) {
  const uint64_t mask = (1ull << kLogRootMax) - 1;
  uint64_t E = (e << (kLogRootMax - logM)) & mask;
  if (inv) E = ((1ull << kLogRootMax) - E) & mask;
  const double2 h = P.twHi[E >> 15];
  const double2 l = P.twLo[E & 32767];
  double w = mm_shoup(h.x, l.x, l.y, P.p);
  return make_double2(w, w * P.pinv);
}

// One Stockham stage of radix R over all lines held in LDS.
// LDS layout: point n of line lam at lds[n * lines + lam].
template <int R, int LOGR>
__device__ __forceinline__ void stockham_stage(double* lds, int logL, int logLines, int logNs,
                                               const Prime& P, bool inv) {
  constexpr int NB = kEPT / R;
  const int nth = blockDim.x, tid = threadIdx.x;
  const int linesMask = (1 << logLines) - 1;
  const int logLR = logL - LOGR;
  const int NsMask = (1 << logNs) - 1;
  const int sh = kLogE - logNs - LOGR;  // 4096 / (Ns * R) as a shift
  double v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int bf = tid + b * nth;
    const int lam = bf & linesMask;
    const int j = bf >> logLines;
#pragma unroll
    for (int r = 0; r < R; ++r) v[b][r] = lds[((j + (r << logLR)) << logLines) + lam];
    if (logNs > 0) {
      const int jm = j & NsMask;
      v[b][0] = reduce(v[b][0], P.p, P.pinv);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        int e = (jm * r) << sh;
        if (inv) e = (kE - e) & (kE - 1);
        const double2 t = P.twL[e];
        v[b][r] = mm_shoup(v[b][r], t.x, t.y, P.p);
      }
    }
    dft_core<R>(v[b], P, inv);
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int bf = tid + b * nth;
    const int lam = bf & linesMask;
    const int j = bf >> logLines;
    const int jm = j & NsMask;
    const int idxD = ((j - jm) << LOGR) + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[((idxD + (r << logNs)) << logLines) + lam] = v[b][r];
  }
  __syncthreads();
}

// Natural-order DFT of length 2^logL of every line in LDS.
__device__ __forceinline__ void line_dft(double* lds, int logL, int logLines, const Prime& P,
                                         bool inv) {
  int logNs = 0;
  const int first = logL % 3;
  if (first == 1) {
    stockham_stage<2, 1>(lds, logL, logLines, 0, P, inv);
    logNs = 1;
  } else if (first == 2) {
    stockham_stage<4, 2>(lds, logL, logLines, 0, P, inv);
    logNs = 2;
  }
  for (; logNs < logL; logNs += 3) stockham_stage<8, 3>(lds, logL, logLines, logNs, P, inv);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_ntt_pass(PassArgs a) {
  __shared__ double lds[kE];
  const int pi = blockIdx.y;
  const Prime P = a.prm[pi];
  const int tid = threadIdx.x, nth = blockDim.x;
  const int logL = a.logL, logS = a.logS, logLines = a.logLines, logG = a.logG;
  const int Gmask = (1 << logG) - 1, Lmask = (1 << logL) - 1;
  const uint64_t ncb = 1ull << (logS - logG);  // column groups per block of L*s points
  const uint64_t w = blockIdx.x;
  const uint64_t blk0 = (w >> (logS - logG)) << a.logB;
  const uint64_t col0 = (w & (ncb - 1)) << logG;
  const double* src = a.src ? a.src + pi * a.N : nullptr;
  double* dst = a.dst + pi * a.N;
  const int logRoot = logL + logS;

#pragma unroll
  for (int i = 0; i < kEPT; ++i) {
    const int f = tid + i * nth;
    const int col = f & Gmask;
    const int t = f >> logG;
    const int n = t & Lmask;
    const int b = t >> logL;
    const uint64_t g = ((blk0 + b) << logRoot) + ((uint64_t)n << logS) + col0 + col;
    double x;
    if (src)
      x = src[g];
    else
      x = g < a.nlimbs ? (double)a.limbs[g] : 0.0;
    if (MODE == kInv && logS > 0) {
      const double2 tw = root_pow((col0 + col) * (uint64_t)n, logRoot, true, P);
      x = mm_shoup(x, tw.x, tw.y, P.p);
    }
    lds[(n << logLines) + (b << logG) + col] = x;
  }
  __syncthreads();
  line_dft(lds, logL, logLines, P, MODE == kInv);
  if (MODE == kFused) {
    const double* spec = a.spec + pi * a.N;
#pragma unroll
    for (int i = 0; i < kEPT; ++i) {
      const int f = tid + i * nth;
      const int col = f & Gmask;
      const int t = f >> logG;
      const int n = t & Lmask;
      const int b = t >> logL;
      const uint64_t g = ((blk0 + b) << logRoot) + ((uint64_t)n << logS) + col0 + col;
      const int li = (n << logLines) + (b << logG) + col;
      double x = mm_gen(lds[li], spec[g], P.p, P.pinv);
      lds[li] = mm_shoup(x, P.ninv.x, P.ninv.y, P.p);
    }
    __syncthreads();
    line_dft(lds, logL, logLines, P, true);
  }
#pragma unroll
  for (int i = 0; i < kEPT; ++i) {
    const int f = tid + i * nth;
    const int col = f & Gmask;
    const int t = f >> logG;
    const int n = t & Lmask;
    const int b = t >> logL;
    const uint64_t g = ((blk0 + b) << logRoot) + ((uint64_t)n << logS) + col0 + col;
    double x = lds[(n << logLines) + (b << logG) + col];
    if (MODE == kFwd && logS > 0) {
      const double2 tw = root_pow((col0 + col) * (uint64_t)n, logRoot, false, P);
      x = mm_shoup(x, tw.x, tw.y, P.p);
    } else {
      x = reduce(x, P.p, P.pinv);
    }
    dst[g] = x;
  }
}

}  // namespace pcg
