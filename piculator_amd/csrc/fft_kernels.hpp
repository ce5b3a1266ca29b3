// FP64 complex FFT multiply (second engine), gfx950.
//
// Digits: each limb is split into two balanced 16-bit digits (values in
// [-2^15, 2^15], the top digit of an operand unbalanced), which keeps the
// convolution coefficients near sqrt(count) * 2^30 instead of count * 2^32.
// Right-angle convolution: a digit sequence x of length 2N is packed as
// X_j = x_j + i x_{j+N}, weighted by w^j with w = e^{i pi / 2N} (so w^N = i),
// transformed with an N-point complex FFT; the inverse transform unweighted
// by w^{-j} / N gives the negacyclic convolution of length 2N, real parts
// digits [0, N) and imaginary parts digits [N, 2N).  With N >= na + nb
// limbs there is no wrap-around, so the result is the exact product once
// every coefficient is rounded to the nearest integer.  The largest distance
// to that integer is reported; above 0.25 the caller recomputes on the NTT.
//
// Transform passes mirror ntt_kernels.hpp (same decimation and geometry)
// with complex arithmetic and 2048 points staged per workgroup.
#pragma once
#include "modfp64.hpp"

#pragma clang fp contract(off)

namespace pcg {

constexpr int kFftLogE = 11;
constexpr int kFftE = 1 << kFftLogE;
constexpr int kFftLogStg = 12;  // stage twiddle table: e^{2 pi i e / 4096}
constexpr int kEPTc = 8;        // points per thread per stage

struct FftTabs {
  const double2* stg;  // e^{2 pi i e / 4096}, e < 4096
  const double2* hi;   // e^{2 pi i (e << 16) / 2^32}, e < 2^16
  const double2* lo;   // e^{2 pi i e / 2^32}, e < 2^16
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by +i (inv == false) or -i (inv == true)
__device__ __forceinline__ double2 cmul_i(double2 a, bool inv) {
  return inv ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// e^{+-2 pi i e / 2^logM}, e < 2^logM, logM <= 32
__device__ __forceinline__ double2 croot(uint64_t e, int logM, bool inv, const FftTabs& t) {
  const uint64_t E = (e << (32 - logM)) & 0xffffffffull;
  const double2 w = cmul(t.hi[E >> 16], t.lo[E & 0xffff]);
  return inv ? cconj(w) : w;
}

template <int R>
__device__ __forceinline__ void cdft_core(double2* v, bool inv);

template <>
__device__ __forceinline__ void cdft_core<2>(double2* v, bool inv) {
  const double2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void cdft_core<4>(double2* v, bool inv) {
  const double2 e0 = cadd(v[0], v[2]), e1 = csub(v[0], v[2]);
  const double2 o0 = cadd(v[1], v[3]), o1 = cmul_i(csub(v[1], v[3]), inv);
  v[0] = cadd(e0, o0);
  v[2] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[3] = csub(e1, o1);
}
template <>
__device__ __forceinline__ void cdft_core<8>(double2* v, bool inv) {
  const double c = 0.70710678118654752440;
  const double s = inv ? -c : c;
  const double2 w1 = make_double2(c, s), w3 = make_double2(-c, s);
  const double2 e0 = cadd(v[0], v[4]), e1 = csub(v[0], v[4]);
  const double2 o0 = cadd(v[2], v[6]), o1 = cmul_i(csub(v[2], v[6]), inv);
  const double2 E0 = cadd(e0, o0), E2 = csub(e0, o0), E1 = cadd(e1, o1), E3 = csub(e1, o1);
  const double2 f0 = cadd(v[1], v[5]), f1 = csub(v[1], v[5]);
  const double2 g0 = cadd(v[3], v[7]), g1 = cmul_i(csub(v[3], v[7]), inv);
  double2 O0 = cadd(f0, g0), O2 = csub(f0, g0), O1 = cadd(f1, g1), O3 = csub(f1, g1);
  O1 = cmul(O1, w1);
  O2 = cmul_i(O2, inv);
  O3 = cmul(O3, w3);
  v[0] = cadd(E0, O0);
  v[4] = csub(E0, O0);
  v[1] = cadd(E1, O1);
  v[5] = csub(E1, O1);
  v[2] = cadd(E2, O2);
  v[6] = csub(E2, O2);
  v[3] = cadd(E3, O3);
  v[7] = csub(E3, O3);
}

template <int R, int LOGR>
__device__ __forceinline__ void cstockham_stage(double2* lds, int logL, int logLines, int logNs,
                                                const FftTabs& tb, bool inv) {
  constexpr int NB = kEPTc / R;
  const int nth = blockDim.x, tid = threadIdx.x;
  const int linesMask = (1 << logLines) - 1;
  const int logLR = logL - LOGR;
  const int NsMask = (1 << logNs) - 1;
  const int sh = kFftLogStg - logNs - LOGR;
  double2 v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int bf = tid + b * nth;
    const int lam = bf & linesMask;
    const int j = bf >> logLines;
#pragma unroll
    for (int r = 0; r < R; ++r) v[b][r] = lds[((j + (r << logLR)) << logLines) + lam];
    if (logNs > 0) {
      const int jm = j & NsMask;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int e = (jm * r) << sh;
        double2 t = tb.stg[e];
        if (inv) t.y = -t.y;
        v[b][r] = cmul(v[b][r], t);
      }
    }
    cdft_core<R>(v[b], inv);
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

__device__ __forceinline__ void cline_dft(double2* lds, int logL, int logLines, const FftTabs& tb,
                                          bool inv) {
  int logNs = 0;
  const int first = logL % 3;
  if (first == 1) {
    cstockham_stage<2, 1>(lds, logL, logLines, 0, tb, inv);
    logNs = 1;
  } else if (first == 2) {
    cstockham_stage<4, 2>(lds, logL, logLines, 0, tb, inv);
    logNs = 2;
  }
  for (; logNs < logL; logNs += 3) cstockham_stage<8, 3>(lds, logL, logLines, logNs, tb, inv);
}

// Balanced 16-bit digit n of a limb vector (see the header comment).
__device__ __forceinline__ double bal_digit(const uint32_t* limbs, uint64_t nl, uint64_t n) {
  if (n >= 2 * nl) return 0.0;
  const uint32_t d = (limbs[n >> 1] >> (16 * (n & 1))) & 0xffffu;
  uint32_t borrow = 0;
  if (n >= 1) borrow = ((limbs[(n - 1) >> 1] >> (16 * ((n - 1) & 1))) & 0xffffu) >= 32768u;
  if (n == 2 * nl - 1) return (double)d + (double)borrow;
  return (double)d - (d >= 32768u ? 65536.0 : 0.0) + (double)borrow;
}

struct FftPassArgs {
  FftTabs tb;
  const double2* src;     // null -> digits from limbs (first forward pass)
  const uint32_t* limbs;
  uint64_t nlimbs;
  double2* dst;
  const double2* spec;    // fused: spectrum of a
  uint64_t N;
  int logN;
  int logL, logS, logLines, logG, logB;
  int unweight;           // last inverse pass: multiply by w^{-n} / N
};

template <int MODE>
__global__ __launch_bounds__(256) void k_fft_pass(FftPassArgs a) {
  __shared__ double2 lds[kFftE];
  const FftTabs tb = a.tb;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int logL = a.logL, logS = a.logS, logLines = a.logLines, logG = a.logG;
  const int Gmask = (1 << logG) - 1, Lmask = (1 << logL) - 1;
  const uint64_t ncb = 1ull << (logS - logG);
  const uint64_t w = blockIdx.x;
  const uint64_t blk0 = (w >> (logS - logG)) << a.logB;
  const uint64_t col0 = (w & (ncb - 1)) << logG;
  const int logRoot = logL + logS;
  constexpr int ept = kEPTc;
#pragma unroll
  for (int i = 0; i < ept; ++i) {
    const int f = tid + i * nth;
    const int col = f & Gmask;
    const int t = f >> logG;
    const int n = t & Lmask;
    const int b = t >> logL;
    const uint64_t g = ((blk0 + b) << logRoot) + ((uint64_t)n << logS) + col0 + col;
    double2 x;
    if (a.src) {
      x = a.src[g];
    } else {
      x = make_double2(bal_digit(a.limbs, a.nlimbs, g), bal_digit(a.limbs, a.nlimbs, g + a.N));
      x = cmul(x, croot(g, a.logN + 2, false, tb));  // right-angle weight w^g
    }
    if (MODE == kInv && logS > 0) x = cmul(x, croot((col0 + col) * (uint64_t)n, logRoot, true, tb));
    lds[(n << logLines) + (b << logG) + col] = x;
  }
  __syncthreads();
  cline_dft(lds, logL, logLines, tb, MODE == kInv);
  if (MODE == kFused) {
#pragma unroll
  for (int i = 0; i < ept; ++i) {
      const int f = tid + i * nth;
      const int col = f & Gmask;
      const int t = f >> logG;
      const int n = t & Lmask;
      const int b = t >> logL;
      const uint64_t g = ((blk0 + b) << logRoot) + ((uint64_t)n << logS) + col0 + col;
      const int li = (n << logLines) + (b << logG) + col;
      lds[li] = cmul(lds[li], a.spec[g]);
    }
    __syncthreads();
    cline_dft(lds, logL, logLines, tb, true);
  }
  const double scale = 1.0 / (double)a.N;
#pragma unroll
  for (int i = 0; i < ept; ++i) {
    const int f = tid + i * nth;
    const int col = f & Gmask;
    const int t = f >> logG;
    const int n = t & Lmask;
    const int b = t >> logL;
    const uint64_t g = ((blk0 + b) << logRoot) + ((uint64_t)n << logS) + col0 + col;
    double2 x = lds[(n << logLines) + (b << logG) + col];
    if (MODE == kFwd && logS > 0) x = cmul(x, croot((col0 + col) * (uint64_t)n, logRoot, false, tb));
    if (a.unweight) {
      x = cmul(x, croot(g, a.logN + 2, true, tb));
      x.x *= scale;
      x.y *= scale;
    }
    a.dst[g] = x;
  }
}

}  // namespace pcg
