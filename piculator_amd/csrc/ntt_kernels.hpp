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
// kEPT (8) points in registers per stage.
#pragma once
#include "limb_buf.hpp"
#include "modfp64.hpp"

#pragma clang fp contract(off)

namespace pcg {

constexpr int kLogE = 12;  // points staged per workgroup (max)
constexpr int kE = 1 << kLogE;
constexpr int kEPT = 8;    // points per thread per stage

// kFusedA / kSquareA: the fused row pass of a large-tile plan that also runs
// operand a's row pass (k_ntt_row_fa); kSquareA when a and b are one operand.
enum PassMode { kFwd = 0, kInv = 1, kFused = 2, kFusedA = 3, kSquareA = 4 };

// One product of a batched launch (device pointers).
struct BatchDesc {
  const uint32_t* a;
  const uint32_t* b;
  uint32_t* out;
  uint32_t na, nb;
};


struct PassArgs {
  const Prime* prm;        // device: the two primes' constants for this N
  const double* src;       // [2][N] or null (then limbs)
  const uint32_t* limbs;   // operand limbs (first forward pass)
  uint64_t nlimbs;
  double* dst;             // [2][N]
  const double* spec;      // fused: spectrum of operand a, [2][N]
  uint64_t N;
  int logL, logS, logLines, logG, logB;
  // batched launches: product blockIdx.z uses descs[i] with i = z << zshift,
  // scratch offset i * zstride doubles, and operand a (0) or b (1) of its
  // descriptor; the fused pass reads a's spectrum at product
  // (i >> sshift) << sshift (sshift 1: products 2k and 2k+1 share operand a,
  // whose spectrum only product 2k computes)
  const BatchDesc* descs;
  int operand;
  int zshift, sshift;
  uint64_t zstride;
  uint64_t ntiles;  // tiles of E points per prime and product
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

// LDS layout of the lines staged by one workgroup.  Column passes (stride
// s > 1): point n of line lam at n * LINES + lam, so consecutive threads touch
// consecutive columns.  Row passes (s == 1): line-major, with one pad double
// after every 8 points (index n + n / 8): the first Stockham stages write at
// a stride of 8 points, which would otherwise hit the same LDS banks.
template <int LOGL, int LOGLINES, bool ROW>
struct Geo {
  static constexpr int L = 1 << LOGL;
  static constexpr int LINES = 1 << LOGLINES;
  static constexpr int E = L * LINES;
  static constexpr int NTH = E / kEPT;
  static constexpr int LP = ROW ? L + L / 8 : L;
  static constexpr int LDS = ROW ? LINES * LP : E;
  __device__ static __forceinline__ int idx(int n, int lam) {
    return ROW ? lam * LP + n + (n >> 3) : (n << LOGLINES) + lam;
  }
  // butterfly bf -> (line, j): the fast-varying index follows the layout
  __device__ static __forceinline__ void split(int bf, int logLR, int& lam, int& j) {
    if (ROW) {
      lam = bf >> logLR;
      j = bf & ((1 << logLR) - 1);
    } else {
      lam = bf & (LINES - 1);
      j = bf >> LOGLINES;
    }
  }
};

// Stage twiddles omega_L^e (e < L) come from an LDS copy of the table for
// lines up to 2^10 points; longer lines keep two 64-entry LDS tables,
// omega_4096^(64 h) and omega_4096^l, and form omega_4096^(64 h + l) with
// one modular product (no global loads inside the stages).
constexpr int kLogLdsTw = 10;
constexpr int kTwSplit = 64;
constexpr int kLdsTwEntries = 1 << kLogLdsTw;

template <int LOGL>
constexpr int lds_tw_size() {
  return LOGL <= kLogLdsTw ? (1 << LOGL) : 2 * kTwSplit;
}

template <int LOGL>
__device__ __forceinline__ double2 stage_tw(const double2* tw, int e, const Prime& P) {
  if constexpr (LOGL <= kLogLdsTw) {
    return tw[e];
  } else {
    const double2 lo = tw[e & (kTwSplit - 1)];
    const double2 hi = tw[kTwSplit + (e >> 6)];
    const double w = mm_shoup(hi.x, lo.x, lo.y, P.p);
    return make_double2(w, w * P.pinv);
  }
}

// One Stockham stage of radix R (Ns = 2^logNs) over every line in LDS.
// tw: the twiddle table, omega_{2^logTw}^e.
template <class GE, int R, int LOGR, int LOGNS>
__device__ __forceinline__ void stockham_stage(double* lds, const double2* tw, const Prime& P,
                                               bool inv) {
  constexpr int LOGL = __builtin_ctz(GE::L);
  constexpr int LOGTW = LOGL <= kLogLdsTw ? LOGL : kLogE;
  constexpr int TWN = 1 << LOGTW;
  constexpr int NB = kEPT / R;
  constexpr int LOGLR = GE::L == 1 ? 0 : (LOGL - LOGR);
  constexpr int NS = 1 << LOGNS;
  constexpr int SH = LOGTW - LOGNS - LOGR;
  const int tid = threadIdx.x;
  double v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    int lam, j;
    GE::split(tid + b * GE::NTH, LOGLR, lam, j);
#pragma unroll
    for (int r = 0; r < R; ++r) v[b][r] = lds[GE::idx(j + (r << LOGLR), lam)];
    if (LOGNS > 0) {
      const int jm = j & (NS - 1);
      v[b][0] = reduce(v[b][0], P.p, P.pinv);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        int e = (jm * r) << SH;
        if (inv) e = (TWN - e) & (TWN - 1);
        const double2 t = stage_tw<LOGL>(tw, e, P);
        v[b][r] = mm_shoup(v[b][r], t.x, t.y, P.p);
      }
    }
    dft_core<R>(v[b], P, inv);
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    int lam, j;
    GE::split(tid + b * GE::NTH, LOGLR, lam, j);
    const int jm = j & (NS - 1);
    const int idxD = ((j - jm) << LOGR) + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[GE::idx(idxD + (r << LOGNS), lam)] = v[b][r];
  }
  __syncthreads();
}

template <class GE, int LOGNS>
__device__ __forceinline__ void radix8_stages(double* lds, const double2* tw, const Prime& P,
                                              bool inv) {
  if constexpr (LOGNS < __builtin_ctz(GE::L)) {
    stockham_stage<GE, 8, 3, LOGNS>(lds, tw, P, inv);
    radix8_stages<GE, LOGNS + 3>(lds, tw, P, inv);
  }
}

// Natural-order DFT of length L of every line in LDS.
template <class GE>
__device__ __forceinline__ void line_dft(double* lds, const double2* tw, const Prime& P, bool inv) {
  constexpr int LOGL = __builtin_ctz(GE::L);
  constexpr int FIRST = LOGL % 3;
  if constexpr (FIRST == 1) {
    stockham_stage<GE, 2, 1, 0>(lds, tw, P, inv);
    radix8_stages<GE, 1>(lds, tw, P, inv);
  } else if constexpr (FIRST == 2) {
    stockham_stage<GE, 4, 2, 0>(lds, tw, P, inv);
    radix8_stages<GE, 2>(lds, tw, P, inv);
  } else {
    radix8_stages<GE, 0>(lds, tw, P, inv);
  }
}

// Pieces of a Stockham stage for the pass kernel, whose first stage takes
// its inputs from registers (loaded straight from HBM) and whose last stage
// leaves its outputs in registers (stored straight to HBM): two LDS round
// trips and four barriers fewer per tile than a line DFT done wholly in LDS.
// v[b][r] is input r of butterfly bf = tid + b * NTH; in the last stage
// (Ns = L / R) every output r of a butterfly lands where its input r came
// from: point j + r (L / R) of line lam.
template <class GE, int LOGR>
__device__ __forceinline__ void stage_pos(int b, int r, int& lam, int& n) {
  constexpr int LOGLR = __builtin_ctz(GE::L) - LOGR;
  int j;
  GE::split(threadIdx.x + b * GE::NTH, LOGLR, lam, j);
  n = j + (r << LOGLR);
}

template <class GE, int R, int LOGR, int LOGNS>
__device__ __forceinline__ void stage_compute(double (&v)[kEPT / R][R], const double2* tw,
                                              const Prime& P, bool inv) {
  constexpr int LOGL = __builtin_ctz(GE::L);
  constexpr int LOGTW = LOGL <= kLogLdsTw ? LOGL : kLogE;
  constexpr int TWN = 1 << LOGTW;
  constexpr int LOGLR = LOGL - LOGR;
  constexpr int NS = 1 << LOGNS;
  constexpr int SH = LOGTW - LOGNS - LOGR;
#pragma unroll
  for (int b = 0; b < kEPT / R; ++b) {
    if (LOGNS > 0) {
      int lam, j;
      GE::split(threadIdx.x + b * GE::NTH, LOGLR, lam, j);
      const int jm = j & (NS - 1);
      v[b][0] = reduce(v[b][0], P.p, P.pinv);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        int e = (jm * r) << SH;
        if (inv) e = (TWN - e) & (TWN - 1);
        const double2 t = stage_tw<LOGL>(tw, e, P);
        v[b][r] = mm_shoup(v[b][r], t.x, t.y, P.p);
      }
    }
    dft_core<R>(v[b], P, inv);
  }
}

template <class GE, int R, int LOGR, int LOGNS>
__device__ __forceinline__ void stage_write(const double (&v)[kEPT / R][R], double* lds) {
  constexpr int LOGLR = __builtin_ctz(GE::L) - LOGR;
  constexpr int NS = 1 << LOGNS;
#pragma unroll
  for (int b = 0; b < kEPT / R; ++b) {
    int lam, j;
    GE::split(threadIdx.x + b * GE::NTH, LOGLR, lam, j);
    const int jm = j & (NS - 1);
    const int idxD = ((j - jm) << LOGR) + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[GE::idx(idxD + (r << LOGNS), lam)] = v[b][r];
  }
}

// Radix-8 stages with Ns = 2^LOGNS .. L / 64 (all but the last), in LDS.
template <class GE, int LOGNS>
__device__ __forceinline__ void middle_stages(double* lds, const double2* tw, const Prime& P,
                                              bool inv) {
  if constexpr (LOGNS + 3 < __builtin_ctz(GE::L)) {
    stockham_stage<GE, 8, 3, LOGNS>(lds, tw, P, inv);
    middle_stages<GE, LOGNS + 3>(lds, tw, P, inv);
  }
}

// The last (radix-8) stage: LDS -> registers, y[r] = output at stage_pos<GE,3>(0, r).
template <class GE>
__device__ __forceinline__ void last_stage_to_regs(const double* lds, double (&y)[1][8],
                                                   const double2* tw, const Prime& P, bool inv) {
  constexpr int LOGL = __builtin_ctz(GE::L);
  static_assert(kEPT == 8, "one radix-8 butterfly per thread in the last stage");
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    int lam, n;
    stage_pos<GE, 3>(0, r, lam, n);
    y[0][r] = lds[GE::idx(n, lam)];
  }
  stage_compute<GE, 8, 3, LOGL - 3>(y, tw, P, inv);
}

// Line DFT whose inputs are the registers x (stage_pos<GE, LOGR1> order) and
// whose outputs are returned in y (stage_pos<GE, 3>(0, r) order).
template <class GE>
__device__ __forceinline__ void line_dft_regs(double (&x)[kEPT], double (&y)[1][8], double* lds,
                                              const double2* tw, const Prime& P, bool inv) {
  constexpr int LOGL = __builtin_ctz(GE::L);
  constexpr int LOGR1 = LOGL % 3 == 1 ? 1 : (LOGL % 3 == 2 ? 2 : 3);
  constexpr int R1 = 1 << LOGR1;
  double v[kEPT / R1][R1];
#pragma unroll
  for (int b = 0; b < kEPT / R1; ++b)
#pragma unroll
    for (int r = 0; r < R1; ++r) v[b][r] = x[b * R1 + r];
  stage_compute<GE, R1, LOGR1, 0>(v, tw, P, inv);
  stage_write<GE, R1, LOGR1, 0>(v, lds);
  __syncthreads();
  middle_stages<GE, LOGR1>(lds, tw, P, inv);
  last_stage_to_regs<GE>(lds, y, tw, P, inv);
}

// Stage-twiddle table of a line length: LDS copy (omega_L^e) when small
// enough, else the global 4096-entry table.  Call before the first barrier.
template <class GE>
__device__ __forceinline__ const double2* stage_twiddles(double2* ltw, const Prime& P) {
  constexpr int LOGL = __builtin_ctz(GE::L);
  if constexpr (LOGL <= kLogLdsTw) {
    for (int e = threadIdx.x; e < GE::L; e += blockDim.x) ltw[e] = P.twL[e << (kLogE - LOGL)];
  } else {
    // exponents of the 4096-entry table (LOGTW = kLogE): e = 64 h + l
    for (int e = threadIdx.x; e < 2 * kTwSplit; e += blockDim.x)
      ltw[e] = e < kTwSplit ? P.twL[e] : P.twL[(e - kTwSplit) * kTwSplit];
  }
  return ltw;
}

// One pass.  Column passes (ROW false) need s >= LINES (one block of L*s
// points spans several workgroups, each taking LINES adjacent columns); row
// passes have s == 1 and take LINES consecutive lines.
//
// A workgroup walks tiles w = blockIdx.x, + gridDim.x, ... (a.ntiles per
// prime and product).  Each thread loads the inputs of its first-stage
// butterflies straight from HBM and stores the outputs of its last-stage
// butterfly straight to HBM; the next tile's inputs are loaded right after
// the current tile's first stage, so they are in flight during its
// remaining stages.
template <int MODE, int LOGL, int LOGLINES, bool ROW>
__global__ __launch_bounds__((1 << (LOGL + LOGLINES)) / kEPT) void k_ntt_pass(PassArgs a) {
  using GE = Geo<LOGL, LOGLINES, ROW>;
  constexpr int LOGR1 = LOGL % 3 == 1 ? 1 : (LOGL % 3 == 2 ? 2 : 3);
  constexpr int R1 = 1 << LOGR1;
  constexpr int NB1 = kEPT / R1;
  __shared__ double lds[GE::LDS];
  __shared__ double2 ltw[lds_tw_size<LOGL>()];
  const int pi = blockIdx.y;
  const Prime P = a.prm[pi];
  const double2* tw = stage_twiddles<GE>(ltw, P);
  const int logS = a.logS;
  uint64_t zoff = 0, soff = 0;
  const uint32_t* limbs = a.limbs;
  uint64_t nlimbs = a.nlimbs;
  if (a.descs) {
    const uint64_t zi = (uint64_t)blockIdx.z << a.zshift;
    const BatchDesc d = a.descs[zi];
    zoff = zi * a.zstride;
    soff = ((zi >> a.sshift) << a.sshift) * a.zstride;
    limbs = a.operand ? d.b : d.a;
    nlimbs = a.operand ? d.nb : d.na;
  }
  const double* src = a.src ? a.src + zoff + pi * a.N : nullptr;
  double* dst = a.dst + zoff + pi * a.N;
  const int logRoot = LOGL + logS;
  const uint64_t ntiles = a.ntiles;

  // first point of tile t's lines (the global step between them is implied
  // by the layout), and the global index of point n of line lam
  auto tile_base = [&](uint64_t t) -> uint64_t {
    if (ROW) return t << (LOGL + LOGLINES);
    const int lgc = logS - LOGLINES;  // column groups per block, log2
    return ((t >> lgc) << (LOGL + logS)) + ((t & ((1ull << lgc) - 1)) << LOGLINES);
  };
  auto gidx = [&](uint64_t base, int n, int lam) -> uint64_t {
    return ROW ? base + ((uint64_t)lam << LOGL) + (uint64_t)n : base + ((uint64_t)n << logS) + lam;
  };
  // inputs of this thread's first-stage butterflies: x[b R1 + r]
  auto load_tile = [&](uint64_t t, double (&x)[kEPT]) {
    const uint64_t b0 = tile_base(t);
#pragma unroll
    for (int b = 0; b < NB1; ++b)
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        int lam, n;
        stage_pos<GE, LOGR1>(b, r, lam, n);
        const uint64_t g = gidx(b0, n, lam);
        x[b * R1 + r] = src ? src[g] : (g < nlimbs ? (double)limbs[g] : 0.0);
      }
  };
  // Inter-pass twiddle omega^{c n} of a column pass, c = this thread's
  // column (fixed: NTH is a multiple of LINES), along the thread's points
  // n = n0 + k * step: base omega^{c n0} times powers of omega^{c step}.
  auto col_twiddle = [&](uint64_t base, int lam, uint64_t n0, uint64_t step, double2& cur,
                         double2& stp) {
    const uint64_t c = (base & ((1ull << logS) - 1)) + (uint64_t)lam;
    cur = root_pow(c * n0, logRoot, MODE == kInv, P);
    stp = root_pow(c * step, logRoot, MODE == kInv, P);
  };

  // One tile: xr holds its first-stage inputs; once they are in LDS, xr
  // receives the loads of tile wload, which stay in flight through this
  // tile's remaining stages.
  auto tile_step = [&](const uint64_t w, double (&xr)[kEPT], const uint64_t wload) {
    const uint64_t base = tile_base(w);
    if (MODE == kInv && !ROW) {
      // pre-twiddle: butterfly b's points are j_b + r L / R1
#pragma unroll
      for (int b = 0; b < NB1; ++b) {
        int lam, n0;
        stage_pos<GE, LOGR1>(b, 0, lam, n0);
        double2 cur, stp;
        col_twiddle(base, lam, (uint64_t)n0, (uint64_t)(GE::L >> LOGR1), cur, stp);
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          double& x = xr[b * R1 + r];
          x = mm_shoup(x, cur.x, cur.y, P.p);
          const double wv = mm_shoup(cur.x, stp.x, stp.y, P.p);
          cur = make_double2(wv, wv * P.pinv);
        }
      }
    }
    // first stage from registers
    double v1[NB1][R1];
#pragma unroll
    for (int b = 0; b < NB1; ++b)
#pragma unroll
      for (int r = 0; r < R1; ++r) v1[b][r] = xr[b * R1 + r];
    stage_compute<GE, R1, LOGR1, 0>(v1, tw, P, MODE == kInv);
    stage_write<GE, R1, LOGR1, 0>(v1, lds);
    __syncthreads();
    if (wload < ntiles) load_tile(wload, xr);
    middle_stages<GE, LOGR1>(lds, tw, P, MODE == kInv);
    double y[1][8];
    last_stage_to_regs<GE>(lds, y, tw, P, MODE == kInv);
    int lam_o, n_o0;
    stage_pos<GE, 3>(0, 0, lam_o, n_o0);
    if (MODE == kFused) {
      // pointwise with a's spectrum and N^-1, then the inverse DFT in LDS
      const double* spec = a.spec + soff + pi * a.N;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const uint64_t g = gidx(base, n_o0 + r * (GE::L >> 3), lam_o);
        const double x = mm_gen(y[0][r], spec[g], P.p, P.pinv);
        y[0][r] = mm_shoup(x, P.ninv.x, P.ninv.y, P.p);
      }
      __syncthreads();  // every last-stage read of LDS is done
#pragma unroll
      for (int r = 0; r < 8; ++r) lds[GE::idx(n_o0 + r * (GE::L >> 3), lam_o)] = y[0][r];
      __syncthreads();
      if constexpr (LOGR1 == 1) {
        stockham_stage<GE, 2, 1, 0>(lds, tw, P, true);
      } else if constexpr (LOGR1 == 2) {
        stockham_stage<GE, 4, 2, 0>(lds, tw, P, true);
      } else {
        stockham_stage<GE, 8, 3, 0>(lds, tw, P, true);
      }
      middle_stages<GE, LOGR1>(lds, tw, P, true);
      last_stage_to_regs<GE>(lds, y, tw, P, true);
    }
    // last-stage outputs straight to HBM: point n_o0 + r L / 8 of line lam_o
    double2 cur = make_double2(1.0, 0.0), stp = make_double2(1.0, 0.0);
    if (MODE == kFwd && !ROW)
      col_twiddle(base, lam_o, (uint64_t)n_o0, (uint64_t)(GE::L >> 3), cur, stp);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      double x = y[0][r];
      if (MODE == kFwd && !ROW) {
        x = mm_shoup(x, cur.x, cur.y, P.p);
        const double wv = mm_shoup(cur.x, stp.x, stp.y, P.p);
        cur = make_double2(wv, wv * P.pinv);
      } else {
        x = reduce(x, P.p, P.pinv);
      }
      dst[gidx(base, n_o0 + r * (GE::L >> 3), lam_o)] = x;
    }
  };

  uint64_t w = blockIdx.x;
  if (w >= ntiles) return;
  const uint64_t g = gridDim.x;
  double xa[kEPT];
  load_tile(w, xa);
  if constexpr (!ROW) {
    // column passes (at most 106 VGPRs) keep two tiles' loads in flight
    double xb[kEPT];
    if (w + g < ntiles) load_tile(w + g, xb);
    for (;;) {
      tile_step(w, xa, w + 2 * g);
      w += g;
      if (w >= ntiles) break;
      __syncthreads();  // the next tile's first stage overwrites LDS
      tile_step(w, xb, w + 2 * g);
      w += g;
      if (w >= ntiles) break;
      __syncthreads();
    }
  } else {
    for (;;) {
      tile_step(w, xa, w + g);
      w += g;
      if (w >= ntiles) break;
      __syncthreads();  // the next tile's first stage overwrites LDS
    }
  }
}

}  // namespace pcg
