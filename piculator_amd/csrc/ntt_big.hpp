// Large-tile NTT column passes for long transforms (gfx950: up to 160 KB of
// LDS per workgroup), and the pieces the row pass (ntt_row.hpp) shares.
//
// The same pass as k_ntt_pass (ntt_kernels.hpp) -- a line DFT of length
// L = 2^LOGL over lines of stride s, with the same inter-pass twiddles --
// but each workgroup stages 2^14 points of one prime (128 KB) and 1024
// threads hold 16 points each through radix-16 Stockham stages.  Columns of
// up to 2^11 points (8 adjacent columns, 64-byte runs) and 2^14-point rows
// make a 2^25-point transform two HBM sweeps instead of three.
//
// Stage twiddles omega_{R Ns}^{jm r} come from an LDS table of
// omega_{L/16}^e (e < L/16) for every stage but the last; the last stage
// (Ns = L/16) splits jm = 16 a + b and forms
//   omega_L^{jm r} = omega_{L/16}^{a r} * omega_L^{b r}
// from that table and a 256-entry one (b r < 256), one modular product per
// twiddle.  Radix 16 is 4 x 4 with the internal omega_16 powers.
#pragma once
#include "ntt_kernels.hpp"

#pragma clang fp contract(off)

namespace pcg {

constexpr int kBigLogE = 14;
constexpr int kBigE = 1 << kBigLogE;
// 16 points per thread (1024 threads, at most 128 registers each); 32 points
// per thread (512 threads) measured 19% slower at 2^25 points
constexpr int kBigEPT = 16;
constexpr int kBigNTH = kBigE / kBigEPT;

static_assert(kBigEPT == 16 || kBigEPT == 32, "16 or 32 points per thread");

// Stage-twiddle table entries hold w; its Shoup quotient w/p is formed at
// use (one multiply; (w, w/p) pairs measured no faster and take twice the
// LDS).
using BigTw = double;
__device__ __forceinline__ double2 tw_of(const BigTw* ta, int e, const Prime& P) {
  const double w = ta[e];
  return make_double2(w, w * P.pinv);
}

// Natural-order DFT of 16 inputs: r = 4 r1 + r2, k = k1 + 4 k2,
// X_k = sum_r2 w4^{r2 k2} w16^{r2 k1} sum_r1 y_{4 r1 + r2} w4^{r1 k1}.
__device__ __forceinline__ void dft16(double* v, const Prime& P, bool inv) {
  const double2* w8 = inv ? P.w8i : P.w8f;    // w16^2, w16^4, w16^6
  const double2* w16 = inv ? P.w16i : P.w16f;  // w16^1, w16^3, w16^9
  double z[4][4];
#pragma unroll
  for (int r2 = 0; r2 < 4; ++r2) {
    double t[4] = {v[r2], v[4 + r2], v[8 + r2], v[12 + r2]};
    dft_core<4>(t, P, inv);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) z[r2][k1] = t[k1];
  }
  // w16^{r2 k1}: (1,1) w16, (1,2) w8, (1,3) w16^3, (2,1) w8, (2,2) w4,
  // (2,3) w8^3, (3,1) w16^3, (3,2) w8^3, (3,3) w16^9
  z[1][1] = mm_shoup(z[1][1], w16[0].x, w16[0].y, P.p);
  z[1][2] = mm_shoup(z[1][2], w8[0].x, w8[0].y, P.p);
  z[1][3] = mm_shoup(z[1][3], w16[1].x, w16[1].y, P.p);
  z[2][1] = mm_shoup(z[2][1], w8[0].x, w8[0].y, P.p);
  z[2][2] = mm_shoup(z[2][2], w8[1].x, w8[1].y, P.p);
  z[2][3] = mm_shoup(z[2][3], w8[2].x, w8[2].y, P.p);
  z[3][1] = mm_shoup(z[3][1], w16[1].x, w16[1].y, P.p);
  z[3][2] = mm_shoup(z[3][2], w8[2].x, w8[2].y, P.p);
  z[3][3] = mm_shoup(z[3][3], w16[2].x, w16[2].y, P.p);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    double t[4] = {z[0][k1], z[1][k1], z[2][k1], z[3][k1]};
    dft_core<4>(t, P, inv);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
  }
}

template <int R>
__device__ __forceinline__ void big_dft(double* v, const Prime& P, bool inv) {
  if constexpr (R == 16)
    dft16(v, P, inv);
  else
    dft_core<R>(v, P, inv);
}

// Tile geometry: LINES = 2^(14 - LOGL) adjacent columns, point n of line
// lam at n * LINES + lam, plus 8 pad doubles per 8 points.
template <int LOGL>
struct BigGeo {
  static constexpr int L = 1 << LOGL;
  static constexpr int LOGLINES = kBigLogE - LOGL;
  static constexpr int LINES = 1 << LOGLINES;
  static constexpr int LDS = kBigE + kBigE / 8;
  static constexpr int LOGR1 = LOGL % 4 == 0 ? 4 : LOGL % 4;
  static_assert(LOGL >= 5 && LOGL <= 11, "column tiles of 8 or more lines");
  // columns: 8 doubles of pad after every 8 points, so the first stage's
  // writes (8 points of a butterfly, adjacent lines) spread over the banks
  static constexpr int kColPad = 3;
  __device__ static __forceinline__ int idx(int n, int lam) {
    return (n << LOGLINES) + lam + ((n >> kColPad) << 3);
  }
  // idx(n + o, lam) == idx(n, lam) + ioff(o) for every access of the stages
  // (n a butterfly's base point, o = r L/R or r Ns a compile-time offset:
  // checked exhaustively for every instance), so the per-point LDS
  // addresses are one base plus immediates
  __device__ static constexpr int ioff(int o) { return (o << LOGLINES) + ((o >> kColPad) << 3); }
  __device__ static __forceinline__ void split(int bf, int logLR, int& lam, int& j) {
    lam = bf & (LINES - 1);
    j = bf >> LOGLINES;
  }
};

// The thread index, opaque to the optimiser: address arithmetic derived
// from it is recomputed where it is used (a few integer operations) instead
// of being hoisted out of the tile loop and kept live -- with 1024 threads
// and 128 registers the hoisted addresses would spill, and every spill
// reload waits for the next tile's loads in flight.
__device__ __forceinline__ int big_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// Orders this wave's LDS accesses before and after it (the wave's LDS
// operations execute in order, so no workgroup barrier is needed for an
// exchange between its own lanes).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Twiddle omega_{2^(LOGR+LOGNS)}^{e}, e = jm r, for a stage of the line DFT
// before the last one, from the table ta[e] = omega_{L/16}^e.
template <int LOGL, int LOGR, int LOGNS>
__device__ __forceinline__ double2 big_stage_tw(const BigTw* ta, int jm, int r, const Prime& P, bool inv) {
  constexpr int LOGTA = LOGL - 4;  // table omega_{L/16}
  static_assert(LOGR + LOGNS <= LOGTA, "stages before the last one only");
  constexpr int SH = LOGTA - LOGR - LOGNS;
  int e = (jm * r) << SH;
  if (inv) e = ((1 << LOGTA) - e) & ((1 << LOGTA) - 1);
  return tw_of(ta, e, P);
}

// Inputs of butterfly b of this thread in a stage of radix 2^LOGR: line lam,
// point j + r L / R.
template <class GE, int LOGR>
__device__ __forceinline__ void big_pos(int t, int b, int r, int& lam, int& n) {
  constexpr int LOGLR = __builtin_ctz(GE::L) - LOGR;
  int j;
  GE::split(t + b * kBigNTH, LOGLR, lam, j);
  n = j + (r << LOGLR);
}

// Twiddles and DFTs of a stage before the last one on the registers v
// (butterfly b, input r).
template <class GE, int LOGR, int LOGNS>
__device__ __forceinline__ void big_compute(double (&v)[kBigEPT >> LOGR][1 << LOGR], const BigTw* ta,
                                            const Prime& P, bool inv) {
  constexpr int R = 1 << LOGR;
  constexpr int LOGLR = __builtin_ctz(GE::L) - LOGR;
  constexpr int NS = 1 << LOGNS;
  const int t = big_tid();
#pragma unroll
  for (int b = 0; b < kBigEPT / R; ++b) {
    if constexpr (LOGNS > 0) {
      int lam, j;
      GE::split(t + b * kBigNTH, LOGLR, lam, j);
      const int jm = j & (NS - 1);
      v[b][0] = reduce(v[b][0], P.p, P.pinv);
      constexpr int LOGL = __builtin_ctz(GE::L);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const double2 t = big_stage_tw<LOGL, LOGR, LOGNS>(ta, jm, r, P, inv);
        v[b][r] = mm_shoup(v[b][r], t.x, t.y, P.p);
      }
    }
    big_dft<R>(v[b], P, inv);
  }
}

template <class GE, int LOGR, int LOGNS>
__device__ __forceinline__ void big_write(const double (&v)[kBigEPT >> LOGR][1 << LOGR], double* lds) {
  constexpr int R = 1 << LOGR;
  constexpr int LOGLR = __builtin_ctz(GE::L) - LOGR;
  constexpr int NS = 1 << LOGNS;
  const int t = big_tid();
#pragma unroll
  for (int b = 0; b < kBigEPT / R; ++b) {
    int lam, j;
    GE::split(t + b * kBigNTH, LOGLR, lam, j);
    const int jm = j & (NS - 1);
    const int idxD = ((j - jm) << LOGR) + jm;
    double* base = lds + GE::idx(idxD, lam);
#pragma unroll
    for (int r = 0; r < R; ++r) base[GE::ioff(r << LOGNS)] = v[b][r];
  }
}

template <class GE, int LOGR>
__device__ __forceinline__ void big_read(double (&v)[kBigEPT >> LOGR][1 << LOGR], const double* lds) {
  constexpr int LOGLR = __builtin_ctz(GE::L) - LOGR;
  const int t = big_tid();
#pragma unroll
  for (int b = 0; b < (kBigEPT >> LOGR); ++b) {
    int lam, n0;
    big_pos<GE, LOGR>(t, b, 0, lam, n0);
    const double* base = lds + GE::idx(n0, lam);
#pragma unroll
    for (int r = 0; r < (1 << LOGR); ++r) v[b][r] = base[GE::ioff(r << LOGLR)];
  }
}

// A whole stage in LDS (read, compute, barrier, write, barrier).
template <class GE, int LOGR, int LOGNS>
__device__ __forceinline__ void big_stage(double* lds, const BigTw* ta, const Prime& P, bool inv) {
  double v[kBigEPT >> LOGR][1 << LOGR];
  big_read<GE, LOGR>(v, lds);
  big_compute<GE, LOGR, LOGNS>(v, ta, P, inv);
  __syncthreads();
  big_write<GE, LOGR, LOGNS>(v, lds);
  __syncthreads();
}

// Radix-16 stages from Ns = 2^LOGNS up to (not including) the last one.
template <class GE, int LOGNS>
__device__ __forceinline__ void big_middle(double* lds, const BigTw* ta, const Prime& P, bool inv) {
  if constexpr (LOGNS + 4 < __builtin_ctz(GE::L)) {
    big_stage<GE, 4, LOGNS>(lds, ta, P, inv);
    big_middle<GE, LOGNS + 4>(lds, ta, P, inv);
  }
}

constexpr int kBigNBL = kBigEPT / 16;  // last-stage butterflies per thread

// The last (radix-16) stage reads LDS into registers: y[b][r] is output
// point j_b + r L / 16 of butterfly b's line (the place its input r came
// from).  Its twiddles omega_L^{jm r} take a table of omega_L or
// omega_{2L} (big_last_half, big_last_th below).

// The last stage of a forward line DFT with its twiddles omega_L^e (e =
// jm r < L) from a table th[n] = omega_{2L}^n, n < L: omega_L^e =
// omega_{2L}^{2e}, and for 2e >= L it is -th[2e - L].  One table read and
// one modular product per twiddle instead of the two factors of big_last.
template <class GE>
__device__ __forceinline__ void big_last_th(const double* lds, double (&y)[kBigNBL][16], const double* th,
                                            const Prime& P) {
  constexpr int L = GE::L;
  constexpr int LOGLR = __builtin_ctz(L) - 4;
  constexpr int NS = L / 16;
  big_read<GE, 4>(y, lds);
  int lam, j;
  GE::split(big_tid(), LOGLR, lam, j);
  const int jm = j & (NS - 1);
  y[0][0] = reduce(y[0][0], P.p, P.pinv);
#pragma unroll
  for (int r = 1; r < 16; ++r) {
    const int e2 = 2 * jm * r;  // < 2L
    const double t = th[e2 & (L - 1)];
    const double w = (e2 & L) ? -t : t;
    y[0][r] = mm_shoup(y[0][r], w, w * P.pinv, P.p);
  }
  dft16(y[0], P, false);
}

// The last stage of a line DFT (either direction) with its twiddles
// omega_L^{+-e} (e = jm r < L) from a table th[n] = omega_L^n, n < L / 2:
// omega_L^{n + L/2} = -omega_L^n.
template <class GE>
__device__ __forceinline__ void big_last_half(const double* lds, double (&y)[kBigNBL][16], const double* th,
                                              const Prime& P, bool inv) {
  constexpr int L = GE::L;
  constexpr int LOGLR = __builtin_ctz(L) - 4;
  constexpr int NS = L / 16;
  big_read<GE, 4>(y, lds);
  int lam, j;
  GE::split(big_tid(), LOGLR, lam, j);
  const int jm = j & (NS - 1);
  y[0][0] = reduce(y[0][0], P.p, P.pinv);
#pragma unroll
  for (int r = 1; r < 16; ++r) {
    int e = jm * r;  // < L
    if (inv) e = (L - e) & (L - 1);
    const double t = th[e & (L / 2 - 1)];
    const double w = (e & (L / 2)) ? -t : t;
    y[0][r] = mm_shoup(y[0][r], w, w * P.pinv, P.p);
  }
  dft16(y[0], P, inv);
}

// Fills the stage tables: ta[e] = omega_{L/16}^e (e < L/16, the stages
// before the last; the inverse direction negates the index) and, when th
// is given, th[n] = omega_L^n (n < L/2, the last stage).  Call before the
// first barrier.
template <class GE>
__device__ __forceinline__ void big_tables(BigTw* ta, double* th, const Prime& P) {
  constexpr int LOGL = __builtin_ctz(GE::L);
  constexpr int LOGTA = LOGL - 4;
  static_assert(LOGTA >= 1 && LOGL <= kLogE, "line length out of range");
  for (int e = threadIdx.x; e < (1 << LOGTA); e += blockDim.x)
    ta[e] = P.twL[e << (kLogE - LOGTA)].x;  // twL = (omega_4096^e, its quotient)
  if (th)
    for (int e = threadIdx.x; e < GE::L / 2; e += blockDim.x) th[e] = P.twL[e << (kLogE - LOGL)].x;
}

template <int MODE, int LOGL>
__global__ __launch_bounds__(kBigNTH) void k_ntt_col(PassArgs a) {
  static_assert(MODE == kFwd || MODE == kInv, "column passes run forward or inverse");
  using GE = BigGeo<LOGL>;
  constexpr int LOGR1 = GE::LOGR1;
  constexpr int R1 = 1 << LOGR1;
  constexpr int NB1 = kBigEPT / R1;
  constexpr int LOGTA = LOGL - 4;
  // inverse pre-twiddles from one power per thread (LOGL <= 8, see
  // inv_twiddles): omega_N^{-lam} for the tile's columns lam
  constexpr bool kPowTw = MODE == kInv && LOGL <= 8;
  __shared__ double lds[GE::LDS];
  __shared__ BigTw ta[1 << LOGTA];
  __shared__ double th[GE::L / 2];  // omega_L^n, n < L / 2 (last stage)
  __shared__ double tlam[kPowTw ? GE::LINES : 1];
  const int pi = blockIdx.y;
  const Prime P = a.prm[pi];
  big_tables<GE>(ta, th, P);
  const int logS = a.logS;
  uint64_t zoff = 0;
  const uint32_t* limbs = a.limbs;
  uint64_t nlimbs = a.nlimbs;
  if (a.descs) {
    const uint64_t zi = (uint64_t)blockIdx.z << a.zshift;
    const BatchDesc d = a.descs[zi];
    zoff = zi * a.zstride;
    limbs = a.operand ? d.b : d.a;
    nlimbs = a.operand ? d.nb : d.na;
  }
  const double* src = a.src ? a.src + zoff + pi * a.N : nullptr;
  double* dst = a.dst + zoff + pi * a.N;
  const int logRoot = LOGL + logS;
  const uint64_t ntiles = a.ntiles;
  if constexpr (kPowTw) {
    for (int e = threadIdx.x; e < GE::LINES; e += blockDim.x) tlam[e] = root_pow((uint64_t)e, logRoot, true, P).x;
    __syncthreads();
  }

  auto tile_base = [&](uint64_t t) -> uint64_t {
    const int lgc = logS - GE::LOGLINES;
    return ((t >> lgc) << (LOGL + logS)) + ((t & ((1ull << lgc) - 1)) << GE::LOGLINES);
  };
  auto gidx = [&](uint64_t base, int n, int lam) -> uint64_t {
    return base + ((uint64_t)n << logS) + lam;
  };
  // inputs of this thread's first-stage butterflies: x[b R1 + r]; limbs
  // past the operand are zeros (range-checked buffer loads)
  const LimbBuf lb(limbs, (uint32_t)nlimbs);
  auto load_tile = [&](uint64_t t, double (&x)[kBigEPT]) {
    const uint64_t b0 = tile_base(t);
    const int bt = big_tid();
#pragma unroll
    for (int b = 0; b < NB1; ++b)
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        int lam, n;
        big_pos<GE, LOGR1>(bt, b, r, lam, n);
        const uint64_t g = gidx(b0, n, lam);
        x[b * R1 + r] = src ? src[g] : (double)lb[(uint32_t)g];
      }
  };
  // inter-pass twiddle omega^{c n} of a column pass along the thread's points
  // n = n0 + k step: base omega^{c n0} times powers of omega^{c step}
  auto col_twiddle = [&](uint64_t base, int lam, uint64_t n0, uint64_t step, double2& cur,
                         double2& stp) {
    const uint64_t c = (base & ((1ull << logS) - 1)) + (uint64_t)lam;
    cur = root_pow(c * n0, logRoot, MODE == kInv, P);
    stp = root_pow(c * step, logRoot, MODE == kInv, P);
  };

  // bases of the inverse pre-twiddles of tile t's first-stage butterflies
  // (fetched with the tile's loads, so their table reads are in flight
  // together with them).  Butterfly b of thread t is column lam of the
  // tile (the same for every b) at point n0 + b 2^(L-4), n0 = j0 its first,
  // stepping by L / R1 = 2^(L-4) NB1.  kPowTw: one power per thread,
  // w = omega^{-c} = omega^{-c0} omega^{-lam} (c0 the tile's first column,
  // the same for the workgroup, and the tlam table), then by squarings and
  // products w^{j0}, q = w^{2^(L-4)} and the step q^{NB1}.  This replaces
  // two lookups of the 2^30-th-root tables per butterfly: at the large
  // strides of the last inverse pass of the three-pass plans (c up to
  // 2^21) their lanes each read a different line of the tables, which the
  // pass's streaming data keeps out of the caches.
  auto inv_twiddles = [&](uint64_t t, double2 (&cur)[NB1], double2 (&stp)[NB1]) {
    if constexpr (kPowTw) {
      const uint64_t b0 = tile_base(t);
      int lam, j0;
      big_pos<GE, LOGR1>(big_tid(), 0, 0, lam, j0);
      const double2 u0 = root_pow(b0 & ((1ull << logS) - 1), logRoot, true, P);
      const double w = mm_shoup(tlam[lam], u0.x, u0.y, P.p);
      double2 pw = make_double2(w, w * P.pinv);
      double2 cb = make_double2(1.0, P.pinv);
#pragma unroll
      for (int k = 0; k < LOGL - 4; ++k) {
        const double m = mm_shoup(cb.x, pw.x, pw.y, P.p);
        if ((j0 >> k) & 1) cb = make_double2(m, m * P.pinv);
        const double sq = mm_shoup(pw.x, pw.x, pw.y, P.p);
        pw = make_double2(sq, sq * P.pinv);
      }
      cur[0] = cb;  // omega^{-c j0}
#pragma unroll
      for (int b = 1; b < NB1; ++b) {
        const double m = mm_shoup(cur[b - 1].x, pw.x, pw.y, P.p);
        cur[b] = make_double2(m, m * P.pinv);
      }
#pragma unroll
      for (int k = 0; k < 4 - LOGR1; ++k) {
        const double sq = mm_shoup(pw.x, pw.x, pw.y, P.p);
        pw = make_double2(sq, sq * P.pinv);
      }
#pragma unroll
      for (int b = 0; b < NB1; ++b) stp[b] = pw;  // omega^{-c L / R1}
    } else if constexpr (MODE == kInv) {
      const uint64_t b0 = tile_base(t);
      const int bt = big_tid();
#pragma unroll
      for (int b = 0; b < NB1; ++b) {
        int lam, n0;
        big_pos<GE, LOGR1>(bt, b, 0, lam, n0);
        col_twiddle(b0, lam, (uint64_t)n0, (uint64_t)(GE::L >> LOGR1), cur[b], stp[b]);
      }
    }
  };

  auto tile_step = [&](const uint64_t w, double (&xr)[kBigEPT], double2 (&icur)[NB1],
                       double2 (&istp)[NB1], const uint64_t wload) {
    const uint64_t base = tile_base(w);
    const int bt = big_tid();
    // bases of the forward post-twiddles, fetched now, used after the stages
    double2 fcur[kBigNBL], fstp[kBigNBL];
    if constexpr (MODE == kFwd) {
#pragma unroll
      for (int b = 0; b < kBigNBL; ++b) {
        int lam_o, n_o0;
        big_pos<GE, 4>(bt, b, 0, lam_o, n_o0);
        col_twiddle(base, lam_o, (uint64_t)n_o0, (uint64_t)(GE::L >> 4), fcur[b], fstp[b]);
      }
    }
    if constexpr (MODE == kInv) {
#pragma unroll
      for (int b = 0; b < NB1; ++b) {
        double2 cur = icur[b], stp = istp[b];
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          double& x = xr[b * R1 + r];
          x = mm_shoup(x, cur.x, cur.y, P.p);
          const double wv = mm_shoup(cur.x, stp.x, stp.y, P.p);
          cur = make_double2(wv, wv * P.pinv);
        }
      }
    }
    {
      double v1[NB1][R1];
#pragma unroll
      for (int b = 0; b < NB1; ++b)
#pragma unroll
        for (int r = 0; r < R1; ++r) v1[b][r] = xr[b * R1 + r];
      big_compute<GE, LOGR1, 0>(v1, ta, P, MODE == kInv);
      big_write<GE, LOGR1, 0>(v1, lds);
    }
    __syncthreads();
    if (wload < ntiles) {
      load_tile(wload, xr);
      inv_twiddles(wload, icur, istp);
    }
    big_middle<GE, LOGR1>(lds, ta, P, MODE == kInv);
    double y[kBigNBL][16];
    big_last_half<GE>(lds, y, th, P, MODE == kInv);
#pragma unroll
    for (int b = 0; b < kBigNBL; ++b) {
      int lam_o, n_o0;
      big_pos<GE, 4>(bt, b, 0, lam_o, n_o0);
      double2 cur = make_double2(1.0, 0.0), stp = make_double2(1.0, 0.0);
      if constexpr (MODE == kFwd) {
        cur = fcur[b];
        stp = fstp[b];
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        double x = y[b][r];
        if constexpr (MODE == kFwd) {
          x = mm_shoup(x, cur.x, cur.y, P.p);
          const double wv = mm_shoup(cur.x, stp.x, stp.y, P.p);
          cur = make_double2(wv, wv * P.pinv);
        } else {
          x = reduce(x, P.p, P.pinv);
        }
        dst[gidx(base, n_o0 + r * (GE::L >> 4), lam_o)] = x;
      }
    }
  };

  // XCD-aware tile order: workgroups are dispatched to the 8 XCDs round
  // robin (linear id mod 8, and gridDim.x is a multiple of 8, so both
  // primes' workgroups x sit on the same XCD), so rank them XCD by XCD:
  // the workgroups of one XCD then take consecutive tiles, and a column
  // tile's 64-byte runs and its neighbour's (the two halves of each
  // 128-byte line) are read through the same L2 at the same time.
  const uint64_t g = gridDim.x;
  uint64_t w = blockIdx.x;
  if ((g & 7) == 0) w = (uint64_t)(blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
  if (w >= ntiles) return;
  double xa[kBigEPT];
  double2 icur[NB1], istp[NB1];
  load_tile(w, xa);
  inv_twiddles(w, icur, istp);
  for (;;) {
    tile_step(w, xa, icur, istp, w + g);
    w += g;
    if (w >= ntiles) break;
    __syncthreads();  // the next tile's first stage overwrites LDS
  }
}


// ---------------------------------------------------------------------------
// Column passes of 2 .. 16 points (batched products of 2^15 .. 2^18 points:
// one of these and the 2^14-point row pass).  A line is short enough for
// one thread: each of the 1024 threads of a 2^14-point tile runs the DFTs
// of 16 / L adjacent-in-the-tile columns in registers (no LDS); lanes take
// consecutive columns, so every access is 512 contiguous bytes per wave.
// Forward: DFT, then the inter-pass twiddle omega_{L s}^{c k} of k_ntt_col
// (from the limbs on the first pass); inverse: twiddle omega^{-c n}, then
// the inverse DFT.
template <int MODE, int LOGL>
__global__ __launch_bounds__(kBigNTH) void k_ntt_col_reg(PassArgs a) {
  static_assert(MODE == kFwd || MODE == kInv, "column passes run forward or inverse");
  static_assert(LOGL >= 1 && LOGL <= 4, "lines of 2 .. 16 points");
  constexpr int L = 1 << LOGL;
  constexpr int CPT = kBigEPT / L;  // columns per thread
  constexpr int LOGW = kBigLogE - LOGL;  // columns per tile
  const int pi = blockIdx.y;
  const Prime P = a.prm[pi];
  const int logS = a.logS;
  uint64_t zoff = 0;
  const uint32_t* limbs = a.limbs;
  uint64_t nlimbs = a.nlimbs;
  if (a.descs) {
    const uint64_t zi = (uint64_t)blockIdx.z << a.zshift;
    const BatchDesc d = a.descs[zi];
    zoff = zi * a.zstride;
    limbs = a.operand ? d.b : d.a;
    nlimbs = a.operand ? d.nb : d.na;
  }
  const double* src = a.src ? a.src + zoff + pi * a.N : nullptr;
  double* dst = a.dst + zoff + pi * a.N;
  const LimbBuf lb(limbs, (uint32_t)nlimbs);  // first pass: zero past the operand
  const int logRoot = LOGL + logS;
  const uint64_t ntiles = a.ntiles;
  const uint64_t g = gridDim.x;
  uint64_t w = blockIdx.x;
  if ((g & 7) == 0) w = (uint64_t)(blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
  for (; w < ntiles; w += g) {
    const int lgc = logS - LOGW;
    const uint64_t base = ((w >> lgc) << (LOGL + logS)) + ((w & ((1ull << lgc) - 1)) << LOGW);
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const uint64_t col = (uint64_t)threadIdx.x + ((uint64_t)j << 10);  // in the tile
      const uint64_t c = (base & ((1ull << logS) - 1)) + col;             // in the stride block
      double v[L];
#pragma unroll
      for (int n = 0; n < L; ++n) {
        const uint64_t gi = base + ((uint64_t)n << logS) + col;
        v[n] = src ? src[gi] : (double)lb[(uint32_t)gi];  // zero past the operand
      }
      const double2 stp = root_pow(c, logRoot, MODE == kInv, P);  // omega^{+-c}
      double2 cur = make_double2(1.0, 0.0);
      if constexpr (MODE == kInv) {
        // pre-twiddle omega^{-c n}
#pragma unroll
        for (int n = 1; n < L; ++n) {
          const double wv = n == 1 ? stp.x : mm_shoup(cur.x, stp.x, stp.y, P.p);
          cur = n == 1 ? stp : make_double2(wv, wv * P.pinv);
          v[n] = mm_shoup(v[n], cur.x, cur.y, P.p);
        }
      }
      big_dft<L>(v, P, MODE == kInv);
#pragma unroll
      for (int k = 0; k < L; ++k) {
        double x = v[k];
        if constexpr (MODE == kFwd) {
          // post-twiddle omega^{c k}
          if (k == 1) cur = stp;
          if (k >= 2) {
            const double wv = mm_shoup(cur.x, stp.x, stp.y, P.p);
            cur = make_double2(wv, wv * P.pinv);
          }
          x = k == 0 ? reduce(x, P.p, P.pinv) : mm_shoup(x, cur.x, cur.y, P.p);
        } else {
          x = reduce(x, P.p, P.pinv);
        }
        dst[base + ((uint64_t)k << logS) + col] = x;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// First forward pass from an operand's limbs (column DFTs of L = 2^LOGL
// points), as two half-length DFTs per tile.
//
// An operand that fills at most the lower half of the transform
// (nlimbs <= N / 2: a balanced product's operands, the shorter one of any
// product) has inputs only at rows n < L / 2 of each column, so its DFT
// splits into two DFTs of L / 2 points:
//   Y[2j]     = sum_{n < L/2} x_n omega_{L/2}^{n j}
//   Y[2j + 1] = sum_{n < L/2} (x_n omega_L^n) omega_{L/2}^{n j}.
// A 2^14-point LDS tile holds L / 2 rows of 2^(15 - LOGL) adjacent columns
// (16 for 2^11-point columns: every store covers whole 128-byte lines,
// where k_ntt_col<kFwd, 11> stores 64-byte runs), and the two half-length
// DFTs run one after the other through the stages of k_ntt_col on the
// same 16 limbs per thread.  Each output row k = 2j + s gets the inter-pass
// twiddle omega_{L s}^{c k} of k_ntt_col, so the output is the same
// transform.  The limbs of the next tile (16 words per thread) are loaded
// while the current tile's stages run.
//
// Batched launches (descs) take each product's operand and length.  A
// product whose operand is longer than N / 2 has upper rows: each DFT's
// inputs then add (x_n + x_{n+L/2}) or subtract (x_n - x_{n+L/2}) them,
// the upper row's limbs loaded again for each (uniform per workgroup,
// latency not hidden).
template <int LOGL>
__global__ __launch_bounds__(kBigNTH) void k_ntt_col_limbs(PassArgs a) {
  using GE = BigGeo<LOGL - 1>;  // one half-length line per column
  constexpr int LH = GE::L;      // L / 2
  constexpr int LOGR1 = GE::LOGR1;
  constexpr int R1 = 1 << LOGR1;
  constexpr int NB1 = kBigEPT / R1;
  constexpr int LOGTA = LOGL - 1 - 4;
  static_assert(kBigNBL == 1, "one last-stage butterfly per thread");
  __shared__ double lds[GE::LDS];
  __shared__ BigTw ta[1 << LOGTA];
  __shared__ double tp[LH];  // omega_L^n, n < L / 2 (odd rows' pre-twiddle, last stage)
  const int pi = blockIdx.y;
  const Prime P = a.prm[pi];
  big_tables<GE>(ta, nullptr, P);
  for (int e = threadIdx.x; e < LH; e += blockDim.x) tp[e] = root_pow((uint64_t)e, LOGL, false, P).x;
  const int logS = a.logS;
  uint64_t zoff = 0;
  const uint32_t* limbs = a.limbs;
  uint64_t nlimbs = a.nlimbs;
  if (a.descs) {
    const uint64_t zi = (uint64_t)blockIdx.z << a.zshift;
    const BatchDesc d = a.descs[zi];
    zoff = zi * a.zstride;
    limbs = a.operand ? d.b : d.a;
    nlimbs = a.operand ? d.nb : d.na;
  }
  double* dst = a.dst + zoff + pi * a.N;
  const int logRoot = LOGL + logS;
  const uint64_t ntiles = a.ntiles;  // tiles of 2^(15 - LOGL) columns
  const uint64_t half = (uint64_t)LH << logS;  // N / 2: the upper rows' offset
  const bool full = nlimbs > half;

  auto tile_base = [&](uint64_t t) -> uint64_t {
    const int lgc = logS - GE::LOGLINES;
    return ((t >> lgc) << (LOGL + logS)) + ((t & ((1ull << lgc) - 1)) << GE::LOGLINES);
  };
  // limbs of this thread's first-stage inputs (rows n < L / 2); limbs past
  // the operand read as zero (range-checked buffer loads, all in flight
  // together; transforms of at most 2^28 points)
  const LimbBuf lb(limbs, (uint32_t)nlimbs);
  auto load_tile = [&](uint64_t t, uint32_t (&x)[kBigEPT]) {
    const uint64_t b0 = tile_base(t);
    const int bt = big_tid();
#pragma unroll
    for (int b = 0; b < NB1; ++b)
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        int lam, n;
        big_pos<GE, LOGR1>(bt, b, r, lam, n);
        x[b * R1 + r] = lb[(uint32_t)(b0 + ((uint64_t)n << logS) + lam)];
      }
  };

  auto tile_step = [&](const uint64_t w, const uint32_t (&xc)[kBigEPT], uint32_t (&xn)[kBigEPT],
                       const uint64_t wload) {
    const uint64_t base = tile_base(w);
    const int bt = big_tid();
    int lam_o, n_o0;
    big_pos<GE, 4>(bt, 0, 0, lam_o, n_o0);
    const uint64_t c = (base & ((1ull << logS) - 1)) + (uint64_t)lam_o;
    // output row k = 2 (n_o0 + r L / 32) + s: twiddle omega^{c k} from
    // omega^{c (2 n_o0 + s)} by steps of omega^{c L / 16}
    const double2 fstp = root_pow(c * (uint64_t)(GE::L >> 3), logRoot, false, P);
#pragma unroll 1
    for (int s = 0; s < 2; ++s) {
      double2 cur = root_pow(c * (uint64_t)(2 * n_o0 + s), logRoot, false, P);
      double v1[NB1][R1];
#pragma unroll
      for (int b = 0; b < NB1; ++b)
#pragma unroll
        for (int r = 0; r < R1; ++r) v1[b][r] = (double)xc[b * R1 + r];
      if (full) {
        const int bt2 = big_tid();
#pragma unroll
        for (int b = 0; b < NB1; ++b)
#pragma unroll
          for (int r = 0; r < R1; ++r) {
            int lam, n;
            big_pos<GE, LOGR1>(bt2, b, r, lam, n);
            const double xh = (double)lb[(uint32_t)(base + ((uint64_t)n << logS) + lam + half)];
            v1[b][r] = s == 0 ? v1[b][r] + xh : v1[b][r] - xh;
          }
      }
      if (s == 1) {
#pragma unroll
        for (int b = 0; b < NB1; ++b)
#pragma unroll
          for (int r = 0; r < R1; ++r) {
            int lam, n;
            big_pos<GE, LOGR1>(big_tid(), b, r, lam, n);
            const double wv = tp[n];
            v1[b][r] = mm_shoup(v1[b][r], wv, wv * P.pinv, P.p);
          }
      }
      big_compute<GE, LOGR1, 0>(v1, ta, P, false);
      __syncthreads();  // LDS readers of the previous half-length DFT are done
      big_write<GE, LOGR1, 0>(v1, lds);
      __syncthreads();
      if (s == 0 && wload < ntiles) load_tile(wload, xn);
      big_middle<GE, LOGR1>(lds, ta, P, false);
      double y[kBigNBL][16];
      big_last_th<GE>(lds, y, tp, P);  // tp[n] = omega_L^n = omega_{2 LH}^n
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const double x = mm_shoup(y[0][r], cur.x, cur.y, P.p);
        const double wv = mm_shoup(cur.x, fstp.x, fstp.y, P.p);
        cur = make_double2(wv, wv * P.pinv);
        const uint64_t k = (uint64_t)(2 * (n_o0 + r * (GE::L >> 4)) + s);
        dst[base + (k << logS) + lam_o] = x;
      }
    }
  };

  // tile order as in k_ntt_col (XCD by XCD)
  const uint64_t g = gridDim.x;
  uint64_t w = blockIdx.x;
  if ((g & 7) == 0) w = (uint64_t)(blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
  if (w >= ntiles) return;
  uint32_t xc[kBigEPT];
  load_tile(w, xc);
  for (;;) {
    uint32_t xn[kBigEPT];
    tile_step(w, xc, xn, w + g);
    w += g;
    if (w >= ntiles) break;
#pragma unroll
    for (int i = 0; i < kBigEPT; ++i) xc[i] = xn[i];
  }
}

// ---------------------------------------------------------------------------
// First forward pass from the limbs for 2^11-point columns with one
// workgroup-wide exchange per half-length DFT (k_ntt_col_limbs<11> takes
// three stages through LDS, four workgroup barriers each).
//
// A tile is 16 adjacent columns cl of 1024 rows (the radix-2 split of
// k_ntt_col_limbs).  Each half-length DFT of 1024 points runs as
//   S1: DFT16 over jj for each m (n = m + 64 jj), twiddle omega_1024^{k1 m}
//   S2: DFT16 over i for each (k1, m0) (m = m0 + 4 i), twiddle omega_64^{k2 m0}
//   S3: DFT4 over m0 for each (k1, k2)
// giving Z[k1 + 16 k2 + 256 k3].  Thread t = 64 w + l runs S1 on column
// t & 15, m = t >> 4 (its 16 loads: 4 rows x 16 columns per wave access);
// one exchange between two workgroup barriers gives wave w all of k1 = w
// for the 16 columns (lane l: column l & 15, m0 = l >> 4), and the
// exchange before S3 stays inside the wave's own 1024 LDS doubles (ordered
// by its in-order LDS queue).  After S3 lane l holds, for its column,
// k2 = 4 (l >> 4) + s' and k3 = 0 .. 3: every store covers 4 rows x 16
// columns (128-byte lines).  All LDS layouts are conflict-free for each
// 32-lane half of a 64-bit access.
// Global-memory views of pointers the compiler cannot place itself (they
// come from argument structs or descriptors): loads through them are global
// loads, which wait on the vector-memory counter only (flat loads also
// hold up the LDS counter).
typedef const __attribute__((address_space(1))) uint32_t GlobalU32;
typedef const __attribute__((address_space(1))) double GlobalF64c;
typedef __attribute__((address_space(1))) double GlobalF64;
// p advanced by boff bytes.  With p wave-uniform and boff one 32-bit
// offset per lane this is the global access form with a scalar base: one
// VGPR of address per thread instead of a 64-bit address per access.
__device__ __forceinline__ const GlobalU32* at_bytes(const GlobalU32* p, uint32_t boff) {
  return (const GlobalU32*)((const __attribute__((address_space(1))) char*)p + boff);
}
__device__ __forceinline__ GlobalF64c* at_bytes(GlobalF64c* p, uint32_t boff) {
  return (GlobalF64c*)((const __attribute__((address_space(1))) char*)p + boff);
}
__device__ __forceinline__ GlobalF64* at_bytes(GlobalF64* p, uint32_t boff) {
  return (GlobalF64*)((__attribute__((address_space(1))) char*)p + boff);
}

__device__ __forceinline__ double col11_tp(const double* tp, int e2) {
  // omega_1024^e = omega_2048^{2e} from tp[n] = omega_2048^n (n < 1024):
  // e2 = 2e < 2048, the upper half negated
  const double t = tp[e2 & 1023];
  return (e2 & 1024) ? -t : t;
}

__global__ __launch_bounds__(kBigNTH) void k_ntt_col_limbs11w(PassArgs a) {
  constexpr int LOGL = 11;
  constexpr int LH = 1 << (LOGL - 1);  // 1024 rows per tile
  __shared__ double xs[kBigE];
  __shared__ double tp[LH];      // omega_2048^n: the odd half's pre-twiddle, S1 twiddles
  __shared__ BigTw t64[64];      // omega_64^e: S2 twiddles
  const int pi = blockIdx.y;
  const Prime P = a.prm[pi];
  for (int e = threadIdx.x; e < LH; e += blockDim.x) tp[e] = root_pow((uint64_t)e, LOGL, false, P).x;
  for (int e = threadIdx.x; e < 64; e += blockDim.x) t64[e] = P.twL[e << (kLogE - 6)].x;
  const int logS = a.logS;
  uint64_t zoff = 0;
  const uint32_t* limbs = a.limbs;
  uint64_t nlimbs = a.nlimbs;
  if (a.descs) {
    const uint64_t zi = (uint64_t)blockIdx.z << a.zshift;
    const BatchDesc d = a.descs[zi];
    zoff = zi * a.zstride;
    limbs = a.operand ? d.b : d.a;
    nlimbs = a.operand ? d.nb : d.na;
  }
  double* dst = a.dst + zoff + pi * a.N;
  const int logRoot = LOGL + logS;
  const uint64_t ntiles = a.ntiles;            // tiles of 16 columns
  const uint64_t half = (uint64_t)LH << logS;  // N / 2: the upper rows' offset
  const bool full = nlimbs > half;

  auto tile_base = [&](uint64_t t) -> uint64_t {
    const int lgc = logS - 4;
    return ((t >> lgc) << (LOGL + logS)) + ((t & ((1ull << lgc) - 1)) << 4);
  };
  // S1 inputs of this thread: rows m + 64 jj of column t & 15.  A tile
  // wholly inside the operand loads them plainly (all 16 in flight
  // together); else limbs past the operand read as zero (range-checked
  // buffer loads).  The limbs are global memory: global loads wait on the
  // vector-memory counter only.
  const GlobalU32* lp = (const GlobalU32*)limbs;  // address-space cast
  const LimbBuf lb(limbs, (uint32_t)nlimbs);       // the ragged blocks: zero past the operand
  // rows m + 64 jj (jj < 16) of this thread's column in the block of 1024
  // rows at u (wave-uniform; a whole block loads at a scalar base plus one
  // 32-bit lane offset)
  auto load_rows = [&](uint64_t u, uint32_t (&x)[16]) {
    const int bt = big_tid();
    const uint32_t lo = ((uint32_t)(bt >> 4) << logS) + (uint32_t)(bt & 15);
    if (u + ((uint64_t)(LH - 1) << logS) + 15 < nlimbs) {
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) x[jj] = *at_bytes(lp + u + ((uint64_t)jj << (logS + 6)), lo * 4u);
    } else {
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) x[jj] = lb[(uint32_t)(u + lo + ((uint64_t)jj << (logS + 6)))];
    }
  };
  auto load_tile = [&](uint64_t t, uint32_t (&x)[16]) { load_rows(tile_base(t), x); };

  auto tile_step = [&](const uint64_t w, const uint32_t (&xc)[16], uint32_t (&xn)[16], const uint64_t wload) {
    const uint64_t base = tile_base(w);
#pragma unroll 1
    for (int s = 0; s < 2; ++s) {
      double v[16];
      {
        const int bt = big_tid();
        const int m = bt >> 4;
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) v[jj] = (double)xc[jj];
        if (full) {
          uint32_t uh[16];
          load_rows(base + half, uh);
#pragma unroll
          for (int jj = 0; jj < 16; ++jj) v[jj] = s == 0 ? v[jj] + (double)uh[jj] : v[jj] - (double)uh[jj];
        }
        if (s == 1) {
#pragma unroll
          for (int jj = 0; jj < 16; ++jj) {
            const double wv = tp[m + 64 * jj];
            v[jj] = mm_shoup(v[jj], wv, wv * P.pinv, P.p);
          }
        }
        // S1 and its twiddles omega_1024^{k1 m}
        dft16(v, P, false);
        v[0] = reduce(v[0], P.p, P.pinv);
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) {
          const double wv = col11_tp(tp, 2 * m * k1);
          v[k1] = mm_shoup(v[k1], wv, wv * P.pinv, P.p);
        }
      }
      __syncthreads();  // every wave is done with the LDS (previous half)
      {
        double* xw = xs + big_tid();  // [k1][m][cl]
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) xw[k1 << 10] = v[k1];
      }
      __syncthreads();
      if (s == 0 && wload < ntiles) load_tile(wload, xn);
      const int bt = big_tid();
      const int wv_ = bt >> 6, lane = bt & 63;
      double* R = xs + (wv_ << 10);  // this wave's 1024 doubles: k1 = wave
      {
        // S2 over i (m = m0 + 4 i), twiddles omega_64^{k2 m0}
        const int m0 = lane >> 4;
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = R[lane + 64 * i];
        dft16(v, P, false);
        v[0] = reduce(v[0], P.p, P.pinv);
#pragma unroll
        for (int k2 = 1; k2 < 16; ++k2) {
          const double2 t = tw_of(t64, m0 * k2, P);
          v[k2] = mm_shoup(v[k2], t.x, t.y, P.p);
        }
        wave_lds_sync();
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) R[k2 * 64 + ((m0 ^ (k2 >> 2)) << 4) + (lane & 15)] = v[k2];
        wave_lds_sync();
        // S3: lane (column lane & 15, q = lane >> 4) takes k2 = 4 q + s'
        const int q = lane >> 4;
#pragma unroll
        for (int sp = 0; sp < 4; ++sp) {
#pragma unroll
          for (int mm = 0; mm < 4; ++mm) v[4 * sp + mm] = R[(4 * q + sp) * 64 + ((mm ^ q) << 4) + (lane & 15)];
          dft_core<4>(v + 4 * sp, P, false);
        }
      }
      // v[4 s' + k3] is Z[j] of column lane & 15, j = wave + 16 (4 q + s') +
      // 256 k3, i.e. output row k = 2 j + s; the inter-pass twiddle
      // omega^{c k} along s' (step omega^{32 c}) and k3 (step omega^{512 c})
      {
        const int q = lane >> 4;
        const uint64_t c = (base & ((1ull << logS) - 1)) + (uint64_t)(lane & 15);
        const uint64_t k0 = (uint64_t)(2 * wv_ + 128 * q + s);
        const double2 st = root_pow(c * 32, logRoot, false, P);
        const double2 sk = root_pow(c * 512, logRoot, false, P);
        double2 row = root_pow(c * k0, logRoot, false, P);
        GlobalF64* db = (GlobalF64*)dst + base;  // address-space cast
        const uint32_t so = (((uint32_t)k0 << logS) + (uint32_t)(lane & 15)) * 8u;
#pragma unroll
        for (int k3 = 0; k3 < 4; ++k3) {
          double2 cur = row;
#pragma unroll
          for (int sp = 0; sp < 4; ++sp) {
            const double x = mm_shoup(v[4 * sp + k3], cur.x, cur.y, P.p);
            *at_bytes(db + ((uint64_t)(32 * sp + 512 * k3) << logS), so) = x;
            if (sp < 3) {
              const double wn = mm_shoup(cur.x, st.x, st.y, P.p);
              cur = make_double2(wn, wn * P.pinv);
            }
          }
          if (k3 < 3) {
            const double wn = mm_shoup(row.x, sk.x, sk.y, P.p);
            row = make_double2(wn, wn * P.pinv);
          }
        }
      }
    }
  };

  // tile order as in k_ntt_col (XCD by XCD)
  const uint64_t g = gridDim.x;
  uint64_t w = blockIdx.x;
  if ((g & 7) == 0) w = (uint64_t)(blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
  if (w >= ntiles) return;
  uint32_t xc[16];
  load_tile(w, xc);
  __syncthreads();  // tables
  for (;;) {
    uint32_t xn[16];
    tile_step(w, xc, xn, w + g);
    w += g;
    if (w >= ntiles) break;
#pragma unroll
    for (int i = 0; i < 16; ++i) xc[i] = xn[i];
  }
}

constexpr int kCol16LogL = 11;

// ---------------------------------------------------------------------------
// Inverse 2^11-point column pass over 16 adjacent columns with one
// workgroup-wide exchange per half-length DFT: the stages of
// k_ntt_col_limbs11w run backwards in direction (inverse DFTs, conjugate
// twiddles).  Its tiles are 16 adjacent columns (k_ntt_col<kInv, 11> takes
// 8 per 2^14-point tile and moves 64-byte runs, half a 128-byte line per
// access; it also takes three stages through LDS, four workgroup barriers
// each).
//
// Column c of a tile holds x[n], n < 2048.  Its inverse DFT with the
// inter-pass twiddle omega_N^{-c n} splits (decimation in frequency) into
// two 1024-point inverse DFTs: output row 2 k + s is
//   sum_{n < 1024} omega_1024^{-n k} y_s[n],
//   y_s[n] = omega_N^{-(c + s S) n} (x[n] + (-1)^s u[n]),
//   u[n] = omega_N^{-1024 c} x[n + 1024],
// with S = 2^logS (omega_2048 = omega_N^S).  Thread t = 64 w + l loads rows
// m + 64 jj and m + 64 jj + 1024 of column t & 15 (m = t >> 4; each wave
// access is 4 rows x 16 columns, whole 128-byte lines) and forms y_0 and
// y_1 in registers.  Each y_s then runs S1, the workgroup exchange, S2 and
// S3 of k_ntt_col_limbs11w, and lane l stores its column's rows reduced.
// The next tile's low rows are loaded after half 0's exchange and its high
// rows after half 1's, so the loads overlap the wave-local stages.  The
// global-address-space views keep the loads and stores off the LDS
// counter: a flat access in flight would be waited for at every LDS wait.
template <bool B>
struct BoolTag {
  static constexpr bool value = B;
};

// omega_1024^{-e} from t[n] = omega_1024^n, n < 512 (omega_1024^{n + 512}
// = -omega_1024^n)
__device__ __forceinline__ double w1024_conj(const double* t, int e) {
  const int f = (-e) & 1023;
  const double w = t[f & 511];
  return (f & 512) ? -w : w;
}

__global__ __launch_bounds__(kBigNTH) void k_ntt_col16_invw(PassArgs a) {
  constexpr int LOGL = kCol16LogL;
  static_assert(LOGL == 11 && kBigE == (1 << 14) && kBigNTH == 1024, "tiles of 16 columns x 2048 rows");
  __shared__ double xs[kBigE];
  __shared__ double t1k[512];  // omega_1024^n: S1 twiddles
  __shared__ BigTw t64[64];    // omega_64^e: S2 twiddles
  const int pi = blockIdx.y;
  const Prime P = a.prm[pi];
  for (int e = threadIdx.x; e < 512; e += blockDim.x) t1k[e] = P.twL[e << (kLogE - 10)].x;
  for (int e = threadIdx.x; e < 64; e += blockDim.x) t64[e] = P.twL[e << (kLogE - 6)].x;
  const int logS = a.logS;
  const uint64_t S = 1ull << logS;
  const uint64_t zoff = a.descs ? ((uint64_t)blockIdx.z << a.zshift) * a.zstride : 0;
  GlobalF64c* src = (GlobalF64c*)(a.src + zoff + pi * a.N);  // address-space casts
  GlobalF64* dst = (GlobalF64*)(a.dst + zoff + pi * a.N);
  const int logRoot = LOGL + logS;
  const uint64_t ntiles = a.ntiles;  // tiles of 16 columns

  auto tile_base = [&](uint64_t t) -> uint64_t {
    const int lgc = logS - 4;
    return ((t >> lgc) << (LOGL + logS)) + ((t & ((1ull << lgc) - 1)) << 4);
  };
  // Addresses are a wave-uniform base plus one 32-bit byte offset per lane
  // (at_bytes; a tile spans 2^(11 + logS) <= 2^28 points).
  // byte offset of row r, column cl in a tile
  auto lane_off = [&](uint32_t r, int cl) -> uint32_t { return ((r << logS) + (uint32_t)cl) * 8u; };
  // rows m + 64 jj + 1024 hi (jj < 16) of this thread's column
  auto load_rows = [&](uint64_t t, int hi, double (&x)[16]) {
    const int bt = big_tid();
    const uint32_t off = lane_off((uint32_t)(bt >> 4), bt & 15);
    GlobalF64c* p = src + tile_base(t) + ((uint64_t)hi << (10 + logS));
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) x[jj] = *at_bytes(p + ((uint64_t)jj << (logS + 6)), off);
  };

  // xl, xh: rows m + 64 jj and m + 64 jj + 1024 of tile w, replaced by
  // y_0 and y_1 (in place); on return (PF), the same rows of tile wn
  auto tile_step = [&](auto prefetch, const uint64_t w, double (&xl)[16], double (&xh)[16], const uint64_t wn) {
    constexpr bool PF = decltype(prefetch)::value;
    const uint64_t base = tile_base(w);
    {
      // y_0 (in xl) and y_1 (in xh) with their twiddles
      // omega_N^{-(c + s S) n} along jj (steps omega_N^{-64 (c + s S)})
      const int bt = big_tid();
      const uint64_t m = (uint64_t)(bt >> 4);
      const uint64_t c = (base & (S - 1)) + (uint64_t)(bt & 15);
      const double2 th = root_pow(c << 10, logRoot, true, P);
      const double2 st0 = root_pow(c << 6, logRoot, true, P);
      const double2 st1 = root_pow((c + S) << 6, logRoot, true, P);
      double2 c0 = root_pow(c * m, logRoot, true, P);
      double2 c1 = root_pow((c + S) * m, logRoot, true, P);
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        const double u = mm_shoup(xh[jj], th.x, th.y, P.p);
        const double lo = xl[jj];
        xl[jj] = mm_shoup(lo + u, c0.x, c0.y, P.p);
        xh[jj] = mm_shoup(lo - u, c1.x, c1.y, P.p);
        if (jj < 15) {
          const double w0 = mm_shoup(c0.x, st0.x, st0.y, P.p);
          const double w1 = mm_shoup(c1.x, st1.x, st1.y, P.p);
          c0 = make_double2(w0, w0 * P.pinv);
          c1 = make_double2(w1, w1 * P.pinv);
        }
      }
    }
    // both halves unrolled: the next tile's rows land in registers that
    // are dead from there on (a loop would keep xh live through half 0)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      double v[16];
      {
        const int m = big_tid() >> 4;
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) v[jj] = s == 0 ? xl[jj] : xh[jj];
        // S1 and its twiddles omega_1024^{-k1 m}
        dft16(v, P, true);
        v[0] = reduce(v[0], P.p, P.pinv);
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) {
          const double wv = w1024_conj(t1k, m * k1);
          v[k1] = mm_shoup(v[k1], wv, wv * P.pinv, P.p);
        }
      }
      __syncthreads();  // every wave is done with the LDS (previous half)
      {
        double* xw = xs + big_tid();  // [k1][m][cl]
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) xw[k1 << 10] = v[k1];
      }
      __syncthreads();
      if constexpr (PF) load_rows(wn, s, s == 0 ? xl : xh);
      const int bt = big_tid();
      const int wv_ = bt >> 6, lane = bt & 63;
      double* R = xs + (wv_ << 10);  // this wave's 1024 doubles: k1 = wave
      {
        // S2 over i (m = m0 + 4 i), twiddles omega_64^{-k2 m0}
        const int m0 = lane >> 4;
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = R[lane + 64 * i];
        dft16(v, P, true);
        v[0] = reduce(v[0], P.p, P.pinv);
#pragma unroll
        for (int k2 = 1; k2 < 16; ++k2) {
          const double2 t = tw_of(t64, (64 - m0 * k2) & 63, P);
          v[k2] = mm_shoup(v[k2], t.x, t.y, P.p);
        }
        wave_lds_sync();
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) R[k2 * 64 + ((m0 ^ (k2 >> 2)) << 4) + (lane & 15)] = v[k2];
        wave_lds_sync();
        // S3: lane (column lane & 15, q = lane >> 4) takes k2 = 4 q + s'
        const int q = lane >> 4;
#pragma unroll
        for (int sp = 0; sp < 4; ++sp) {
#pragma unroll
          for (int mm = 0; mm < 4; ++mm) v[4 * sp + mm] = R[(4 * q + sp) * 64 + ((mm ^ q) << 4) + (lane & 15)];
          dft_core<4>(v + 4 * sp, P, true);
        }
      }
      // v[4 s' + k3] is output row 2 (wave + 16 (4 q + s') + 256 k3) + s of
      // column lane & 15
      {
        const int q = lane >> 4;
        const uint32_t off = lane_off((uint32_t)(2 * wv_ + 128 * q + s), lane & 15);
#pragma unroll
        for (int k3 = 0; k3 < 4; ++k3)
#pragma unroll
          for (int sp = 0; sp < 4; ++sp) {
            *at_bytes(dst + base + ((uint64_t)(32 * sp + 512 * k3) << logS), off) = reduce(v[4 * sp + k3], P.p, P.pinv);
          }
      }
    }
  };

  // tile order as in k_ntt_col (XCD by XCD)
  const uint64_t g = gridDim.x;
  uint64_t w = blockIdx.x;
  if ((g & 7) == 0) w = (uint64_t)(blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
  if (w >= ntiles) return;
  double xl[16], xh[16];
  load_rows(w, 0, xl);
  load_rows(w, 1, xh);
  __syncthreads();  // tables
  for (;;) {
    const uint64_t wn = w + g;
    if (wn >= ntiles) break;
    tile_step(BoolTag<true>{}, w, xl, xh, wn);
    w = wn;
  }
  tile_step(BoolTag<false>{}, w, xl, xh, w);
}

}  // namespace pcg
