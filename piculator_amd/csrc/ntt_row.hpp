// Row pass of the large-tile plan (2^14-point lines, stride 1) with one
// cross-wave exchange per transform.
//
// A 2^14-point DFT is four digit stages, decimation in frequency with the
// digits of n = n1 2^10 + n2 2^6 + n3 2^2 + n4 taken from the top:
//   S1: DFT16 over n1 for each m = n mod 2^10, twiddle omega_{2^14}^{k1 m}
//   S2: DFT16 over n2 for each (k1, m2 = n mod 64), twiddle omega_1024^{k2 m2}
//   S3: DFT16 over n3 for each (k1, k2, n4), twiddle omega_64^{k3 n4}
//   S4: DFT4 over n4,
// giving X[k1 + 16 k2 + 256 k3 + 4096 k4].  Thread t = 64 w + lane of the
// 1024 holds 16 points.  S1 runs on thread t = m (its loads are 16
// coalesced rows of the line); its outputs go to wave k1 through LDS (one
// workgroup barrier on each side), after which wave k1 holds all of
// subproblem k1, so the exchanges before S3 and S4 stay inside the wave:
// each wave rewrites only the 1024 LDS doubles it has just read, ordered by
// the wave's own in-order LDS queue, with no workgroup barrier.  The
// inverse runs the same stages mirrored: twiddle^-1 then DFT^-1 in each
// stage's layout, exchanges reversed.
//
// The spectrum stays in the thread layout of S4: register s 4 + k4 of
// thread (k1 = w, lane = 4 k2 + q) holds frequency k1 + 16 k2 +
// 256 (4 q + s) + 4096 k4 of the line; the pointwise product does not
// depend on the order, and the inverse starts from the same layout.
//
// LDS layouts (doubles; conflict-free for every 16- and 32-lane group of a
// 64-bit access):
//   cross-wave:  [k1][n2][lane]                 = xs[(k1 16 + n2) 64 + lane]
//   before S3:   wave w, row n3, column 4 k2 + n4 swizzled by n3 << 2
//   before S4:   wave w, row 4 s + n4, column 4 k2 + q swizzled by n4
#pragma once
#include "ntt_big.hpp"

#pragma clang fp contract(off)

namespace pcg {

constexpr int kRowNTH = kBigNTH;  // 1024 threads x 16 points
static_assert(kBigEPT == 16, "the row pass holds 16 points per thread");

// Stage-twiddle tables in LDS (i = 0 forward, 1 inverse: exponent sign s =
// +1, -1).  Each stage's table is laid out by (k, lane index) so that the
// lanes of one read touch consecutive or equal entries (no bank conflicts;
// a table indexed by k m2 would put lanes k entries apart) and its address
// is a constant offset from the lane's base.
//   ta[e]        = omega_1024^e                 (e < 1024)  S1, factor a
//   t1[i][k][b]  = omega_{2^14}^{s k b}         (b < 16)    S1, factor b
//   t2[i][k][m2] = omega_1024^{s k m2}          (m2 < 64)   S2
//   t3[i][k][n4] = omega_64^{s k n4}            (n4 < 4)    S3
struct RowTw {
  BigTw ta[1024];
  double t1[2][16][16];
  double t2[2][16][64];
  double t3[2][16][4];
};

// Fills the tables (all threads of the workgroup; a barrier must follow
// before the first use).
__device__ __forceinline__ void row_tw_init(RowTw& T, const Prime& P) {
  const int nth = blockDim.x;
  for (int e = threadIdx.x; e < 1024; e += nth) T.ta[e] = P.twL[e << (kLogE - 10)].x;
  for (int x = threadIdx.x; x < 512; x += nth) {
    const int i = x >> 8, k = (x >> 4) & 15, b = x & 15;
    T.t1[i][k][b] = root_pow((uint64_t)(k * b), kBigLogE, i != 0, P).x;
  }
  for (int x = threadIdx.x; x < 2048; x += nth) {
    const int i = x >> 10, k = (x >> 6) & 15, m2 = x & 63;
    int e = (k * m2) & 1023;
    if (i) e = (1024 - e) & 1023;
    T.t2[i][k][m2] = P.twL[e << (kLogE - 10)].x;
  }
  for (int x = threadIdx.x; x < 128; x += nth) {
    const int i = x >> 6, k = (x >> 2) & 15, n4 = x & 3;
    int e = (16 * k * n4) & 1023;
    if (i) e = (1024 - e) & 1023;
    T.t3[i][k][n4] = P.twL[e << (kLogE - 10)].x;
  }
}

// S1 twiddles: v[k1] *= omega_{2^14}^{+-k1 m}, m = 16 a + b, as
// omega_{2^14}^{b k1} (b k1 < 256) times omega_1024^{a k1} (a k1 < 1024).
__device__ __forceinline__ void row_s1_twiddle(double (&v)[16], int m, const RowTw& T, const Prime& P,
                                               bool inv) {
  const int a = m >> 4, b = m & 15;
  const int i = inv ? 1 : 0;
  v[0] = reduce(v[0], P.p, P.pinv);
#pragma unroll
  for (int k = 1; k < 16; ++k) {
    const double wb = T.t1[i][k][b];
    const int ea = (a * k) & 1023;
    const double2 wa = tw_of(T.ta, inv ? (1024 - ea) & 1023 : ea, P);
    const double x = mm_shoup(v[k], wb, wb * P.pinv, P.p);
    v[k] = mm_shoup(x, wa.x, wa.y, P.p);
  }
}

// S2 twiddles: v[k2] *= omega_1024^{+-k2 m2}.
__device__ __forceinline__ void row_s2_twiddle(double (&v)[16], int m2, const RowTw& T, const Prime& P,
                                               bool inv) {
  const double* t = T.t2[inv ? 1 : 0][0] + m2;
  v[0] = reduce(v[0], P.p, P.pinv);
#pragma unroll
  for (int k = 1; k < 16; ++k) {
    const double w = t[k * 64];
    v[k] = mm_shoup(v[k], w, w * P.pinv, P.p);
  }
}

// S3 twiddles: v[k3] *= omega_64^{+-k3 n4}.
__device__ __forceinline__ void row_s3_twiddle(double (&v)[16], int n4, const RowTw& T, const Prime& P,
                                               bool inv) {
  const double* t = T.t3[inv ? 1 : 0][0] + n4;
  v[0] = reduce(v[0], P.p, P.pinv);
#pragma unroll
  for (int k = 1; k < 16; ++k) {
    const double w = t[k * 4];
    v[k] = mm_shoup(v[k], w, w * P.pinv, P.p);
  }
}

// Exchange before S3 (forward direction): this lane (n3 = lane >> 2,
// n4 = lane & 3) holds k2 = 0..15 and receives n3 = 0..15 as lane 4 k2 + n4.
// R is the wave's 1024 doubles.
__device__ __forceinline__ void row_w1_put(const double (&v)[16], double* R, int lane) {
  const int n3 = lane >> 2, n4 = lane & 3;
  double* rb = R + n3 * 64 + n4;
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) rb[(k2 ^ n3) << 2] = v[k2];
}
__device__ __forceinline__ void row_w1_get(double (&v)[16], const double* R, int lane) {
#pragma unroll
  for (int n3 = 0; n3 < 16; ++n3) v[n3] = R[n3 * 64 + (lane ^ (n3 << 2))];
}
// Exchange before S4: this lane (k2 = lane >> 2, n4 = lane & 3) holds
// k3 = 4 q + s and receives register 4 s + n4 as lane 4 k2 + q.
__device__ __forceinline__ void row_w2_put(const double (&v)[16], double* R, int lane) {
  const int k2 = lane >> 2, n4 = lane & 3;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double* rb = R + n4 * 64 + k2 * 4 + (q ^ n4);
#pragma unroll
    for (int s = 0; s < 4; ++s) rb[s * 256] = v[4 * q + s];
  }
}
__device__ __forceinline__ void row_w2_get(double (&v)[16], const double* R, int lane) {
#pragma unroll
  for (int row = 0; row < 16; ++row) v[row] = R[row * 64 + (lane ^ (row & 3))];
}
// The same two exchanges in the inverse direction (put where the forward
// got, get where it put).
__device__ __forceinline__ void row_w2_put_inv(const double (&v)[16], double* R, int lane) {
#pragma unroll
  for (int row = 0; row < 16; ++row) R[row * 64 + (lane ^ (row & 3))] = v[row];
}
__device__ __forceinline__ void row_w2_get_inv(double (&v)[16], const double* R, int lane) {
  const int k2 = lane >> 2, n4 = lane & 3;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double* rb = R + n4 * 64 + k2 * 4 + (q ^ n4);
#pragma unroll
    for (int s = 0; s < 4; ++s) v[4 * q + s] = rb[s * 256];
  }
}
__device__ __forceinline__ void row_w1_put_inv(const double (&v)[16], double* R, int lane) {
#pragma unroll
  for (int n3 = 0; n3 < 16; ++n3) R[n3 * 64 + (lane ^ (n3 << 2))] = v[n3];
}
__device__ __forceinline__ void row_w1_get_inv(double (&v)[16], const double* R, int lane) {
  const int n3 = lane >> 2, n4 = lane & 3;
  const double* rb = R + n3 * 64 + n4;
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) v[k2] = rb[(k2 ^ n3) << 2];
}

// ---------------------------------------------------------------------------
// The fused row pass: the last forward pass of both operands, the pointwise
// product and the first inverse pass.
//
// The pass reads a after its column pass and runs a's row DFT in the tile
// (a's spectrum stays in registers, in the layout the pointwise product
// needs), then b's: a's spectrum never goes through HBM, one sweep of a's
// points fewer per product than a row pass of a of its own (read 16 +
// write 16 bytes per point) for no more arithmetic.
// SQUARE: a and b are one operand (src == spec): one row DFT, the pointwise
// square.  LIMBS: a one-pass plan (N = 2^14, batched): both operands are
// read from their limbs (descs), zero past their lengths.  Per tile, the
// loads of b are issued before a's row DFT and the next tile's a during
// b's cross-wave exchange.

// Forward row DFT (S1 .. S4) of the 16 points t + 1024 r of thread t, into
// the S4 layout; between() runs between the two workgroup barriers of the
// cross-wave exchange.
template <class F>
__device__ __forceinline__ void row_forward(double (&v)[16], double* xs, const RowTw& T, const Prime& P,
                                            F&& between) {
  dft16(v, P, false);
  row_s1_twiddle(v, big_tid(), T, P, false);
  __syncthreads();  // every wave is done with the LDS
  {
    const int t = big_tid();
    double* xw = xs + t;  // [k1][w][lane]
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) xw[k1 << 10] = v[k1];
  }
  __syncthreads();
  between();
  {
    const int t = big_tid();
    const int lane = t & 63;
    const double* xw = xs + ((t >> 6) << 10) + lane;  // [w][n2][lane]
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = xw[n2 << 6];
    dft16(v, P, false);
    row_s2_twiddle(v, lane, T, P, false);
    double* R = xs + ((t >> 6) << 10);
    wave_lds_sync();
    row_w1_put(v, R, lane);
    wave_lds_sync();
    row_w1_get(v, R, lane);
  }
  dft16(v, P, false);
  row_s3_twiddle(v, big_tid() & 3, T, P, false);
  {
    const int t = big_tid();
    double* R = xs + ((t >> 6) << 10);
    wave_lds_sync();
    row_w2_put(v, R, t & 63);
    wave_lds_sync();
    row_w2_get(v, R, t & 63);
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) dft_core<4>(v + 4 * s, P, false);
}

// Inverse row DFT of the pointwise products in v (S4 layout), stored
// reduced at points t + 1024 r of the tile at d.
__device__ __forceinline__ void row_inverse_store(double (&v)[16], double* xs, const RowTw& T, const Prime& P,
                                                  double* d) {
#pragma unroll
  for (int s = 0; s < 4; ++s) dft_core<4>(v + 4 * s, P, true);
  {
    const int t = big_tid();
    double* R = xs + ((t >> 6) << 10);
    wave_lds_sync();
    row_w2_put_inv(v, R, t & 63);
    wave_lds_sync();
    row_w2_get_inv(v, R, t & 63);
  }
  row_s3_twiddle(v, big_tid() & 3, T, P, true);
  dft16(v, P, true);
  {
    const int t = big_tid();
    const int lane = t & 63;
    double* R = xs + ((t >> 6) << 10);
    wave_lds_sync();
    row_w1_put_inv(v, R, lane);
    wave_lds_sync();
    row_w1_get_inv(v, R, lane);
    row_s2_twiddle(v, lane, T, P, true);
  }
  dft16(v, P, true);
  __syncthreads();  // every wave is done with its own rows of LDS
  {
    const int t = big_tid();
    double* xw = xs + ((t >> 6) << 10) + (t & 63);  // [w][n2][lane]
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) xw[n2 << 6] = v[n2];
  }
  __syncthreads();
  {
    const int t = big_tid();
    const double* xw = xs + t;  // [k1][w][lane]
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) v[k1] = xw[k1 << 10];
  }
  row_s1_twiddle(v, big_tid(), T, P, true);
  dft16(v, P, true);
  double* dd = d + big_tid();
#pragma unroll
  for (int r = 0; r < 16; ++r) dd[r << 10] = reduce(v[r], P.p, P.pinv);
}

template <bool SQUARE, bool LIMBS>
__global__ __launch_bounds__(kRowNTH) void k_ntt_row_fa(PassArgs a) {
  static_assert(!(SQUARE && LIMBS), "squares take a's column pass output");
  __shared__ double xs[kBigE];
  __shared__ RowTw T;
  const int pi = blockIdx.y;
  const Prime P = a.prm[pi];
  row_tw_init(T, P);
  uint64_t zoff = 0, soff = 0;  // batched: as k_ntt_row
  BatchDesc bd{};
  if (a.descs) {
    const uint64_t zi = (uint64_t)blockIdx.z << a.zshift;
    zoff = zi * a.zstride;
    soff = ((zi >> a.sshift) << a.sshift) * a.zstride;
    if constexpr (LIMBS) bd = a.descs[zi];
  }
  const double* src = LIMBS ? nullptr : a.src + zoff + pi * a.N;   // b after its column pass
  const double* sra = LIMBS ? nullptr : a.spec + soff + pi * a.N;  // a after its column pass (src when SQUARE)
  double* dst = a.dst + zoff + pi * a.N;
  const uint64_t ntiles = a.ntiles;

  // tile w of operand a (op 0) or b (op 1); LIMBS: the whole line
  auto load_tile = [&](int op, uint64_t w, double (&x)[16]) {
    if constexpr (LIMBS) {
      const LimbBuf lb(op ? bd.b : bd.a, op ? bd.nb : bd.na);  // zero past the operand
      const uint32_t t = (uint32_t)big_tid();
#pragma unroll
      for (int r = 0; r < 16; ++r) x[r] = (double)lb[t + ((uint32_t)r << 10)];
    } else {
      const double* s = (op ? src : sra) + (w << kBigLogE) + big_tid();
#pragma unroll
      for (int r = 0; r < 16; ++r) x[r] = s[r << 10];
    }
  };

  // tile order as in k_ntt_col (XCD by XCD)
  const uint64_t g = gridDim.x;
  uint64_t w = blockIdx.x;
  if ((g & 7) == 0) w = (uint64_t)(blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
  if (w >= ntiles) return;
  double xa[16];
  load_tile(SQUARE ? 1 : 0, w, xa);
  __syncthreads();  // tables
  for (;;) {
    const uint64_t wn = w + g;
    double v[16];
    if constexpr (SQUARE) {
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = xa[r];
      row_forward(v, xs, T, P, [&] {
        if (wn < ntiles) load_tile(1, wn, xa);
      });
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const double x = mm_gen(v[r], reduce(v[r], P.p, P.pinv), P.p, P.pinv);
        v[r] = mm_shoup(x, P.ninv.x, P.ninv.y, P.p);
      }
    } else {
      double xb[16], sa[16];
      load_tile(1, w, xb);
#pragma unroll
      for (int r = 0; r < 16; ++r) sa[r] = xa[r];
      row_forward(sa, xs, T, P, [] {});
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        sa[r] = reduce(sa[r], P.p, P.pinv);
        v[r] = xb[r];
      }
      row_forward(v, xs, T, P, [&] {
        if (wn < ntiles) load_tile(0, wn, xa);
      });
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const double x = mm_gen(v[r], sa[r], P.p, P.pinv);
        v[r] = mm_shoup(x, P.ninv.x, P.ninv.y, P.p);
      }
    }
    row_inverse_store(v, xs, T, P, dst + (w << kBigLogE));
    w = wn;
    if (w >= ntiles) break;
  }
}

}  // namespace pcg
