// Batched products: one workgroup per product, carries resolved inside the
// workgroup, so a whole size class of a tree level is one launch.
//
//  k_batch_school : zL = na + nb <= 1024 limbs.  Thread t forms the column
//                   sums of limbs 4t..4t+3 (exact 96-bit accumulation of
//                   32x32 products) from operands staged in LDS.
//  k_batch_ntt    : zL <= 4096 limbs.  Two-prime NTT of N = 2^ceil(log2 zL)
//                   points entirely in LDS (one line per prime, both primes
//                   at once on separate waves), CRT, carry.
#pragma once
#include "carry_kernels.hpp"
#include "ntt_kernels.hpp"

#pragma clang fp contract(off)

namespace pcg {

constexpr int kSchoolMaxLimbs = 1024;  // zL bound of k_batch_school
constexpr int kSmallNttMaxLog = 12;    // N bound of k_batch_ntt (rows padded by one)

// 96-bit sums (lo 64 bits + carry count) of columns k0 .. k0 + C - 1 over
// a's limbs i = ilo .. ihi: each step multiplies a_i by a window of C limbs
// of b, w[q] = b[k0 + q - i] = sb[4 + k0 + q - i] (sb holds b after four
// zero limbs and is zero past its end, for C + 3 more limbs), then slides
// the window down one limb.  The next step's a limb and window limb are read
// before this step's products, so their LDS latency overlaps the
// multiply-adds; sa[ihi + 1] is read but not used (ihi + 1 <= na <= the
// array's length - 1, as nb >= 1).
// (hi0:lo0) += a * b0 and (hi1:lo1) += a * b1 as 96-bit sums: each a
// 64-bit multiply-add with its carry out (to an SGPR pair) added into the
// high word.  The waits between each carry's write and its read are
// explicit (the hazard recognizer does not look inside asm): the second
// multiply-add and one wait state.
__device__ __forceinline__ void mac96x2(uint64_t& lo0, uint32_t& hi0, uint64_t& lo1, uint32_t& hi1, uint32_t a,
                                        uint32_t b0, uint32_t b1) {
  uint64_t c0, c1;
  asm("v_mad_u64_u32 %0, %4, %6, %7, %0\n\t"
      "v_mad_u64_u32 %1, %5, %6, %8, %1\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32_e64 %2, %4, 0, %2, %4\n\t"
      "v_addc_co_u32_e64 %3, %5, 0, %3, %5"
      : "+v"(lo0), "+v"(lo1), "+v"(hi0), "+v"(hi1), "=&s"(c0), "=&s"(c1)
      : "v"(a), "v"(b0), "v"(b1));
}

template <int C>
__device__ __forceinline__ void school_columns(const uint32_t* sa, const uint32_t* sb, int k0, int ilo,
                                               int ihi, uint64_t (&lo)[C], uint32_t (&hi)[C]) {
  uint32_t w[C];
#pragma unroll
  for (int q = 0; q < C; ++q) w[q] = sb[4 + k0 + q - ilo];
  uint32_t an = sa[ilo];
#pragma unroll 2
  for (int i = ilo; i <= ihi; ++i) {
    const uint32_t ai = an;
    an = sa[i + 1];
    const uint32_t bn = sb[3 + k0 - i];  // b[k0 - (i + 1)]
#pragma unroll
    for (int q = 0; q < C; q += 2) mac96x2(lo[q], hi[q], lo[q + 1], hi[q + 1], ai, w[q], w[q + 1]);
    static_assert(C % 2 == 0, "columns go in pairs");
#pragma unroll
    for (int q = C - 1; q > 0; --q) w[q] = w[q - 1];
    w[0] = bn;
  }
}

__global__ __launch_bounds__(256) void k_batch_school(const BatchDesc* descs) {
  // sb is stored 3 limbs to the right with zeros around it, so the window
  // b[k0 + q - i] (q = 0..3) is readable for every i of the loop below.
  __shared__ uint32_t sa[kSchoolMaxLimbs], sbp[kSchoolMaxLimbs + 8];
  __shared__ uint32_t sw1[kSchoolMaxLimbs], sw2[kSchoolMaxLimbs];
  __shared__ CarryFn wtot[4];
  const BatchDesc d = descs[blockIdx.x];
  const int tid = threadIdx.x;
  const int na = (int)d.na, nb = (int)d.nb, zL = na + nb;
  for (int i = tid; i < na; i += 256) sa[i] = d.a[i];
  for (int i = tid; i < kSchoolMaxLimbs + 8; i += 256) {
    const int j = i - 4;
    sbp[i] = (j >= 0 && j < nb) ? d.b[j] : 0u;
  }
  __syncthreads();
  // Columns k0 .. k0 + 3 as 96-bit sums (lo 64 bits + hi word).  Each step
  // of i reads one limb of a and slides the window of b by one limb.
  const int k0 = 4 * tid;
  uint64_t lo[4] = {0, 0, 0, 0};
  uint32_t hi[4] = {0, 0, 0, 0};
  int ilo = k0 - nb + 1;
  if (ilo < 0) ilo = 0;
  int ihi = k0 + 3 < na - 1 ? k0 + 3 : na - 1;
  if (k0 < zL && ilo <= ihi && nb > 0) {
    school_columns<4>(sa, sbp, k0, ilo, ihi, lo, hi);
  }
  uint32_t w0[4], w1[4], w2[4], D[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    w0[q] = (uint32_t)lo[q];
    w1[q] = (uint32_t)(lo[q] >> 32);
    w2[q] = hi[q];
  }
  uint64_t c, spill;
  block_carry<4>(w0, w1, w2, sw1, sw2, wtot, D, c, spill);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = k0 + q;
    if (k < zL) d.out[k] = D[q];
  }
}

// One Karatsuba level over the schoolbook column sums (the crossover class
// of the schoolbook and the NTT), zL = na + nb <= 1024, one product per
// workgroup of 256 threads.  With h = ceil(max(na, nb) / 2), a = a0 + a1 B^h
// and b = b0 + b1 B^h (B = 2^32; a1 or b1 may be empty):
//   a b = z0 + (z1 - z0 - z2) B^h + z2 B^{2h},
//   z0 = a0 b0,  z2 = a1 b1,  z1 = (a0 + a1)(b0 + b1),
// three products of about h x h limbs, 3/4 of the schoolbook's
// multiply-adds.  The digit sums a0 + a1, b0 + b1 are resolved to 32-bit
// limbs first (one carry scan each), so the three products run on the
// schoolbook's 32 x 32-bit column sums; the columns combine column by
// column in 128-bit arithmetic.  A combined column may be negative, so
// every column k < zL carries the offset K B - K (K B in column 0), which
// adds K B^zL to the total: above the product's limbs, so the carry
// resolution of k_batch_school gives them unchanged.
constexpr int kKaraMaxH = kSchoolMaxLimbs / 2;  // h <= 512
constexpr uint64_t kKaraK = 1ull << 43;         // |negative part| < 2 h 2^64 <= 2^74 < K B - K

__global__ __launch_bounds__(256) void k_batch_kara(const BatchDesc* descs) {
  // a-side operands with one limb of slack (school_columns reads sa[ihi + 1]);
  // b-side operands after four zeros and zero past their end
  __shared__ uint32_t sa0[kKaraMaxH + 2], sa1[kKaraMaxH + 2], ssa[kKaraMaxH + 2];
  __shared__ uint32_t sb0p[kKaraMaxH + 12], sb1p[kKaraMaxH + 12], ssbp[kKaraMaxH + 12];
  __shared__ uint64_t clo[3][kSchoolMaxLimbs];  // column sums of z0, z2, z1
  __shared__ uint32_t chi[3][kSchoolMaxLimbs];
  __shared__ uint32_t sw1[kSchoolMaxLimbs], sw2[kSchoolMaxLimbs];
  __shared__ CarryFn wtot[4];
  const BatchDesc d = descs[blockIdx.x];
  const int tid = threadIdx.x;
  const int na = (int)d.na, nb = (int)d.nb, zL = na + nb;
  const int h = ((na > nb ? na : nb) + 1) >> 1;
  const int la0 = na < h ? na : h, la1 = na - la0;
  const int lb0 = nb < h ? nb : h, lb1 = nb - lb0;
  for (int i = tid; i < kKaraMaxH + 2; i += 256) {
    sa0[i] = i < la0 ? d.a[i] : 0u;
    sa1[i] = i < la1 ? d.a[h + i] : 0u;
    ssa[i] = 0u;
  }
  for (int i = tid; i < kKaraMaxH + 12; i += 256) {
    const int j = i - 4;
    sb0p[i] = (j >= 0 && j < lb0) ? d.b[j] : 0u;
    sb1p[i] = (j >= 0 && j < lb1) ? d.b[h + j] : 0u;
    ssbp[i] = 0u;
  }
  __syncthreads();
  // digit sums (h limbs and a carry): thread t adds limbs 2t, 2t + 1; the
  // carries are one scan of carry functions over the workgroup
  auto digit_sum = [&](const uint32_t* x, int xo, const uint32_t* y, int yo, uint32_t* out, int oo) {
    uint64_t sv[2];
    CarryFn f;
    f.base = 0;
    f.T = 0;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int i = 2 * tid + q;
      sv[q] = i < h ? (uint64_t)x[xo + i] + y[yo + i] : 0;
      CarryFn g;
      g.base = sv[q] >> 32;
      g.T = (1ull << 32) - (sv[q] & 0xffffffffull);
      f = cf_compose(f, g);
    }
    const CarryFn pre = cf_block_exclusive(f, wtot);
    uint64_t c = cf_eval0(pre);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int i = 2 * tid + q;
      const uint64_t t = (sv[q] & 0xffffffffull) + c;
      if (i < h) out[oo + i] = (uint32_t)t;
      c = (sv[q] >> 32) + (t >> 32);
      if (i == h - 1) out[oo + h] = (uint32_t)c;  // the carry out of the top limb
    }
  };
  digit_sum(sa0, 0, sa1, 0, ssa, 0);
  digit_sum(sb0p, 4, sb1p, 4, ssbp, 4);
  __syncthreads();
  const int lsa = la1 ? h + 1 : la0, lsb = lb1 ? h + 1 : lb0;
  // column sums of the three products, columns 4t .. 4t + 3 each
  const int k0 = 4 * tid;
#pragma unroll 1
  for (int p = 0; p < 3; ++p) {
    const uint32_t* x = p == 0 ? sa0 : (p == 1 ? sa1 : ssa);
    const uint32_t* yp = p == 0 ? sb0p : (p == 1 ? sb1p : ssbp);
    const int lx = p == 0 ? la0 : (p == 1 ? la1 : lsa);
    const int ly = p == 0 ? lb0 : (p == 1 ? lb1 : lsb);
    uint64_t lo[4] = {0, 0, 0, 0};
    uint32_t hi[4] = {0, 0, 0, 0};
    int ilo = k0 - ly + 1;
    if (ilo < 0) ilo = 0;
    const int ihi = k0 + 3 < lx - 1 ? k0 + 3 : lx - 1;
    if (lx > 0 && ly > 0 && k0 < lx + ly - 1 && ilo <= ihi) school_columns<4>(x, yp, k0, ilo, ihi, lo, hi);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      clo[p][k0 + q] = lo[q];
      chi[p][k0 + q] = hi[q];
    }
  }
  __syncthreads();
  auto col = [&](int p, int k) -> unsigned __int128 {
    return (k >= 0 && k < kSchoolMaxLimbs) ? ((unsigned __int128)chi[p][k] << 64) | clo[p][k] : 0;
  };
  uint32_t w0[4], w1[4], w2[4], D[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = k0 + q;
    unsigned __int128 acc = 0;
    if (k < zL) {
      acc = ((unsigned __int128)kKaraK << 32) - (k ? kKaraK : 0);
      acc += col(0, k) + col(1, k - 2 * h);
      acc += col(2, k - h);
      acc -= col(0, k - h) + col(1, k - h);
    }
    w0[q] = (uint32_t)acc;
    w1[q] = (uint32_t)(acc >> 32);
    w2[q] = (uint32_t)(acc >> 64);
  }
  uint64_t c, spill;
  block_carry<4>(w0, w1, w2, sw1, sw2, wtot, D, c, spill);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = k0 + q;
    if (k < zL) d.out[k] = D[q];
  }
}

// Schoolbook products of zL <= C G limbs with G threads per product (G = 16
// or 64, a part of one wave), C adjacent columns per thread (4 or 8) and
// 256 / G products per workgroup, for the many small products of the lower
// tree levels.  The column sums are those of k_batch_school with a window of
// C limbs; the carries are resolved inside the group with shuffles: each
// thread composes the carry functions of its C limbs, the group scans them
// (width-G shuffles), and each thread adds its carry-in.
__device__ __forceinline__ LimbFn lf_shfl_up_w(LimbFn f, int d, int width) {
  LimbFn r;
  r.base = __shfl_up(f.base, d, width);
  r.tm = __shfl_up(f.tm, d, width);
  return r;
}

template <int G, int C>
__global__ __launch_bounds__(256) void k_batch_school_g(const BatchDesc* descs, uint32_t count) {
  constexpr int PPB = 256 / G;  // products per workgroup
  constexpr int MAXL = C * G;   // zL bound
  constexpr int BPAD = MAXL + C + 4;  // b after four zeros, zeros to the window's reach
  __shared__ uint32_t sa[PPB][MAXL], sbp[PPB][BPAD];
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  const uint32_t z = blockIdx.x * PPB + grp;
  BatchDesc d;
  d.a = nullptr;
  d.b = nullptr;
  d.out = nullptr;
  d.na = 0;
  d.nb = 0;
  if (z < count) d = descs[z];
  const int na = (int)d.na, nb = (int)d.nb, zL = na + nb;
  for (int i = lane; i < na; i += G) sa[grp][i] = d.a[i];
  for (int i = lane; i < BPAD; i += G) {
    const int j = i - 4;
    sbp[grp][i] = (j >= 0 && j < nb) ? d.b[j] : 0u;
  }
  __syncthreads();
  const int k0 = C * lane;
  uint64_t lo[C];
  uint32_t hi[C];
#pragma unroll
  for (int q = 0; q < C; ++q) {
    lo[q] = 0;
    hi[q] = 0;
  }
  int ilo = k0 - nb + 1;
  if (ilo < 0) ilo = 0;
  const int ihi = k0 + C - 1 < na - 1 ? k0 + C - 1 : na - 1;
  if (k0 < zL && ilo <= ihi && nb > 0) {
    school_columns<C>(sa[grp], sbp[grp], k0, ilo, ihi, lo, hi);
  }
  // limb k receives w0_k + w1_{k-1} + w2_{k-2}; the previous thread's top
  // words come by shuffle (zero for the group's first thread)
  uint32_t x0[C], x1[C], x2[C];
#pragma unroll
  for (int q = 0; q < C; ++q) {
    x0[q] = (uint32_t)lo[q];
    x1[q] = (uint32_t)(lo[q] >> 32);
    x2[q] = hi[q];
  }
  uint32_t p1 = __shfl_up(x1[C - 1], 1, G), p22 = __shfl_up(x2[C - 2], 1, G),
           p23 = __shfl_up(x2[C - 1], 1, G);
  if (lane == 0) p1 = p22 = p23 = 0;
  uint64_t s[C];
  s[0] = (uint64_t)x0[0] + p1 + p22;
  s[1] = (uint64_t)x0[1] + x1[0] + p23;
#pragma unroll
  for (int q = 2; q < C; ++q) s[q] = (uint64_t)x0[q] + x1[q - 1] + x2[q - 2];
  // s < 3 * 2^32: carries of at most 3, 32-bit limb functions
  LimbFn f = lf_of_sum(s[0]);
#pragma unroll
  for (int q = 1; q < C; ++q) f = lf_compose(f, lf_of_sum(s[q]));
  LimbFn incl = f;
#pragma unroll
  for (int dd = 1; dd < G; dd <<= 1) {
    const LimbFn o = lf_shfl_up_w(incl, dd, G);
    if (lane >= dd) incl = lf_compose(o, incl);
  }
  LimbFn excl = lf_shfl_up_w(incl, 1, G);
  if (lane == 0) excl = lf_identity();
  uint64_t c = lf_eval0(excl);
#pragma unroll
  for (int q = 0; q < C; ++q) {
    const uint64_t t = (s[q] & 0xffffffffull) + c;
    if (k0 + q < zL) d.out[k0 + q] = (uint32_t)t;
    c = (s[q] >> 32) + (t >> 32);
  }
}

// Two lines in LDS, one per prime, each thread working in one of them: the
// first L/8 threads on prime 0's line, the next L/8 on prime 1's (whole
// waves, so a wave's prime is uniform).  Butterfly bf = t + b NTH of a stage
// is butterfly (t mod L/8) + b L/8 of line t / (L/8): within a line, the
// thread-to-butterfly map of a one-line row tile.
template <int LOGL>
struct GeoPrimes {
  static constexpr int L = 1 << LOGL;
  static constexpr int LINES = 2;
  static constexpr int E = 2 * L;
  static constexpr int NTH = E / kEPT;  // L / 4
  static constexpr int LP = L + L / 8;
  static constexpr int LDS = 2 * LP;
  __device__ static __forceinline__ int idx(int n, int lam) { return lam * LP + n + (n >> 3); }
  __device__ static __forceinline__ void split(int bf, int logLR, int& lam, int& j) {
    const int t = bf & (NTH - 1), b = bf >> (LOGL - 2);
    lam = t >> (LOGL - 3);
    j = (t & ((L >> 3) - 1)) + (b << (LOGL - 3));
  }
};

// One product per workgroup (N = 2^logN, 2^10 <= N <= 2^12); all products
// of one launch share logN, prm points at the constants for this N.  Both
// primes at once: N / 4 threads, the first half on prime 0's transforms and
// the second half on prime 1's (whole waves, so each wave holds one prime's
// points and its twiddle table).  The operands are read once, in
// first-stage order, into registers.  The forward DFTs of a and b run from
// registers to registers (one LDS line per prime for the middle stages), the
// pointwise product with N^-1 stays in registers, one pass through LDS puts
// it in first-stage order for the inverse DFT.  The prime-1 results go
// through LDS to the prime-0 threads for the CRT, whose words go through
// LDS into limb order for the block carry over all threads.  LDS: two padded
// lines (or the CRT staging) and a stage-twiddle table per prime.  Capped at
// 128 VGPRs (four waves per SIMD): at 2^11 that fits two workgroups per CU.
template <int LOGN>
__global__ __launch_bounds__((2 << LOGN) / kEPT) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_ntt(
    const BatchDesc* descs, const Prime* prm, double2 crt) {
  using GE = GeoPrimes<LOGN>;
  constexpr int N = 1 << LOGN;
  constexpr int LOGR1 = LOGN % 3 == 1 ? 1 : (LOGN % 3 == 2 ? 2 : 3);
  constexpr int R1 = 1 << LOGR1;
  constexpr int NB1 = kEPT / R1;
  constexpr int NTH = GE::NTH;
  constexpr int LDSD = GE::LDS > (5 * N) / 2 ? GE::LDS : (5 * N) / 2;  // doubles
  __shared__ double lds[LDSD];
  __shared__ CarryFn wtot[NTH / 64];
  __shared__ double2 ltw[2][lds_tw_size<LOGN>()];
  const BatchDesc d = descs[blockIdx.x];
  const int tid = threadIdx.x;
  const int pw = __builtin_amdgcn_readfirstlane(tid >> (LOGN - 3));  // this wave's prime
  const int na = (int)d.na, nb = (int)d.nb, zL = na + nb;
  const Prime P0 = prm[0], P1 = prm[1];
  const Prime& P = pw ? P1 : P0;
  stage_twiddles<GE>(ltw[0], P0);
  stage_twiddles<GE>(ltw[1], P1);
  const double2* tw = ltw[pw];
  uint32_t ua[kEPT], ub[kEPT];
  {
    const LimbBuf la(d.a, (uint32_t)na), lb(d.b, (uint32_t)nb);  // zero past na, nb
#pragma unroll
    for (int b = 0; b < NB1; ++b)
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        int lam, n;
        stage_pos<GE, LOGR1>(b, r, lam, n);
        ua[b * R1 + r] = la[(uint32_t)n];
        ub[b * R1 + r] = lb[(uint32_t)n];
      }
  }
  double y[1][8];
  {
    double x[kEPT], yb[1][8];
#pragma unroll
    for (int q = 0; q < kEPT; ++q) x[q] = (double)ua[q];
    line_dft_regs<GE>(x, y, lds, tw, P, false);
#pragma unroll
    for (int q = 0; q < kEPT; ++q) x[q] = (double)ub[q];
    __syncthreads();
    line_dft_regs<GE>(x, yb, lds, tw, P, false);
    // pointwise (natural order: point t' + r N / 8) with N^-1, then the
    // inverse DFT's first-stage order through LDS
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const double t = mm_gen(y[0][r], yb[0][r], P.p, P.pinv);
      y[0][r] = mm_shoup(t, P.ninv.x, P.ninv.y, P.p);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int lam, n;
      stage_pos<GE, 3>(0, r, lam, n);
      lds[GE::idx(n, lam)] = y[0][r];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB1; ++b)
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        int lam, n;
        stage_pos<GE, LOGR1>(b, r, lam, n);
        x[b * R1 + r] = lds[GE::idx(n, lam)];
      }
    __syncthreads();
    line_dft_regs<GE>(x, y, lds, tw, P, true);
  }
  // CRT of coefficient n: prime 1's results by point, then the words
  __syncthreads();
  double* r1 = lds;                                     // N doubles
  uint32_t* sw = reinterpret_cast<uint32_t*>(lds + N);  // w0 [0, N), w1 [N, 2N), w2 [2N, 3N)
  if (pw == 1) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int lam, n;
      stage_pos<GE, 3>(0, r, lam, n);
      r1[n] = y[0][r];
    }
  }
  __syncthreads();
  if (pw == 0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int lam, n;
      stage_pos<GE, 3>(0, r, lam, n);
      uint32_t c0 = 0, c1 = 0, c2 = 0;
      if (n < zL) crt_words(y[0][r], r1[n], P0, P1, crt, c0, c1, c2);
      sw[n] = c0;
      sw[N + n] = c1;
      sw[2 * N + n] = c2;
    }
  }
  __syncthreads();
  constexpr int LPT = N / NTH;  // 4 limbs per thread
  uint32_t w0[LPT], w1[LPT], w2[LPT], D[LPT];
#pragma unroll
  for (int q = 0; q < LPT; ++q) {
    const int k = tid * LPT + q;
    w0[q] = sw[k];
    w1[q] = sw[N + k];
    w2[q] = sw[2 * N + k];
  }
  __syncthreads();  // block_carry's scratch reuses the LDS
  uint64_t c, spill;
  block_carry<LPT>(w0, w1, w2, sw, sw + N, wtot, D, c, spill);
#pragma unroll
  for (int q = 0; q < LPT; ++q) {
    const int k = tid * LPT + q;
    if (k < zL) d.out[k] = D[q];
  }
}

// CRT and carry of batched multi-pass products, chunk-parallel: block
// (chunk, z) resolves limbs [1024 chunk, 1024 chunk + 1024) of product z
// with carry-in 0 and records the chunk's carry function (the form of
// k_carry_local); k_batch_carry_scan scans each product's chunk functions
// with one wave and adds the carry-ins; above 64 chunks
// k_batch_carry_scan_wide (one workgroup) scans and k_batch_carry_fix adds.  r: inverse-transform
// outputs of product z at r + z * zstride, prime q at offset q * N.
// agg / cin: nch entries per product.  Wrapped products (top != null): a
// product with w = zL - 1 - N > 0 holds the cyclic convolution of length N,
// and its top coefficients t = k - N + w - 1 at top + z * zstride (prime 1
// at + M) are subtracted below w and used from N up (as coeff_residues).
__global__ __launch_bounds__(kCarryThreads) void k_batch_carry_local(
    const BatchDesc* descs, const double* r, uint64_t zstride, uint64_t N, const Prime* prm,
    double2 crt, CarryFn* agg, uint32_t nch, const double* top, uint64_t M) {
  __shared__ uint32_t sw1[kCarryBlock], sw2[kCarryBlock];
  __shared__ CarryFn wtot[kCarryThreads / 64];
  __shared__ uint32_t d01[2];
  const uint32_t z = blockIdx.y;
  const BatchDesc d = descs[z];
  const uint64_t zL = (uint64_t)d.na + d.nb;
  const uint64_t base = (uint64_t)blockIdx.x * kCarryBlock;
  if (base >= zL) return;  // whole block: the chunk lies past this product
  const double* rz = r + (uint64_t)z * zstride;
  const double* tz = top ? top + (uint64_t)z * zstride : nullptr;
  const uint64_t w = (top && zL > N + 1) ? zL - 1 - N : 0;  // wrapped coefficients
  const uint64_t kend = w ? zL - 1 : (zL < N ? zL : N);
  const Prime P0 = prm[0], P1 = prm[1];
  const int tid = threadIdx.x;
  uint32_t w0[kCarryLPT], w1[kCarryLPT], w2[kCarryLPT], D[kCarryLPT];
  // both residues of the thread's coefficients below N, loaded before any
  // branch so that all of them are in flight together (the index clamped,
  // the value of a clamped load unused)
  double m0[kCarryLPT], m1[kCarryLPT];
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    const uint64_t k = base + tid * kCarryLPT + q;
    const uint64_t kc = k < N ? k : N - 1;
    m0[q] = rz[kc];
    m1[q] = rz[N + kc];
  }
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    const uint64_t k = base + tid * kCarryLPT + q;
    w0[q] = w1[q] = w2[q] = 0;
    if (k < kend) {
      double r0, r1;
      if (k < N) {
        r0 = m0[q];
        r1 = m1[q];
        if (k < w) {
          r0 -= tz[k + w - 1];
          r1 -= tz[M + k + w - 1];
        }
      } else {
        r0 = tz[k - N + w - 1];
        r1 = tz[M + k - N + w - 1];
      }
      crt_words(r0, r1, P0, P1, crt, w0[q], w1[q], w2[q]);
    }
  }
  uint64_t c, spill;
  block_carry<kCarryLPT>(w0, w1, w2, sw1, sw2, wtot, D, c, spill);
  bool ones = true;
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    const int li = tid * kCarryLPT + q;
    const uint64_t k = base + li;
    if (k < zL) d.out[k] = D[q];
    if (li >= 2 && D[q] != 0xffffffffu) ones = false;
    if (li < 2) d01[li] = D[q];
  }
  const int all_ones = __syncthreads_and(ones ? 1 : 0);
  if (tid == kCarryThreads - 1) {
    CarryFn g;
    g.base = spill + c;
    g.T = kInf;
    if (all_ones) {
      const uint64_t low = (uint64_t)d01[0] | ((uint64_t)d01[1] << 32);
      if (low) g.T = 0 - low;  // 2^64 - low
    }
    agg[(uint64_t)z * nch + blockIdx.x] = g;
  }
}

// Top descriptors of wrapped products: the w = zL - 1 - N highest limbs of
// each operand (an empty product when the coefficients fit N).
__global__ void k_wrap_descs(const BatchDesc* descs, uint64_t count, uint64_t N, BatchDesc* tops) {
  const uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= count) return;
  const BatchDesc d = descs[z];
  const uint64_t zL = (uint64_t)d.na + d.nb;
  BatchDesc t;
  t.out = nullptr;
  if (zL > N + 1) {
    const uint32_t w = (uint32_t)(zL - 1 - N);
    t.a = d.a + (d.na - w);
    t.b = d.b + (d.nb - w);
    t.na = w;
    t.nb = w;
  } else {
    t.a = d.a;
    t.b = d.b;
    t.na = 0;
    t.nb = 0;
  }
  tops[z] = t;
}

// One wave per product: exclusive scan of its chunk functions, 64 at a time,
// each chunk's carry-in rippled into its limbs (no k_batch_carry_fix).
__global__ __launch_bounds__(64) void k_batch_carry_scan(const BatchDesc* descs, const CarryFn* agg,
                                                         uint64_t* cin, uint32_t nch) {
  const uint32_t z = blockIdx.x;
  const BatchDesc d = descs[z];
  const uint64_t zL = (uint64_t)d.na + d.nb;
  const uint32_t nc = (uint32_t)((zL + kCarryBlock - 1) / kCarryBlock);
  const int lane = threadIdx.x;
  CarryFn run;
  run.base = 0;
  run.T = 0;
  for (uint32_t c0 = 0; c0 < nc; c0 += 64) {
    const uint32_t idx = c0 + lane;
    CarryFn f;
    f.base = 0;
    f.T = 0;
    if (idx < nc) f = agg[(uint64_t)z * nch + idx];
    CarryFn incl = f;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
      const CarryFn o = cf_shfl_up(incl, s);
      if (lane >= s) incl = cf_compose(o, incl);
    }
    CarryFn excl = cf_shfl_up(incl, 1);
    if (lane == 0) excl.T = 0;
    if (idx < nc) {
      // the chunk's carry-in, rippled into its limbs here (as k_batch_carry_fix)
      uint64_t c = cf_eval0(cf_compose(run, excl));
      uint64_t k = (uint64_t)idx * kCarryBlock;
      const uint64_t end = k + kCarryBlock < zL ? k + kCarryBlock : zL;
      while (c != 0 && k < end) {
        const uint64_t v = (uint64_t)d.out[k] + (c & 0xffffffffull);
        d.out[k] = (uint32_t)v;
        c = (c >> 32) + (v >> 32);
        ++k;
      }
    }
    CarryFn tot;
    tot.base = __shfl(incl.base, 63, 64);
    tot.T = __shfl(incl.T, 63, 64);
    run = cf_compose(run, tot);
  }
}

// The same scan with one workgroup per product, for products of more than
// 64 chunks (cf_scan_runs).
__global__ __launch_bounds__(256) void k_batch_carry_scan_wide(const BatchDesc* descs, const CarryFn* agg,
                                                               uint64_t* cin, uint32_t nch) {
  __shared__ CarryFn wtot[4], tot;
  const uint32_t z = blockIdx.x;
  const BatchDesc d = descs[z];
  const uint64_t zL = (uint64_t)d.na + d.nb;
  const uint32_t nc = (uint32_t)((zL + kCarryBlock - 1) / kCarryBlock);
  (void)cf_scan_runs(agg + (uint64_t)z * nch, nc, cin + (uint64_t)z * nch, wtot, &tot);
}

// One thread per (product, chunk): ripple the chunk's carry-in.
__global__ __launch_bounds__(256) void k_batch_carry_fix(const BatchDesc* descs, const uint64_t* cin,
                                                         uint32_t nch, uint64_t count) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t z = t / nch, ch = t % nch;
  if (z >= count) return;
  const BatchDesc d = descs[z];
  const uint64_t zL = (uint64_t)d.na + d.nb;
  uint64_t k = ch * kCarryBlock;
  if (k >= zL) return;
  uint64_t c = cin[z * nch + ch];
  uint64_t end = k + kCarryBlock;
  if (end > zL) end = zL;
  while (c != 0 && k < end) {
    const uint64_t v = (uint64_t)d.out[k] + (c & 0xffffffffull);
    d.out[k] = (uint32_t)v;
    c = (c >> 32) + (v >> 32);
    ++k;
  }
}

}  // namespace pcg
