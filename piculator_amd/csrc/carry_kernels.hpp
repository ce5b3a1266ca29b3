// CRT reconstruction and exact carry propagation (device side).
//
// After the inverse transform each point i holds the convolution coefficient
// c_i modulo both primes; c_i < p0 * p1 (~2^94), so CRT gives c_i exactly as
// a 96-bit integer (w0, w1, w2) of 32-bit words.  The product is
//     sum_i c_i 2^{32 i},
// so limb k receives s_k = w0_k + w1_{k-1} + w2_{k-2} < 3 * 2^32 plus the
// carry from limb k-1.  With carries below 2^32, limb k maps a carry-in c to
//     carry_out = hi(s_k) + [c >= 2^32 - lo(s_k)]
// i.e. a function f(c) = base + [c >= T].  Such functions are closed under
// composition, which makes carry propagation an associative scan:
//     (o1,T1) then (o2,T2) = (o2+1, inf)  if o1 >= T2
//                            (o2,  T1)   if o1 + 1 == T2
//                            (o2,  inf)  otherwise.
// The same form describes a whole block of limbs (T finite only when all but
// its two lowest limbs are 0xFFFFFFFF), so blocks are resolved locally with
// carry-in 0, block functions are scanned, and each block adds its carry-in.
#pragma once
#include "limb_buf.hpp"
#include "modfp64.hpp"

#pragma clang fp contract(off)

namespace pcg {

constexpr int kCarryThreads = 256;
constexpr int kCarryLPT = 4;                           // limbs per thread
constexpr int kCarryBlock = kCarryThreads * kCarryLPT;  // limbs per block
constexpr uint64_t kInf = ~0ull;

struct CarryFn {
  uint64_t base, T;  // T == 0 marks the identity function
};

__device__ __forceinline__ CarryFn cf_compose(CarryFn f1, CarryFn f2) {
  if (f1.T == 0) return f2;
  if (f2.T == 0) return f1;
  CarryFn r;
  if (f1.base >= f2.T) {
    r.base = f2.base + 1;
    r.T = kInf;
  } else if (f1.base + 1 >= f2.T) {
    r.base = f2.base;
    r.T = f1.T;
  } else {
    r.base = f2.base;
    r.T = kInf;
  }
  return r;
}

// The same functions for carries between single limbs, in 32-bit fields:
// a limb's carry-in c is at most 4, so f(c) = base + [c >= T] needs T only
// up to 2^32 (T = 2^32 - lo(s) >= 1), and T = 2^32 acts as never.  Stored
// as tm = T - 1 (0xFFFFFFFF: never); base == kLimbId marks the identity.
constexpr uint32_t kLimbId = 0xFFFFFFFFu;
constexpr uint32_t kLimbNever = 0xFFFFFFFFu;
struct LimbFn {
  uint32_t base, tm;
};

// The function of a limb whose sum (without carry-in) is s < 4 * 2^32.
__device__ __forceinline__ LimbFn lf_of_sum(uint64_t s) {
  LimbFn f;
  f.base = (uint32_t)(s >> 32);
  f.tm = ~(uint32_t)s;  // 2^32 - lo(s) - 1
  return f;
}

__device__ __forceinline__ LimbFn lf_identity() {
  LimbFn f;
  f.base = kLimbId;
  f.tm = kLimbNever;
  return f;
}

// Selects only, field by field (a wave's lanes take different cases; a
// select of whole structs would go through scratch memory).  When
// f1.base > tm2 the test f1.base == tm2 fails, so tm becomes never.
__device__ __forceinline__ LimbFn lf_compose(LimbFn f1, LimbFn f2) {
  const bool id1 = f1.base == kLimbId, id2 = f2.base == kLimbId;
  uint32_t base = f2.base + (f1.base > f2.tm ? 1u : 0u);
  uint32_t tm = f1.base == f2.tm ? f1.tm : kLimbNever;
  base = id2 ? f1.base : base;
  tm = id2 ? f1.tm : tm;
  LimbFn r;
  r.base = id1 ? f2.base : base;
  r.tm = id1 ? f2.tm : tm;
  return r;
}

__device__ __forceinline__ uint32_t lf_eval0(LimbFn f) { return f.base == kLimbId ? 0u : f.base; }

// The 64-bit form of a limb function, for block functions scanned across
// blocks whose carries stay below 2^32 - 1 (T = 2^32 acts as never there).
__device__ __forceinline__ CarryFn cf_of_limb(LimbFn f) {
  const bool id = f.base == kLimbId;
  CarryFn r;
  r.base = id ? 0 : f.base;
  r.T = id ? 0 : (f.tm == kLimbNever ? kInf : (uint64_t)f.tm + 1);
  return r;
}

__device__ __forceinline__ LimbFn lf_shfl_up(LimbFn f, int d) {
  LimbFn r;
  r.base = __shfl_up(f.base, d, 64);
  r.tm = __shfl_up(f.tm, d, 64);
  return r;
}

// Exclusive scan of one LimbFn per thread across the workgroup (<= 1024
// threads); `wave_tot` is LDS scratch of (threads / 64) entries.
__device__ __forceinline__ LimbFn lf_block_exclusive(LimbFn f, LimbFn* wave_tot) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  LimbFn incl = f;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const LimbFn o = lf_shfl_up(incl, d);
    if (lane >= d) incl = lf_compose(o, incl);
  }
  LimbFn excl = lf_shfl_up(incl, 1);
  if (lane == 0) excl = lf_identity();
  if (lane == 63) wave_tot[wv] = incl;
  __syncthreads();
  LimbFn pre = lf_identity();
  for (int i = 0; i < wv; ++i) pre = lf_compose(pre, wave_tot[i]);
  __syncthreads();
  return lf_compose(pre, excl);
}

__device__ __forceinline__ uint64_t cf_eval0(CarryFn f) { return f.T == 0 ? 0 : f.base; }

__device__ __forceinline__ CarryFn cf_shfl_up(CarryFn f, int d) {
  CarryFn r;
  r.base = __shfl_up(f.base, d, 64);
  r.T = __shfl_up(f.T, d, 64);
  return r;
}

// Exclusive scan of one CarryFn per thread across the workgroup (<= 1024
// threads); `wave_tot` is LDS scratch of (threads / 64) entries.
__device__ __forceinline__ CarryFn cf_block_exclusive(CarryFn f, CarryFn* wave_tot) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  CarryFn incl = f;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    CarryFn o = cf_shfl_up(incl, d);
    if (lane >= d) incl = cf_compose(o, incl);
  }
  CarryFn excl = cf_shfl_up(incl, 1);
  if (lane == 0) excl.T = 0;
  if (lane == 63) wave_tot[wv] = incl;
  __syncthreads();
  CarryFn pre;
  pre.base = 0;
  pre.T = 0;
  for (int i = 0; i < wv; ++i) pre = cf_compose(pre, wave_tot[i]);
  (void)nw;
  __syncthreads();
  return cf_compose(pre, excl);
}

// Exclusive scan of the n chunk functions a[0..n) by one whole workgroup
// (any n): thread t composes the run [t K, t K + K), K = ceil(n / threads),
// the runs are scanned across the workgroup, and each thread walks its run
// again from its prefix, writing out[i] = the carry into chunk i.  Loads go
// four at a time so a long run is not one memory latency per chunk.
// Returns the composition of all n (every thread).  wave_tot: LDS of
// threads / 64 entries; tot: one LDS entry.
__device__ __forceinline__ CarryFn cf_scan_runs(const CarryFn* a, uint32_t n, uint64_t* out,
                                                CarryFn* wave_tot, CarryFn* tot) {
  const uint32_t K = (n + blockDim.x - 1) / blockDim.x;
  const uint32_t c0 = min(n, (uint32_t)threadIdx.x * K), c1 = min(n, c0 + K);
  CarryFn f;
  f.base = 0;
  f.T = 0;
  uint32_t c = c0;
  for (; c + 4 <= c1; c += 4) {
    const CarryFn g0 = a[c], g1 = a[c + 1], g2 = a[c + 2], g3 = a[c + 3];
    f = cf_compose(f, cf_compose(cf_compose(g0, g1), cf_compose(g2, g3)));
  }
  for (; c < c1; ++c) f = cf_compose(f, a[c]);
  CarryFn run = cf_block_exclusive(f, wave_tot);
  for (c = c0; c + 4 <= c1; c += 4) {
    const CarryFn g0 = a[c], g1 = a[c + 1], g2 = a[c + 2], g3 = a[c + 3];
    out[c] = cf_eval0(run);
    run = cf_compose(run, g0);
    out[c + 1] = cf_eval0(run);
    run = cf_compose(run, g1);
    out[c + 2] = cf_eval0(run);
    run = cf_compose(run, g2);
    out[c + 3] = cf_eval0(run);
    run = cf_compose(run, g3);
  }
  for (; c < c1; ++c) {
    out[c] = cf_eval0(run);
    run = cf_compose(run, a[c]);
  }
  // the last thread's run ends at n (an empty run starts there)
  if (threadIdx.x == blockDim.x - 1) *tot = run;
  __syncthreads();
  return *tot;
}

struct CarryArgs {
  const Prime* prm;  // the two primes (p, pinv)
  const double* r;   // [2][N] inverse-transform outputs
  uint64_t N;
  uint64_t zL;       // output limbs
  uint32_t* out;     // zL limbs
  CarryFn* agg;      // per block
  uint64_t* cin;     // per block carry-in
  uint64_t nblk;
  double2 crt;       // p0^{-1} mod p1 as (w, w/p1)
  uint32_t* status;  // nonzero on overflow (null: no check)
  uint64_t* carry_out;  // carry out of the top block (null: not stored)
  CarryFn* tot;         // scan scratch: one entry per 1024 blocks
  // FFT engine: unweighted inverse-transform output and round-off maximum
  const double2* z;
  unsigned long long* maxdist;  // bits of a non-negative double
  // Wrapped product (top != null): r holds the cyclic convolution of length
  // N < zL - 1, so coefficient k < top_w also holds coefficient k + N; top
  // holds the exact coefficients t = k - N + top_w - 1 of those (prime 1 at
  // top + top_stride), which are subtracted below N and used from N up.
  const double* top;
  uint64_t top_stride;
  uint64_t top_w;
};

// Residues of coefficient k of a product in a.r (two primes, N each) given
// r0, r1 = a.r[k], a.r[N + k] when k < N (loaded by the caller): a wrapped
// product (a.top) subtracts its short product's coefficient k + w - 1 below
// w and takes coefficient k - N + w - 1 of it from N up.
__device__ __forceinline__ void coeff_residues_top(const CarryArgs& a, uint64_t k, double& r0, double& r1) {
  if (k < a.N) {
    if (a.top && k < a.top_w) {
      const uint64_t t = k + a.top_w - 1;
      r0 -= a.top[t];
      r1 -= a.top[a.top_stride + t];
    }
  } else {
    const uint64_t t = k - a.N + a.top_w - 1;
    r0 = a.top[t];
    r1 = a.top[a.top_stride + t];
  }
}

// Limb k of the FFT engine: digits 2k, 2k+1 are real parts of points 2k, 2k+1
// (below N) or imaginary parts of points 2k-N, 2k+1-N.  The rounded digit
// coefficients give s_k = c0 + 2^16 c1 (signed); the constant
// K 2^32 - K (K 2^32 at k = 0) is added so every limb value is non-negative,
// which adds K 2^(32 zL) to the total, above the product's limbs.
__device__ __forceinline__ void fft_words(const double2* z, uint64_t N, uint64_t k, uint32_t& w0,
                                          uint32_t& w1, uint32_t& w2, double& dist) {
  const uint64_t d0 = 2 * k;
  double x0, x1;
  if (d0 < N) {
    x0 = z[d0].x;
    x1 = z[d0 + 1].x;
  } else {
    x0 = z[d0 - N].y;
    x1 = z[d0 + 1 - N].y;
  }
  double r0 = __builtin_rint(x0), r1 = __builtin_rint(x1);
  dist = fmax(fabs(x0 - r0), fabs(x1 - r1));
  const double lim = 1152921504606846976.0;  // 2^60: garbage stays finite
  r0 = fmin(fmax(r0, -lim), lim);
  r1 = fmin(fmax(r1, -lim), lim);
  const __int128 K = (__int128)1 << 40;
  __int128 s = (__int128)(int64_t)r0 + ((__int128)(int64_t)r1 << 16) + (K << 32);
  if (k > 0) s -= K;
  const unsigned __int128 t = (unsigned __int128)s;
  w0 = (uint32_t)t;
  w1 = (uint32_t)(t >> 32);
  w2 = (uint32_t)(t >> 64);
}

// An integer 0 <= x < 2^52 held in a double, as an integer: x + 2^52 is
// exact and its mantissa field is x.
__device__ __forceinline__ uint64_t small_to_u64(double x) {
  return (uint64_t)__double_as_longlong(x + 4503599627370496.0) & ((1ull << 52) - 1);
}

// c = r0 + p0 t with t = (r1 - r0) p0^-1 mod p1: only r0 (the base) and t
// need their canonical residues; r1 enters through the difference, which the
// Shoup product reduces whatever representative r1 is.
__device__ __forceinline__ void crt_words(double r0, double r1, const Prime& P0, const Prime& P1,
                                          double2 crt, uint32_t& w0, uint32_t& w1, uint32_t& w2) {
  r0 = canon(r0, P0.p, P0.pinv);
  double t = mm_shoup(r1 - r0, crt.x, crt.y, P1.p);
  t = canon(t, P1.p, P1.pinv);
  const uint64_t tu = small_to_u64(t), r0u = small_to_u64(r0), p0u = small_to_u64(P0.p);
  uint64_t lo = p0u * tu;
  uint64_t hi = __umul64hi(p0u, tu);
  lo += r0u;
  hi += lo < r0u;
  w0 = (uint32_t)lo;
  w1 = (uint32_t)(lo >> 32);
  w2 = (uint32_t)hi;
}

// Resolves the carries of LPT * blockDim.x consecutive limbs held LPT per
// thread (thread t holds limbs [LPT t, LPT t + LPT)) given their sums
// s < 4 * 2^32 (limb value plus everything added into that limb position)
// and a carry-in of 0: D receives the limbs, cout (meaningful in the last
// thread) the carry out of the top limb.
template <int LPT>
__device__ __forceinline__ void block_carry_sums(const uint64_t* s, CarryFn* wtot, uint32_t* D,
                                                 uint64_t& cout) {
  // carries between these limbs are at most 4: 32-bit limb functions, the
  // scan's LDS scratch taken from wtot
  LimbFn f = lf_of_sum(s[0]);
#pragma unroll
  for (int q = 1; q < LPT; ++q) f = lf_compose(f, lf_of_sum(s[q]));
  const LimbFn pre = lf_block_exclusive(f, reinterpret_cast<LimbFn*>(wtot));
  uint64_t c = lf_eval0(pre);
#pragma unroll
  for (int q = 0; q < LPT; ++q) {
    const uint64_t t = (s[q] & 0xffffffffull) + c;
    D[q] = (uint32_t)t;
    c = (s[q] >> 32) + (t >> 32);
  }
  cout = c;
}

// Same from CRT words: limb k receives s_k = w0_k + w1_{k-1} + w2_{k-2}.
// Thread t's first two sums take its left neighbour's top words, by a
// shuffle inside the wave and, for lane 0, through LDS from the previous
// wave's lane 63 (sw1: blockDim.x / 64 words, sw2: twice that); the rest
// come from its own registers.  (Exchanging every word through LDS at a
// stride of LPT words per thread put a quarter of the LDS cycles of the
// CRT/carry kernels in bank conflicts.)  wtot is LDS scratch of
// blockDim.x / 64.  blockDim.x is a multiple of 64.  A carry-in cin
// (< 2^64) is folded into the first two limbs: their sums stay below
// 3.5 * 2^32, so carries stay small.  `spill` is the value of the words
// that spill past the top limb (the same in every thread).
template <int LPT>
__device__ __forceinline__ void block_carry(const uint32_t* w0, const uint32_t* w1,
                                            const uint32_t* w2, uint32_t* sw1, uint32_t* sw2,
                                            CarryFn* wtot, uint32_t* D, uint64_t& cout,
                                            uint64_t& spill, uint64_t cin = 0) {
  static_assert(LPT >= 2, "two words reach past a thread's first limb");
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int last = (blockDim.x >> 6) - 1;  // the last wave
  uint32_t p1 = __shfl_up(w1[LPT - 1], 1, 64);
  uint32_t p22 = __shfl_up(w2[LPT - 2], 1, 64);
  uint32_t p23 = __shfl_up(w2[LPT - 1], 1, 64);
  if (lane == 63) {
    sw1[wv] = w1[LPT - 1];
    sw2[2 * wv] = w2[LPT - 2];
    sw2[2 * wv + 1] = w2[LPT - 1];
  }
  __syncthreads();
  if (lane == 0) {
    if (wv > 0) {
      p1 = sw1[wv - 1];
      p22 = sw2[2 * wv - 2];
      p23 = sw2[2 * wv - 1];
    } else {
      p1 = (uint32_t)cin;  // the carry-in takes the first thread's neighbour's place
      p22 = 0;
      p23 = (uint32_t)(cin >> 32);
    }
  }
  spill = (uint64_t)sw1[last] + (uint64_t)sw2[2 * last] + ((uint64_t)sw2[2 * last + 1] << 32);
  uint64_t s[LPT];
  s[0] = (uint64_t)w0[0] + p1 + p22;
  s[1] = (uint64_t)w0[1] + w1[0] + p23;
#pragma unroll
  for (int q = 2; q < LPT; ++q) s[q] = (uint64_t)w0[q] + w1[q - 1] + w2[q - 2];
  block_carry_sums<LPT>(s, wtot, D, cout);
}

template <bool FFT>
__global__ __launch_bounds__(kCarryThreads) void k_carry_local(CarryArgs a) {
  __shared__ uint32_t sw1[kCarryBlock], sw2[kCarryBlock];
  __shared__ CarryFn wtot[kCarryThreads / 64];
  __shared__ uint32_t d01[2];
  const int tid = threadIdx.x;
  const uint64_t base = (uint64_t)blockIdx.x * kCarryBlock;
  uint32_t w0[kCarryLPT], w1[kCarryLPT], w2[kCarryLPT], D[kCarryLPT];
  if (FFT) {
    double dmax = 0.0;
#pragma unroll
    for (int q = 0; q < kCarryLPT; ++q) {
      const uint64_t k = base + tid * kCarryLPT + q;
      w0[q] = w1[q] = w2[q] = 0;
      if (k < a.zL) {
        double dist;
        fft_words(a.z, a.N, k, w0[q], w1[q], w2[q], dist);
        dmax = fmax(dmax, dist);
      }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, d, 64));
    // one contended address for all blocks: only a wave whose maximum
    // exceeds the current one issues the atomic (a plain load first), so
    // the atomics stay few instead of one per wave
    if ((tid & 63) == 0) {
      const unsigned long long v = (unsigned long long)__double_as_longlong(dmax);
      if (v > __atomic_load_n(a.maxdist, __ATOMIC_RELAXED)) atomicMax(a.maxdist, v);
    }
  } else {
    const Prime P0 = a.prm[0], P1 = a.prm[1];
    // coefficients 0 .. zL - 2 (the top limb only receives carries); from
    // N up they exist only in a wrapped product
    const uint64_t kend = a.top ? a.zL - 1 : (a.zL - 1 < a.N ? a.zL - 1 : a.N);
    // both residues of the thread's coefficients below N, loaded before any
    // branch so that all of them are in flight together (the index
    // clamped, the value of a clamped load unused)
    double m0[kCarryLPT], m1[kCarryLPT];
#pragma unroll
    for (int q = 0; q < kCarryLPT; ++q) {
      const uint64_t k = base + tid * kCarryLPT + q;
      const uint64_t kc = k < a.N ? k : a.N - 1;
      m0[q] = a.r[kc];
      m1[q] = a.r[a.N + kc];
    }
#pragma unroll
    for (int q = 0; q < kCarryLPT; ++q) {
      const uint64_t k = base + tid * kCarryLPT + q;
      w0[q] = w1[q] = w2[q] = 0;
      if (k < kend) {
        double r0 = m0[q], r1 = m1[q];
        coeff_residues_top(a, k, r0, r1);
        crt_words(r0, r1, P0, P1, a.crt, w0[q], w1[q], w2[q]);
      }
    }
  }
  uint64_t c, spill;
  block_carry<kCarryLPT>(w0, w1, w2, sw1, sw2, wtot, D, c, spill);
  bool ones = true;
#pragma unroll
  for (int q = 0; q < kCarryLPT; ++q) {
    const int li = tid * kCarryLPT + q;
    const uint64_t k = base + li;
    if (k < a.zL) a.out[k] = D[q];
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
    a.agg[blockIdx.x] = g;
  }
}

// Exclusive scan of the block functions agg[0 .. nblk) in three launches:
// k_scan_partial reduces each run of 1024 functions, k_scan_totals scans
// those totals (one workgroup), k_scan_apply forms every block's prefix.
// Every thread handles one function, so the loads are coalesced and none
// waits on another.
constexpr int kScanChunk = 1024;

__global__ __launch_bounds__(kScanChunk) void k_scan_partial(CarryArgs a) {
  __shared__ CarryFn wtot[kScanChunk / 64];
  const uint64_t i = (uint64_t)blockIdx.x * kScanChunk + threadIdx.x;
  CarryFn f;
  f.base = 0;
  f.T = 0;
  if (i < a.nblk) f = a.agg[i];
  const CarryFn ex = cf_block_exclusive(f, wtot);
  if (threadIdx.x == kScanChunk - 1) a.tot[blockIdx.x] = cf_compose(ex, f);
}

// One workgroup: tot[c] becomes the composition of all chunks before c;
// the total is the carry out of the top block.
__global__ __launch_bounds__(kScanChunk) void k_scan_totals(CarryArgs a, uint64_t ntot) {
  __shared__ CarryFn wtot[kScanChunk / 64];
  __shared__ CarryFn chunk_tot;
  CarryFn run;
  run.base = 0;
  run.T = 0;
  for (uint64_t c0 = 0; c0 < ntot; c0 += kScanChunk) {
    const uint64_t i = c0 + threadIdx.x;
    CarryFn f;
    f.base = 0;
    f.T = 0;
    if (i < ntot) f = a.tot[i];
    const CarryFn ex = cf_block_exclusive(f, wtot);
    if (threadIdx.x == kScanChunk - 1) chunk_tot = cf_compose(ex, f);
    if (i < ntot) a.tot[i] = cf_compose(run, ex);
    __syncthreads();
    run = cf_compose(run, chunk_tot);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // carry out of the top block must vanish for an exact product
    if (a.status && cf_eval0(run) != 0) *a.status = 1;
    if (a.carry_out) *a.carry_out = cf_eval0(run);
  }
}

__global__ __launch_bounds__(kScanChunk) void k_scan_apply(CarryArgs a) {
  __shared__ CarryFn wtot[kScanChunk / 64];
  const uint64_t i = (uint64_t)blockIdx.x * kScanChunk + threadIdx.x;
  CarryFn f;
  f.base = 0;
  f.T = 0;
  if (i < a.nblk) f = a.agg[i];
  const CarryFn ex = cf_block_exclusive(f, wtot);
  if (i < a.nblk) a.cin[i] = cf_eval0(cf_compose(a.tot[blockIdx.x], ex));
}

__global__ __launch_bounds__(256) void k_carry_fix(CarryArgs a) {
  const uint64_t blk = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (blk >= a.nblk) return;
  uint64_t c = a.cin[blk];
  if (c == 0) return;
  uint64_t k = blk * kCarryBlock;
  uint64_t end = k + kCarryBlock;
  if (end > a.zL) end = a.zL;
  // The first two limbs take c < 2^64; past them the carry is 0 or 1.
  for (int i = 0; i < 2 && c != 0 && k < end; ++i, ++k) {
    const uint64_t t = (uint64_t)a.out[k] + (c & 0xffffffffull);
    a.out[k] = (uint32_t)t;
    c = (c >> 32) + (t >> 32);
  }
  if (c == 1 && a.zL < (1ull << 29)) {
    // A carry of 1 turns limbs equal to 2^32 - 1 into 0 and stops at the
    // first other limb.  Such runs can fill whole blocks (a borrow through
    // the zero limbs of a short minuend), so 16 limbs per round, their
    // loads in flight together (range-checked: zero past zL), instead of
    // one dependent load and store per limb.
    const LimbBuf lb(a.out, (uint32_t)a.zL);
    while (c != 0 && k < end) {
      uint32_t v[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) v[j] = lb[(uint32_t)(k + j)];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        if (c != 0 && k + j < end) {
          const bool ones = v[j] == 0xffffffffu;
          a.out[k + j] = ones ? 0u : v[j] + 1u;
          if (!ones) c = 0;
        }
      }
      k += 16;
    }
    if (k > end) k = end;
  }
  while (c != 0 && k < end) {
    const uint64_t t = (uint64_t)a.out[k] + (c & 0xffffffffull);
    a.out[k] = (uint32_t)t;
    c = (c >> 32) + (t >> 32);
    ++k;
  }
  // A carry left at the block end is the block's overflow bit (0 or 1); the
  // scan already added it to the next block's carry-in.
  if (c > 1 && a.status) *a.status = 2;
}

}  // namespace pcg
