// picugpu: C-ABI big-integer multiply engine for MI355X (gfx950).
//
// Host side: tables, transform planning, workspace, launches and the C ABI
// declared in include/picugpu.h.  The engine replaces the convolution block
// of BigFloat::mul (reference BigFloat.cpp:377-441): the reference converts
// limbs to 16-bit digits and runs a 63-bit single-prime NTT (NTT.cpp) or an
// 8-bit-digit FP64 FFT (FFT.cpp); here each 32-bit limb is one point of a
// two-prime NTT computed in FP64 (see modfp64.hpp), followed by CRT and an
// exact carry scan.  The product is the same exact integer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/picugpu.h"
#include "batch_kernels.hpp"
#include "bsr_kernels.hpp"
#include "carry_kernels.hpp"
#include "df_kernels.hpp"
#include "fft_kernels.hpp"
#include "ntt_big.hpp"
#include "ntt_row.hpp"
#include "ntt_kernels.hpp"

using namespace pcg;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      return fail(PCG_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                              \
  } while (0)

// ---------------------------------------------------------------- host math
uint64_t mulmod_u64(uint64_t a, uint64_t b, uint64_t p) {
  return (uint64_t)(((unsigned __int128)a * b) % p);
}
uint64_t powmod_u64(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1;
  a %= p;
  while (e) {
    if (e & 1) r = mulmod_u64(r, a, p);
    a = mulmod_u64(a, a, p);
    e >>= 1;
  }
  return r;
}
double2 shoup_pair(uint64_t w, uint64_t p) { return make_double2((double)w, (double)w / (double)p); }

struct Pass {
  int logL, logS, logLines, logG, logB;
};

// Passes of one transform in a fixed array (at most 4 for 2^30 points), so
// planning a product never touches the heap.
struct PassList {
  static constexpr size_t kCap = 6;
  Pass v[kCap];
  size_t n = 0;
  void push_back(const Pass& p) { v[n++] = p; }
  size_t size() const { return n; }
  const Pass& operator[](size_t i) const { return v[i]; }
};

struct Plan {
  int logN;
  int logE;
  bool big = false;  // large-tile passes (k_ntt_col, k_ntt_row: 2^14 points per tile)
  PassList passes;
};

// NTT pass plan: strided passes take at most 2^8-point lines, so a
// workgroup's 4096 staged points cover >= 16 adjacent columns (128-byte
// runs, which stream at ~80% of a contiguous copy on MI355X; 64-byte runs
// reach ~65%); the last pass (contiguous rows) takes up to 2^12.  Where
// 2^9-point columns (8 adjacent, 64-byte runs) save a whole pass, the plan
// takes them: one sweep fewer outweighs the slower strided pattern.
constexpr int kMaxLogCol = 8;
constexpr int kMaxLogColWide = 9;

int plan_passes(int logN, int maxcol) {
  int m = 1;
  while (logN > maxcol * (m - 1) + kLogE) ++m;
  return m;
}

// Large-tile plan (k_ntt_col, k_ntt_row): a column pass of 2^(logN-14) points (8 or
// more adjacent columns per tile) and a row pass of 2^14: two sweeps; from
// 2^26 points two column passes of 2^6 .. 2^7 points and the row pass.  Used
// from 2^21 to 2^28 points, where it measured faster than the 4096-point
// passes (2^22: 0.25 vs 0.31 ms per product; 2^26: 4.46 vs 5.03 ms, 2^27:
// 8.75 vs 10.04 ms); 2^20 is faster on two 4096-point passes.
constexpr int kBigMinLogN = 21;
constexpr int kBigMaxLogN = 28;

Plan make_big_plan(int logN) {
  Plan pl;
  pl.logN = logN;
  pl.logE = kBigLogE;
  pl.big = true;
  // two passes up to 2^25 points; above, two column passes of 2^6 .. 2^7
  // points before the 2^14-point row pass
  const int cols = logN - kBigLogE;
  const int np = cols <= 11 ? 2 : 3;
  const int ls[3] = {np == 2 ? cols : cols / 2, np == 2 ? kBigLogE : cols - cols / 2, kBigLogE};
  int rem = logN;
  for (int i = 0; i < np; ++i) {
    Pass ps;
    ps.logL = ls[i];
    rem -= ps.logL;
    ps.logS = rem;
    ps.logLines = kBigLogE - ps.logL;
    ps.logG = ps.logS >= ps.logLines ? ps.logLines : ps.logS;
    ps.logB = ps.logLines - ps.logG;
    pl.passes.push_back(ps);
  }
  return pl;
}

// Plan of the products of a batched launch at 2^logN points: the large
// tiles above 2^20 points (many products in flight, so a product's own
// tile count matters less than for one product), below that the geometry
// the autotuner picked (batch_geom): at 2^14 points one fused row pass
// straight from both operands' limbs, from 2^15 a column pass (of 2 .. 16
// points in registers up to 2^18) before the row pass.  Measured with a
// fixed threshold for pi to 1e9 digits, large tiles from 2^15 / 2^16 /
// 2^17 / 2^18 / 2^19 points: 1.257 / 1.246 / 1.239 / 1.239 / 1.245 s.
constexpr int kGeomMinLog = 13, kGeomMaxLog = 20;  // autotuned geometries (batch_geom)
int batch_pass_geom(int logN);
Plan make_plan(int logN);
// One fused row pass of 2^14 points from both operands' limbs.
Plan make_onepass_plan() {
  Plan pl;
  pl.logN = kBigLogE;
  pl.logE = kBigLogE;
  pl.big = true;
  Pass ps;
  ps.logL = kBigLogE;
  ps.logS = 0;
  ps.logLines = 0;
  ps.logG = 0;
  ps.logB = 0;
  pl.passes.push_back(ps);
  return pl;
}
Plan make_batch_plan(int logN) {
  const bool big = logN > kGeomMaxLog ? logN <= kBigMaxLogN : batch_pass_geom(logN) == 1;
  if (big && logN == kBigLogE) return make_onepass_plan();
  if (big) return make_big_plan(logN);
  return make_plan(logN);
}

Plan make_plan(int logN) {
  if (logN >= kBigMinLogN && logN <= kBigMaxLogN) return make_big_plan(logN);
  Plan pl;
  pl.logN = logN;
  pl.logE = logN < kLogE ? logN : kLogE;
  int m = plan_passes(logN, kMaxLogCol);
  int maxcol = kMaxLogCol;
  if (plan_passes(logN, kMaxLogColWide) < m) {
    maxcol = kMaxLogColWide;
    m = plan_passes(logN, maxcol);
  }
  int last = (logN + m - 1) / m;
  if (last < logN - maxcol * (m - 1)) last = logN - maxcol * (m - 1);
  if (last > kLogE) last = kLogE;
  std::vector<int> ls;
  const int rest = logN - last;
  for (int i = 0; i < m - 1; ++i) ls.push_back(rest / (m - 1) + (i < rest % (m - 1) ? 1 : 0));
  ls.push_back(last);
  int rem = logN;
  for (int i = 0; i < m; ++i) {
    Pass ps;
    ps.logL = ls[i];
    rem -= ps.logL;
    ps.logS = rem;
    ps.logLines = pl.logE - ps.logL;
    ps.logG = ps.logS >= ps.logLines ? ps.logLines : ps.logS;
    ps.logB = ps.logLines - ps.logG;
    pl.passes.push_back(ps);
  }
  return pl;
}

// The same decomposition for the FFT engine (2048 points per workgroup).
Plan make_fft_plan(int logN) {
  Plan pl;
  const int logEmax = kFftLogE;
  pl.logN = logN;
  pl.logE = logN < logEmax ? logN : logEmax;
  int m = (logN + logEmax - 1) / logEmax;
  if (m < 1) m = 1;
  const int base = logN / m, extra = logN % m;
  int rem = logN;
  for (int i = 0; i < m; ++i) {
    Pass ps;
    ps.logL = base + (i < extra ? 1 : 0);
    rem -= ps.logL;
    ps.logS = rem;
    ps.logLines = pl.logE - ps.logL;
    ps.logG = ps.logS >= ps.logLines ? ps.logLines : ps.logS;
    ps.logB = ps.logLines - ps.logG;
    pl.passes.push_back(ps);
  }
  return pl;
}

// Kernel instance of an NTT pass: column passes (stride > 1) for line
// lengths 2^4..2^9 with 4096 staged points, row passes (stride 1) for
// 2^6..2^12 with 4096 staged points or a single line of 2^8..2^11.
using PassFn = void (*)(PassArgs);
template <int MODE, int LOGL, int LOGLINES, bool ROW>
PassFn pass_fn() {
  return &k_ntt_pass<MODE, LOGL, LOGLINES, ROW>;
}
#define PCG_COL(M, L) \
  case L: return pass_fn<M, L, kLogE - L, false>();
#define PCG_ROW(M, L) \
  case L: return pass_fn<M, L, kLogE - L, true>();
template <int M>
PassFn col_pass_fn(int logL) {
  switch (logL) {
    PCG_COL(M, 4) PCG_COL(M, 5) PCG_COL(M, 6) PCG_COL(M, 7) PCG_COL(M, 8) PCG_COL(M, 9)
  }
  return nullptr;
}
template <int M>
PassFn row_pass_fn(int logL, int logLines) {
  if (logLines == 0) {
    switch (logL) {
      case 8: return pass_fn<M, 8, 0, true>();
      case 9: return pass_fn<M, 9, 0, true>();
      case 10: return pass_fn<M, 10, 0, true>();
      case 11: return pass_fn<M, 11, 0, true>();
      case 12: return pass_fn<M, 12, 0, true>();
    }
    return nullptr;
  }
  if (logL + logLines != kLogE) return nullptr;
  switch (logL) {
    PCG_ROW(M, 6) PCG_ROW(M, 7) PCG_ROW(M, 8) PCG_ROW(M, 9) PCG_ROW(M, 10) PCG_ROW(M, 11)
  }
  return nullptr;
}
#undef PCG_COL
#undef PCG_ROW

// Large-tile pass instances: columns of 2^7 .. 2^11 points, rows of 2^14.
template <int M>
PassFn big_col_fn(int logL) {
  switch (logL) {
    case 1: return &k_ntt_col_reg<M, 1>;
    case 2: return &k_ntt_col_reg<M, 2>;
    case 3: return &k_ntt_col_reg<M, 3>;
    case 4: return &k_ntt_col_reg<M, 4>;
    case 5: return &k_ntt_col<M, 5>;
    case 6: return &k_ntt_col<M, 6>;
    case 7: return &k_ntt_col<M, 7>;
    case 8: return &k_ntt_col<M, 8>;
    case 9: return &k_ntt_col<M, 9>;
    case 10: return &k_ntt_col<M, 10>;
    case 11: return &k_ntt_col<M, 11>;
  }
  return nullptr;
}

// First forward passes from the limbs (two half-length DFTs per 2^14-point
// tile).  The 2^11 pass addresses a tile with 32-bit byte offsets (logS <=
// 17).
PassFn big_col_limbs_fn(int logL, int logS) {
  if (logL == 11 && logS > 17) return nullptr;
  switch (logL) {
    case 6: return &k_ntt_col_limbs<6>;
    case 7: return &k_ntt_col_limbs<7>;
    case 8: return &k_ntt_col_limbs<8>;
    case 9: return &k_ntt_col_limbs<9>;
    case 10: return &k_ntt_col_limbs<10>;
    case 11: return &k_ntt_col_limbs11w;
  }
  return nullptr;
}

constexpr int kMinLogN = 8;

struct Ctx {
  std::mutex mu;
  bool ready = false;
  int device = -1;
  hipStream_t stream = nullptr;
  double2* tw[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  Prime* d_prime = nullptr;  // [31][2]
  double2 crt;
  // workspace
  double* spec = nullptr;  // [2 operands][2 primes][N]
  size_t spec_points = 0;
  double* top_spec = nullptr;  // the short product of a wrapped one, [2][2][M]
  size_t top_points = 0;
  // Generation of the transforms in spec / top_spec: bumped whenever a
  // product writes operand a's half of them or they are reallocated.
  uint64_t spec_gen = 1;
  CarryFn* agg = nullptr;
  uint64_t* cin = nullptr;
  size_t nblk_cap = 0;
  uint32_t* status = nullptr;
  // staging for host-pointer calls
  uint32_t* d_stage = nullptr;
  size_t stage_limbs = 0;
  // batch pools: pinned host mirror and device copy, plus descriptors
  uint32_t* h_pool = nullptr;
  uint32_t* d_pool = nullptr;
  size_t pool_limbs = 0;
  BatchDesc* h_desc = nullptr;
  BatchDesc* d_desc = nullptr;
  size_t desc_cap = 0;
  // scratch of batched multi-pass products
  double* bspec = nullptr;
  size_t bspec_doubles = 0;
  BatchDesc* wrap_desc = nullptr;  // top descriptors of wrapped batched products
  size_t wrap_desc_cap = 0;
  CarryFn* bagg = nullptr;  // chunk carry functions of batched products
  uint64_t* bcin = nullptr;
  size_t bcarry_cap = 0;
  // bumped whenever bspec, wrap_desc, bagg or bcin is reallocated (captured
  // subtree schedules name them)
  uint64_t scratch_gen = 1;
  // profiling
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t tev[2] = {nullptr, nullptr};
  double prof_ms[3] = {0, 0, 0};
  double prof_total = 0;
  bool prof_valid = false;
  double fft_maxdist = 0;
  int fft_recomputed = 0;
  // FFT engine tables and round-off slot
  double2* ftab[3] = {nullptr, nullptr, nullptr};
  unsigned long long* d_maxdist = nullptr;
  int num_cu = 256;
  // device values: carry-scan blocks and readback slots
  CarryFn* df_agg = nullptr;
  uint64_t* df_cin = nullptr;
  size_t df_nblk_cap = 0;
  unsigned long long* df_slot = nullptr;  // [4] device
  unsigned long long* h_slot = nullptr;   // [4] pinned host
  uint32_t* h_status = nullptr;           // pinned host copy of *status
  hipEvent_t ev_wait = nullptr;           // polled by wait_stream
  // device address of the static scratch array of pcg_mul_dev_noalloc on
  // this context's device (hipGetSymbolAddress at pcg_init)
  unsigned char* noalloc = nullptr;
};

Ctx g_ctx;

// Waits for everything enqueued on the library stream by polling an event
// from this thread.  The runtime's stream synchronisation sleeps until an
// interrupt wakes it; the device-resident operations read a length or a top
// limb back after every few kernels, and that wake-up is a large share of
// each such round trip while the GPU idles.
int wait_stream(Ctx& c) {
  HIPCHK(hipEventRecord(c.ev_wait, c.stream));
  for (;;) {
    const hipError_t e = hipEventQuery(c.ev_wait);
    if (e == hipSuccess) return PCG_OK;
    if (e != hipErrorNotReady) HIPCHK(e);
  }
}

// ---- scratch of the allocation-free product (pcg_mul_dev_noalloc) ----
// Products of na + nb <= 2^20 limbs run the plain (unwrapped) transform of
// at most 2^20 points: spectra [2 operands][2 primes][2^20] doubles, then the
// carry scan's block functions, scan totals and carry-ins.
constexpr uint64_t kNoAllocMaxLimbs = 1ull << 20;
constexpr int kNoAllocMaxLogN = 20;
constexpr size_t kNoAllocSpecBytes = (size_t)32 << kNoAllocMaxLogN;
constexpr size_t kNoAllocBlocks = kNoAllocMaxLimbs / kCarryBlock + 1;
constexpr size_t kNoAllocAggEntries = kNoAllocBlocks + kNoAllocBlocks / kScanChunk + 2;
constexpr size_t kNoAllocAggBytes = (kNoAllocAggEntries * sizeof(CarryFn) + 255) & ~(size_t)255;
constexpr size_t kNoAllocCinBytes = (kNoAllocBlocks * sizeof(uint64_t) + 255) & ~(size_t)255;
constexpr size_t kNoAllocScratchBytes = kNoAllocSpecBytes + kNoAllocAggBytes + kNoAllocCinBytes;
}  // namespace

// In host code `scratch` names a host-side stand-in, not device memory: only
// the address from hipGetSymbolAddress (Ctx::noalloc) reaches kernels.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

int ceil_log2(uint64_t x) {
  int k = 0;
  while ((1ull << k) < x) ++k;
  return k;
}

int build_tables(Ctx& c) {
  const uint64_t primes[2] = {kP0, kP1};
  std::vector<Prime> hp(31 * 2);
  for (int pi = 0; pi < 2; ++pi) {
    const uint64_t p = primes[pi];
    const uint64_t w30 = powmod_u64(kGen, (p - 1) >> kLogRootMax, p);
    std::vector<double2> lo(32768), hi(32768), tl(kE);
    uint64_t x = 1;
    for (int i = 0; i < 32768; ++i) {
      lo[i] = shoup_pair(x, p);
      x = mulmod_u64(x, w30, p);
    }
    const uint64_t w15 = x;  // w30^(2^15)
    x = 1;
    for (int i = 0; i < 32768; ++i) {
      hi[i] = shoup_pair(x, p);
      x = mulmod_u64(x, w15, p);
    }
    const uint64_t w12 = powmod_u64(w30, 1ull << (kLogRootMax - kLogE), p);
    x = 1;
    std::vector<uint64_t> tl_int(kE);
    for (int i = 0; i < kE; ++i) {
      tl[i] = shoup_pair(x, p);
      tl_int[i] = x;
      x = mulmod_u64(x, w12, p);
    }
    for (int t = 0; t < 3; ++t) {
      std::vector<double2>& v = t == 0 ? tl : (t == 1 ? hi : lo);
      HIPCHK(hipMalloc(&c.tw[pi][t], v.size() * sizeof(double2)));
      HIPCHK(hipMemcpy(c.tw[pi][t], v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice));
    }
    for (int logN = 0; logN <= 30; ++logN) {
      Prime& P = hp[logN * 2 + pi];
      P.p = (double)p;
      P.pinv = 1.0 / (double)p;
      for (int j = 0; j < 3; ++j) {
        P.w8f[j] = tl[(j + 1) * (kE / 8)];
        P.w8i[j] = tl[(kE - (j + 1) * (kE / 8)) & (kE - 1)];
        const int e16 = j == 0 ? 1 : (j == 1 ? 3 : 9);
        P.w16f[j] = tl[e16 * (kE / 16)];
        P.w16i[j] = tl[(kE - e16 * (kE / 16)) & (kE - 1)];
      }
      const uint64_t ninv = powmod_u64(powmod_u64(2, logN, p), p - 2, p);
      P.ninv = shoup_pair(ninv, p);
      P.twL = c.tw[pi][0];
      P.twHi = c.tw[pi][1];
      P.twLo = c.tw[pi][2];
    }
  }
  // FFT engine: e^{2 pi i e / 4096}, e^{2 pi i e / 2^16}, e^{2 pi i e / 2^32}
  {
    const long double tau = 6.283185307179586476925286766559005768L;
    const size_t sz[3] = {4096, 65536, 65536};
    const long double den[3] = {4096.0L, 65536.0L, 4294967296.0L};
    for (int t = 0; t < 3; ++t) {
      std::vector<double2> v(sz[t]);
      for (size_t e = 0; e < sz[t]; ++e) {
        const long double ang = tau * (long double)e / den[t];
        v[e] = make_double2((double)cosl(ang), (double)sinl(ang));
      }
      HIPCHK(hipMalloc(&c.ftab[t], v.size() * sizeof(double2)));
      HIPCHK(hipMemcpy(c.ftab[t], v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&c.d_maxdist, sizeof(unsigned long long)));
  }
  HIPCHK(hipMalloc(&c.d_prime, hp.size() * sizeof(Prime)));
  HIPCHK(hipMemcpy(c.d_prime, hp.data(), hp.size() * sizeof(Prime), hipMemcpyHostToDevice));
  const uint64_t inv = powmod_u64(kP0 % kP1, kP1 - 2, kP1);
  c.crt = shoup_pair(inv, kP1);
  return PCG_OK;
}

int ensure_workspace(Ctx& c, int logN, uint64_t zL) {
  const size_t pts = (size_t)1 << logN;
  if (pts > c.spec_points) {
    HIPCHK(hipStreamSynchronize(c.stream));  // earlier products may still use it
    if (c.spec) HIPCHK(hipFree(c.spec));
    c.spec = nullptr;
    c.spec_points = 0;
    ++c.spec_gen;
    if (hipMalloc(&c.spec, pts * 4 * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PCG_ERR_OOM, "cannot allocate %zu bytes of transform scratch", pts * 32);
    }
    c.spec_points = pts;
  }
  const size_t nblk = (zL + kCarryBlock - 1) / kCarryBlock;
  if (nblk > c.nblk_cap) {
    HIPCHK(hipStreamSynchronize(c.stream));
    if (c.agg) HIPCHK(hipFree(c.agg));
    if (c.cin) HIPCHK(hipFree(c.cin));
    c.agg = nullptr;
    c.cin = nullptr;
    c.nblk_cap = 0;
    HIPCHK(hipMalloc(&c.agg, (nblk + nblk / kScanChunk + 1) * sizeof(CarryFn)));  // + scan totals
    HIPCHK(hipMalloc(&c.cin, nblk * sizeof(uint64_t)));
    c.nblk_cap = nblk;
  }
  return PCG_OK;
}

int ensure_stage(Ctx& c, size_t limbs) {
  if (limbs <= c.stage_limbs) return PCG_OK;
  HIPCHK(hipStreamSynchronize(c.stream));
  if (c.d_stage) HIPCHK(hipFree(c.d_stage));
  c.d_stage = nullptr;
  c.stage_limbs = 0;
  if (hipMalloc(&c.d_stage, limbs * sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PCG_ERR_OOM, "cannot allocate %zu limbs of staging", limbs);
  }
  c.stage_limbs = limbs;
  return PCG_OK;
}

__global__ void k_zero_u32(uint32_t* p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int check_sizes(uint64_t na, uint64_t nb) {
  const uint64_t mn = na < nb ? na : nb;
  if (na + nb > PCG_MAX_TOTAL_LIMBS || mn > PCG_MAX_MIN_LIMBS)
    return fail(PCG_ERR_TOO_LARGE, "FFT length is too small. (operands %llu x %llu limbs exceed the engine limits)",
                (unsigned long long)na, (unsigned long long)nb);
  return PCG_OK;
}

PassFn select_pass(int mode, const Plan& pl, const Pass& ps) {
  if (pl.big) {
    if (ps.logS == 0) {
      if (ps.logL != kBigLogE) return nullptr;
      if (mode == kFusedA) return &k_ntt_row_fa<false, false>;
      if (mode == kSquareA) return &k_ntt_row_fa<true, false>;
      return nullptr;
    }
    if (ps.logS < ps.logLines) return nullptr;  // a tile's columns within one block
    if (mode == kInv && ps.logL == kCol16LogL && ps.logS >= 4 && ps.logS <= 17)
      return &k_ntt_col16_invw;  // 16 adjacent columns per tile (32-bit byte offsets within it)
    if (mode == kFwd) return big_col_fn<kFwd>(ps.logL);
    if (mode == kInv) return big_col_fn<kInv>(ps.logL);
    return nullptr;
  }
  if (ps.logS == 0) {
    if (mode == kFwd) return row_pass_fn<kFwd>(ps.logL, ps.logLines);
    if (mode == kFused) return row_pass_fn<kFused>(ps.logL, ps.logLines);
  } else if (ps.logG == ps.logLines && pl.logE == kLogE) {
    if (mode == kFwd) return col_pass_fn<kFwd>(ps.logL);
    if (mode == kInv) return col_pass_fn<kInv>(ps.logL);
  }
  return nullptr;
}

// Every pass of the plan has a kernel instance for each mode it runs in
// (large-tile plans: the last pass only as the fused pass with a's row DFT).
bool plan_supported(const Plan& pl) {
  const size_t m = pl.passes.size();
  for (size_t j = 0; j < m; ++j) {
    if ((j + 1 < m || !pl.big) && !select_pass(kFwd, pl, pl.passes[j])) return false;
    if (j + 1 < m && !select_pass(kInv, pl, pl.passes[j])) return false;
    if (j + 1 == m && !select_pass(pl.big ? kFusedA : kFused, pl, pl.passes[j])) return false;
  }
  return true;
}

// Scan of the block carry functions (three launches, see carry_kernels.hpp);
// a.tot needs (nblk + 1023) / 1024 entries.
void launch_carry_scan(Ctx& c, const CarryArgs& a) {
  const uint64_t ntot = (a.nblk + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(k_scan_partial, dim3((unsigned)ntot), dim3(kScanChunk), 0, c.stream, a);
  hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kScanChunk), 0, c.stream, a, ntot);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)ntot), dim3(kScanChunk), 0, c.stream, a);
}

// Resident workgroups per compute unit of a pass kernel instance, cached in
// a fixed table (no heap use: the allocation-free entry point plans through
// it); an instance beyond the table's capacity is queried every time.
int pass_occupancy(PassFn fn, int threads) {
  static std::mutex mu;
  static PassFn keys[128];
  static int vals[128];
  static int used = 0;
  std::lock_guard<std::mutex> lk(mu);
  for (int i = 0; i < used; ++i)
    if (keys[i] == fn) return vals[i];
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(fn), threads, 0) !=
          hipSuccess ||
      nb < 1) {
    (void)hipGetLastError();
    nb = 2;
  }
  if (used < 128) {
    keys[used] = fn;
    vals[used] = nb;
    ++used;
  }
  return nb;
}

int launch_pass(Ctx& c, int mode, const Plan& pl, const Pass& ps, const double* src,
                const uint32_t* limbs, uint64_t nlimbs, double* dst, const double* spec,
                const BatchDesc* descs = nullptr, int operand = 0, uint64_t zstride = 0,
                unsigned count = 1, int zshift = 0, int sshift = 0) {
  PassArgs a;
  a.descs = descs;
  a.operand = operand;
  a.zshift = zshift;
  a.sshift = sshift;
  a.zstride = zstride;
  a.prm = c.d_prime + pl.logN * 2;
  a.src = src;
  a.limbs = limbs;
  a.nlimbs = nlimbs;
  a.dst = dst;
  a.spec = spec;
  a.N = 1ull << pl.logN;
  a.logL = ps.logL;
  a.logS = ps.logS;
  a.logLines = ps.logLines;
  a.logG = ps.logG;
  a.logB = ps.logB;
  PassFn fn = select_pass(mode, pl, ps);
  if (!fn)
    return fail(PCG_ERR_INTERNAL, "no NTT pass kernel for logN=%d logL=%d logS=%d mode=%d",
                pl.logN, ps.logL, ps.logS, mode);
  const int E = 1 << pl.logE;
  const int threads = E / (pl.big ? kBigEPT : kEPT);
  a.ntiles = a.N / E;
  if (fn == &k_ntt_col16_invw) {
    if (!src) return fail(PCG_ERR_INTERNAL, "inverse column pass without input");
    a.ntiles /= 2;  // tiles of 16 columns
  }
  if (fn == &k_ntt_row_fa<false, false> && !src) {
    // a one-pass plan: the fused pass transforms both operands from their limbs
    if (!descs) return fail(PCG_ERR_INTERNAL, "one-pass large-tile products run batched only");
    fn = &k_ntt_row_fa<false, true>;
  }
  if (pl.big && mode == kFwd && !src && ps.logS >= kBigLogE + 1 - ps.logL && (descs || nlimbs >= 1)) {
    // first pass from the limbs: two half-length DFTs per tile
    if (PassFn h = big_col_limbs_fn(ps.logL, ps.logS)) {
      fn = h;
      a.ntiles = a.N >> (kBigLogE + 1);  // tiles of 2^14 limbs (2^15 output points)
    }
  }
  // Persistent tiles: as many workgroups as stay resident (occupancy of
  // this instance x compute units), spread over the primes and products.
  int per_cu = pass_occupancy(fn, threads);
  uint64_t gx = ((uint64_t)per_cu * (uint64_t)c.num_cu) / (2ull * count);
  if (gx < 1) gx = 1;
  if (gx > a.ntiles) gx = a.ntiles;
  dim3 grid((unsigned)gx, 2, count), block(threads);
  hipLaunchKernelGGL(fn, grid, block, 0, c.stream, a);
  return PCG_OK;
}

// Transform length of a product of zL limbs: its zL - 1 coefficients.
int ntt_log_points(uint64_t zL) {
  const int logN = ceil_log2(zL > 1 ? zL - 1 : 1);
  return logN < kMinLogN ? kMinLogN : logN;
}

// Passes of operand a's forward transform that run on their own: all of
// them, but for large-tile plans the last (row) pass runs inside the fused
// pass (k_ntt_row_fa), which reads a after its column pass.
size_t a_own_passes(const Plan& pl) { return pl.big ? pl.passes.size() - 1 : pl.passes.size(); }

// Forward transform of operand a into A ([2][N]): its spectrum, or for
// large-tile plans a after its column pass (what the fused pass reads).
int ntt_forward_into(Ctx& c, const Plan& pl, const uint32_t* d_a, uint64_t na, double* A) {
  for (size_t j = 0; j < a_own_passes(pl); ++j) {
    const int rc = launch_pass(c, kFwd, pl, pl.passes[j], j == 0 ? nullptr : A, d_a, na, A, nullptr);
    if (rc) return rc;
  }
  return PCG_OK;
}

// With a's transform in A (ntt_forward_into): forward transform of b (its
// last pass fused with the pointwise product and the first inverse pass)
// and the inverse passes; B receives the convolution's residues ([2][N]).
int ntt_tail_into(Ctx& c, const Plan& pl, const uint32_t* d_b, uint64_t nb, const double* A, double* B) {
  const size_t m = pl.passes.size();
  int rc;
  for (size_t j = 0; j + 1 < m; ++j) {
    rc = launch_pass(c, kFwd, pl, pl.passes[j], j == 0 ? nullptr : B, d_b, nb, B, nullptr);
    if (rc) return rc;
  }
  rc = launch_pass(c, pl.big ? kFusedA : kFused, pl, pl.passes[m - 1], m == 1 ? nullptr : B, d_b, nb, B, A);
  if (rc) return rc;
  for (size_t j = m - 1; j-- > 0;) {
    rc = launch_pass(c, kInv, pl, pl.passes[j], B, nullptr, 0, B, nullptr);
    if (rc) return rc;
  }
  return PCG_OK;
}

// A product whose coefficient count zL - 1 exceeds a power of two N1 by at
// most N1 / 4 runs as the cyclic convolution of length N1 (the top
// coefficients wrap onto the bottom ones) plus the short product of the top
// w = zL - 1 - N1 limbs of each operand, which gives those top coefficients
// exactly; the carry subtracts them below N1 and uses them from N1 up.
// Cost: a transform of N1 and one of 2^ceil(log2(2w - 1)) <= N1 / 2 points
// instead of one of 2 N1 (the products of the largest tree levels and of
// the Newton steps sit a few percent above a power of two).
struct WrapPlan {
  bool use = false;
  int logN1 = 0, logM = 0;
  uint64_t w = 0;
};
constexpr int kWrapMinLog = 20;

WrapPlan wrap_plan(uint64_t na, uint64_t nb) {
  WrapPlan wp;
  const uint64_t zL = na + nb;
  const int logN = ntt_log_points(zL);
  if (logN - 1 < kWrapMinLog) return wp;
  const uint64_t N1 = 1ull << (logN - 1);
  const uint64_t w = zL - 1 - N1;  // >= 1: logN is the length the coefficients need
  if (w > N1 / 4 || na > N1 || nb > N1 || na < w || nb < w) return wp;
  wp.use = true;
  wp.logN1 = logN - 1;
  wp.logM = ntt_log_points(2 * w);  // 2w - 1 coefficients
  wp.w = w;
  return wp;
}

// Scratch of the short product ([2 operands][2 primes][M]).
int ensure_top_scratch(Ctx& c, int logM) {
  const size_t pts = (size_t)1 << logM;
  if (pts <= c.top_points) return PCG_OK;
  HIPCHK(hipStreamSynchronize(c.stream));
  if (c.top_spec) HIPCHK(hipFree(c.top_spec));
  c.top_spec = nullptr;
  c.top_points = 0;
  ++c.spec_gen;
  if (hipMalloc(&c.top_spec, pts * 4 * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PCG_ERR_OOM, "cannot allocate %zu bytes of transform scratch", pts * 32);
  }
  c.top_points = pts;
  return PCG_OK;
}

// The short product of a wrapped one: the top w limbs of a and b convolved
// into c.top_spec's second half (residues, no carry).
int ntt_wrap_top(Ctx& c, const WrapPlan& wp, const uint32_t* d_a, uint64_t na, const uint32_t* d_b,
                 uint64_t nb, bool a_kept = false) {
  const Plan pm = make_plan(wp.logM);
  const uint64_t M = 1ull << wp.logM;
  double* A2 = c.top_spec;
  double* B2 = c.top_spec + 2 * M;
  if (!a_kept) {
    const int rc = ntt_forward_into(c, pm, d_a + (na - wp.w), wp.w, A2);
    if (rc) return rc;
  }
  return ntt_tail_into(c, pm, d_b + (nb - wp.w), wp.w, A2, B2);
}

// Scratch of the carry scan: block functions, scan totals, carry-ins.
struct CarryScratch {
  CarryFn* agg;
  CarryFn* tot;
  uint64_t* cin;
};

// CRT and carries of the residues in B into d_out[0, zL); wp.use: B holds a
// wrapped convolution and c.top_spec the top coefficients.
int ntt_carry(Ctx& c, const Plan& pl, const double* B, uint64_t zL, uint32_t* d_out, const WrapPlan& wp,
              const CarryScratch& cs) {
  CarryArgs ca;
  memset(&ca, 0, sizeof ca);
  ca.prm = c.d_prime + pl.logN * 2;
  ca.r = B;
  ca.N = 1ull << pl.logN;
  ca.zL = zL;
  ca.out = d_out;
  ca.agg = cs.agg;
  ca.cin = cs.cin;
  ca.nblk = (zL + kCarryBlock - 1) / kCarryBlock;
  ca.tot = cs.tot;
  ca.crt = c.crt;
  ca.status = c.status;
  if (wp.use) {
    const uint64_t M = 1ull << wp.logM;
    ca.top = c.top_spec + 2 * M;
    ca.top_stride = M;
    ca.top_w = wp.w;
  }
  hipLaunchKernelGGL(k_carry_local<false>, dim3((unsigned)ca.nblk), dim3(kCarryThreads), 0, c.stream, ca);
  launch_carry_scan(c, ca);
  hipLaunchKernelGGL(k_carry_fix, dim3((unsigned)((ca.nblk + 255) / 256)), dim3(256), 0, c.stream,
                     ca);
  HIPCHK(hipGetLastError());
  return PCG_OK;
}

int ntt_carry(Ctx& c, const Plan& pl, const double* B, uint64_t zL, uint32_t* d_out, const WrapPlan& wp) {
  return ntt_carry(c, pl, B, zL, d_out, wp, CarryScratch{c.agg, c.agg + c.nblk_cap, c.cin});
}

// The operand whose transform spec's first half (and top_spec's, for a
// wrapped product) holds, kept for the next product (PCG_KEEP_B).
struct KeptSpectrum {
  const uint32_t* t = nullptr;
  uint64_t L = 0;
  int logN = 0;
  WrapPlan wp;
  uint64_t gen = 0;
};
KeptSpectrum g_kept;

// Plan of a product: the full transform, or the wrapped one.
struct MulPlan {
  WrapPlan wp;
  int logN;  // the main transform
};
MulPlan mul_plan(uint64_t na, uint64_t nb) {
  MulPlan mp;
  mp.wp = wrap_plan(na, nb);
  mp.logN = mp.wp.use ? mp.wp.logN1 : ntt_log_points(na + nb);
  return mp;
}

int ensure_mul_workspace(Ctx& c, const MulPlan& mp, uint64_t zL) {
  int rc = ensure_workspace(c, mp.logN, zL);
  if (!rc && mp.wp.use) rc = ensure_top_scratch(c, mp.wp.logM);
  return rc;
}

// Forward a (into c.spec's first half) is done; the rest of the product
// of a (na limbs) and b into d_out.
int ntt_mul_tail(Ctx& c, const Plan& pl, const MulPlan& mp, const uint32_t* d_a, uint64_t na,
                 const uint32_t* d_b, uint64_t nb, uint32_t* d_out, bool profile, bool a_kept = false) {
  const uint64_t N = 1ull << pl.logN;
  if (mp.wp.use) {
    const int rc = ntt_wrap_top(c, mp.wp, d_a, na, d_b, nb, a_kept);
    if (rc) return rc;
  }
  int rc = ntt_tail_into(c, pl, d_b, nb, c.spec, c.spec + 2 * N);
  if (rc) return rc;
  if (profile) HIPCHK(hipEventRecord(c.ev[2], c.stream));
  rc = ntt_carry(c, pl, c.spec + 2 * N, na + nb, d_out, mp.wp);
  if (rc) return rc;
  if (profile) HIPCHK(hipEventRecord(c.ev[3], c.stream));
  return PCG_OK;
}

// Enqueue the whole NTT product on the library stream (device pointers).
// keep: kKeepA (record operand a's transform for the next product) and/or
// kReuseA (operand a's transform is the one recorded: skip it when the plan
// matches and nothing has overwritten it since).
constexpr uint32_t kKeepA = 1, kReuseA = 2;
int ntt_mul_enqueue(Ctx& c, const uint32_t* d_a, uint64_t na, const uint32_t* d_b, uint64_t nb,
                    uint32_t* d_out, bool profile, uint32_t keep = 0) {
  const uint64_t zL = na + nb;
  const MulPlan mp = mul_plan(na, nb);
  int rc = ensure_mul_workspace(c, mp, zL);
  if (rc) return rc;
  const Plan pl = make_plan(mp.logN);
  if (profile) HIPCHK(hipEventRecord(c.ev[0], c.stream));
  const size_t m = pl.passes.size();
  const KeptSpectrum& K = g_kept;
  // the main transform of a depends on its limbs and the length only; the
  // short one of a wrapped product on the wrap too
  const bool reuse = (keep & kReuseA) && K.t == d_a && K.L == na && K.gen == c.spec_gen &&
                     K.logN == mp.logN;
  const bool top_kept = reuse && mp.wp.use && K.wp.use && K.wp.logM == mp.wp.logM && K.wp.w == mp.wp.w;
  if (!reuse) ++c.spec_gen;
  if (keep & kKeepA)
    g_kept = KeptSpectrum{d_a, na, mp.logN, mp.wp, c.spec_gen};
  else if (reuse && mp.wp.use)
    g_kept.wp = mp.wp;  // the short transform is redone for this product's wrap
  if (reuse) return ntt_mul_tail(c, pl, mp, d_a, na, d_b, nb, d_out, profile, top_kept);
  if (d_a == d_b && na == nb && m >= 2) {
    // a square: the fused pass takes a's forward transform but its last
    // pass as its input, so the last pass also writes into A (out of place)
    // and b's forward passes are skipped (2 m instead of 3 m - 1 sweeps).
    // Large-tile plans run that last pass inside the fused pass (one row
    // DFT, the pointwise square: 2 m - 1 sweeps); a's column pass writes A,
    // which is a's transform as ntt_forward_into leaves it (PCG_KEEP_B).
    const uint64_t N = 1ull << pl.logN;
    double* A = c.spec;
    double* B = c.spec + 2 * N;
    double* X = pl.big ? A : B;
    for (size_t j = 0; j + 1 < m; ++j) {
      rc = launch_pass(c, kFwd, pl, pl.passes[j], j == 0 ? nullptr : X, d_a, na, X, nullptr);
      if (rc) return rc;
    }
    if (!pl.big) {
      rc = launch_pass(c, kFwd, pl, pl.passes[m - 1], B, nullptr, 0, A, nullptr);
      if (rc) return rc;
    }
    if (profile) HIPCHK(hipEventRecord(c.ev[1], c.stream));
    if (mp.wp.use) {
      rc = ntt_wrap_top(c, mp.wp, d_a, na, d_b, nb);
      if (rc) return rc;
    }
    rc = pl.big ? launch_pass(c, kSquareA, pl, pl.passes[m - 1], A, nullptr, 0, B, A)
                : launch_pass(c, kFused, pl, pl.passes[m - 1], B, nullptr, 0, B, A);
    for (size_t j = m - 1; !rc && j-- > 0;) rc = launch_pass(c, kInv, pl, pl.passes[j], B, nullptr, 0, B, nullptr);
    if (rc) return rc;
    if (profile) HIPCHK(hipEventRecord(c.ev[2], c.stream));
    rc = ntt_carry(c, pl, B, zL, d_out, mp.wp);
    if (rc) return rc;
    if (profile) HIPCHK(hipEventRecord(c.ev[3], c.stream));
    return PCG_OK;
  }
  rc = ntt_forward_into(c, pl, d_a, na, c.spec);
  if (rc) return rc;
  if (profile) HIPCHK(hipEventRecord(c.ev[1], c.stream));
  return ntt_mul_tail(c, pl, mp, d_a, na, d_b, nb, d_out, profile);
}


void launch_fft_pass(Ctx& c, int mode, const Plan& pl, const Pass& ps, const double2* src,
                     const uint32_t* limbs, uint64_t nlimbs, double2* dst, const double2* spec,
                     bool unweight) {
  FftPassArgs a;
  a.tb.stg = c.ftab[0];
  a.tb.hi = c.ftab[1];
  a.tb.lo = c.ftab[2];
  a.src = src;
  a.limbs = limbs;
  a.nlimbs = nlimbs;
  a.dst = dst;
  a.spec = spec;
  a.N = 1ull << pl.logN;
  a.logN = pl.logN;
  a.logL = ps.logL;
  a.logS = ps.logS;
  a.logLines = ps.logLines;
  a.logG = ps.logG;
  a.logB = ps.logB;
  a.unweight = unweight ? 1 : 0;
  const int E = 1 << pl.logE;
  dim3 grid((unsigned)(a.N / E)), block(E / kEPTc);
  if (mode == kFwd)
    hipLaunchKernelGGL(k_fft_pass<kFwd>, grid, block, 0, c.stream, a);
  else if (mode == kInv)
    hipLaunchKernelGGL(k_fft_pass<kInv>, grid, block, 0, c.stream, a);
  else
    hipLaunchKernelGGL(k_fft_pass<kFused>, grid, block, 0, c.stream, a);
}

// Enqueue the FFT-engine product; the round-off maximum lands in d_maxdist.
int fft_mul_enqueue(Ctx& c, const uint32_t* d_a, uint64_t na, const uint32_t* d_b, uint64_t nb,
                    uint32_t* d_out) {
  const uint64_t zL = na + nb;
  int logN = ceil_log2(zL);
  if (logN < kMinLogN) logN = kMinLogN;
  ++c.spec_gen;  // the FFT's data overwrite the kept transform
  int rc = ensure_workspace(c, logN, zL);
  if (rc) return rc;
  const Plan pl = make_fft_plan(logN);
  const uint64_t N = 1ull << logN;
  double2* A = reinterpret_cast<double2*>(c.spec);
  double2* B = A + N;
  const size_t m = pl.passes.size();
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c.stream,
                     reinterpret_cast<uint32_t*>(c.d_maxdist), (uint64_t)2);
  for (size_t j = 0; j < m; ++j)
    launch_fft_pass(c, kFwd, pl, pl.passes[j], j == 0 ? nullptr : A, d_a, na, A, nullptr, false);
  for (size_t j = 0; j + 1 < m; ++j)
    launch_fft_pass(c, kFwd, pl, pl.passes[j], j == 0 ? nullptr : B, d_b, nb, B, nullptr, false);
  launch_fft_pass(c, kFused, pl, pl.passes[m - 1], m == 1 ? nullptr : B, d_b, nb, B, A, m == 1);
  for (size_t j = m - 1; j-- > 0;)
    launch_fft_pass(c, kInv, pl, pl.passes[j], B, nullptr, 0, B, nullptr, j == 0);
  CarryArgs ca;
  memset(&ca, 0, sizeof ca);
  ca.z = B;
  ca.N = N;
  ca.zL = zL;
  ca.out = d_out;
  ca.agg = c.agg;
  ca.cin = c.cin;
  ca.nblk = (zL + kCarryBlock - 1) / kCarryBlock;
  ca.tot = c.agg + c.nblk_cap;
  ca.status = nullptr;
  ca.maxdist = c.d_maxdist;
  hipLaunchKernelGGL(k_carry_local<true>, dim3((unsigned)ca.nblk), dim3(kCarryThreads), 0, c.stream, ca);
  launch_carry_scan(c, ca);
  hipLaunchKernelGGL(k_carry_fix, dim3((unsigned)((ca.nblk + 255) / 256)), dim3(256), 0, c.stream, ca);
  HIPCHK(hipGetLastError());
  return PCG_OK;
}

int mul_dev_locked(Ctx& c, const uint32_t* d_a, uint64_t na, const uint32_t* d_b, uint64_t nb,
                   uint32_t* d_out, uint32_t flags, uint32_t keep = 0) {
  int rc = check_sizes(na, nb);
  if (rc) return rc;
  const uint64_t zL = na + nb;
  if (zL == 0) return PCG_OK;
  if (na == 0 || nb == 0) {
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((zL + 255) / 256)), dim3(256), 0, c.stream, d_out, zL);
    HIPCHK(hipGetLastError());
    return PCG_OK;
  }
  const uint32_t eng = flags & PCG_ENGINE_MASK;
  if (eng == 3) return fail(PCG_ERR_INVALID, "unknown engine flag");
  if (eng == PCG_ENGINE_FFT) {
    // FP64 FFT, then the round-off test decides whether the NTT recomputes.
    rc = fft_mul_enqueue(c, d_a, na, d_b, nb, d_out);
    if (rc) return rc;
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, c.d_maxdist, sizeof bits, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    double dist;
    memcpy(&dist, &bits, sizeof dist);
    c.fft_maxdist = dist;
    c.fft_recomputed = !(dist <= 0.25);
    if (!c.fft_recomputed) return PCG_OK;
  }
  const bool prof = (flags & PCG_PROFILE) != 0;
  rc = ntt_mul_enqueue(c, d_a, na, d_b, nb, d_out, prof, keep);
  if (rc) return rc;
  if (prof) {
    HIPCHK(hipEventSynchronize(c.ev[3]));
    float ms;
    for (int i = 0; i < 3; ++i) {
      HIPCHK(hipEventElapsedTime(&ms, c.ev[i], c.ev[i + 1]));
      c.prof_ms[i] = ms;
    }
    HIPCHK(hipEventElapsedTime(&ms, c.ev[0], c.ev[3]));
    c.prof_total = ms;
    c.prof_valid = true;
  }
  return PCG_OK;
}


int ensure_pool(Ctx& c, size_t limbs, size_t ndesc) {
  if (limbs > c.pool_limbs) {
    HIPCHK(hipStreamSynchronize(c.stream));
    if (c.h_pool) HIPCHK(hipHostFree(c.h_pool));
    if (c.d_pool) HIPCHK(hipFree(c.d_pool));
    c.h_pool = nullptr;
    c.d_pool = nullptr;
    c.pool_limbs = 0;
    size_t want = limbs + limbs / 4;
    if (hipHostMalloc((void**)&c.h_pool, want * 4, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&c.d_pool, want * 4) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PCG_ERR_OOM, "cannot allocate batch pools of %zu limbs", want);
    }
    c.pool_limbs = want;
  }
  if (ndesc > c.desc_cap) {
    HIPCHK(hipStreamSynchronize(c.stream));
    if (c.h_desc) HIPCHK(hipHostFree(c.h_desc));
    if (c.d_desc) HIPCHK(hipFree(c.d_desc));
    c.h_desc = nullptr;
    c.d_desc = nullptr;
    c.desc_cap = 0;
    size_t want = ndesc + ndesc / 4 + 16;
    if (hipHostMalloc((void**)&c.h_desc, want * sizeof(BatchDesc), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&c.d_desc, want * sizeof(BatchDesc)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PCG_ERR_OOM, "cannot allocate %zu batch descriptors", want);
    }
    c.desc_cap = want;
  }
  return PCG_OK;
}

// Size classes of the batched paths.  Small products (zL = na + nb in
// (2^(b-1), 2^b], b <= 12) go to the class g_small_class[b]: 0 the
// 256-thread schoolbook (zL <= 1024), 1, 2 and 3 the grouped schoolbook with
// 16 threads x 4 columns, 64 x 4 and 64 x 8 per product (zL <= 64, <= 256,
// <= 512), 4 one Karatsuba level over the schoolbook (zL <= 1024),
// 10..12 the single-workgroup NTT of 2^k points, 110..112 the batched pass
// kernels at 2^(k - 100) points (spectra through HBM); the autotuner
// (tune_small) picks among them
// per bucket on the device at first use.  Classes 13..21 are the multi-pass
// NTT run for the whole class at once (in groups bounded by the batched
// scratch), -1 one product at a time (from 2^22 points one product's tiles
// fill the GPU by themselves).
constexpr int kBatchMaxLog = 21;
constexpr int kSmallNttMinLog = 10;
int g_small_class[kSmallNttMaxLog + 1];
bool g_small_ready = false;
uint64_t g_table_gen = 1;  // bumped whenever g_small_class changes

// Pass geometry of batched multi-pass products at 2^b points, b in
// [kGeomMinLog, kGeomMaxLog] (the class table's entries 13 .. 20): 0 the
// 4096-point passes (512 threads x 8 points), 1 the large-tile passes
// (2^14-point tiles, 1024 threads x 16 points; at 2^14 points one fused
// pass from the limbs; at 2^13 points the same one-pass product at 2^14
// points, twice the points in one sweep instead of five).  The autotuner
// times both per size; below, only the 4096-point passes exist, above
// only the large tiles.
int batch_geom_default(int b) { return b == 14 || b >= 17 ? 1 : 0; }
bool batch_geom_valid(int b, int g) { return g == 0 || g == 1; }
int g_batch_geom[kGeomMaxLog + 1] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 1, 1, 1};
int batch_pass_geom(int logN) {
  return logN >= kGeomMinLog && logN <= kGeomMaxLog ? g_batch_geom[logN] : (logN > kGeomMaxLog ? 1 : 0);
}
// A batched launch of class 2^13 on geometry 1 runs at 2^14 points (its
// wrapped products need no wrap there: zL - 1 <= 5/4 2^13).
bool batch_onepass13(int logN) { return logN == kGeomMinLog && g_batch_geom[kGeomMinLog] == 1; }

// Largest zL of a schoolbook class.
int school_max(int cls) { return cls == 1 ? 64 : cls == 2 ? 256 : cls == 3 ? 512 : kSchoolMaxLimbs; }

int small_class_default(int b) {
  if ((1 << b) <= 64) return 1;
  if ((1 << b) <= 256) return 2;
  if ((1 << b) <= kSchoolMaxLimbs) return 0;
  return b < kSmallNttMinLog ? kSmallNttMinLog : b;
}

constexpr int kPassClass = 100;  // + log2 N: batched pass kernels

constexpr int kKaraClass = 4;  // one Karatsuba level over the schoolbook (k_batch_kara)

bool small_class_valid(int b, int cls) {
  if (cls >= 0 && cls <= 3) return (1 << b) <= school_max(cls);
  if (cls == kKaraClass) return (1 << b) <= kSchoolMaxLimbs;
  if (cls >= kPassClass + kSmallNttMinLog && cls <= kPassClass + kSmallNttMaxLog) return cls - kPassClass >= b;
  return cls >= kSmallNttMinLog && cls <= kSmallNttMaxLog && cls >= b;
}

// Slot of a small class in launch order (0..9).
int small_slot(int cls) {
  if (cls <= 3) return cls;
  if (cls == kKaraClass) return 10;
  if (cls < kPassClass) return cls - kSmallNttMinLog + 4;
  return cls - kPassClass - kSmallNttMinLog + 7;
}
int small_slot_class(int slot) {
  if (slot <= 3) return slot;
  if (slot == 10) return kKaraClass;
  if (slot <= 6) return slot - 4 + kSmallNttMinLog;
  return slot - 7 + kSmallNttMinLog + kPassClass;
}

int batch_class(uint64_t na, uint64_t nb) {
  const uint64_t zL = na + nb;
  const int lg = ceil_log2(zL);
  if (lg <= kSmallNttMaxLog) return g_small_class[lg];
  if (lg <= kBatchMaxLog) return lg;
  return -1;
}

// Class of a product of at most zL limbs on the device-resident path:
// the small classes above, else the transform length (log2).
int bsr_class(const std::array<int, kSmallNttMaxLog + 1>& tab, uint64_t zL) {
  const int lg = ceil_log2(zL);
  if (lg <= kSmallNttMaxLog) return tab[lg];
  return lg < kMinLogN ? kMinLogN : lg;
}

// Launch key of a subtree product whose operands have at most ta and tb
// limbs: its class (bsr_class) times 32 plus, for a multi-pass product whose
// ta + tb - 1 coefficients exceed a power of two N1 >= 2^12 by at most N1 / 4,
// log2 of the short top product's transform (the product then runs wrapped
// at N1 points, as ntt_mul_enqueue's wrapped products).
constexpr int kWrapKey = 32;
constexpr int kBatchWrapMinLog = 12;
int bsr_key(const std::array<int, kSmallNttMaxLog + 1>& tab, uint64_t ta, uint64_t tb) {
  const int cls = bsr_class(tab, ta + tb);
  if (cls <= kSmallNttMaxLog || cls >= kPassClass) return cls * kWrapKey;
  const uint64_t C = ta + tb - 1;  // coefficients
  const int lg = ceil_log2(C);
  if (lg < cls && lg > kSmallNttMaxLog) return lg * kWrapKey;  // coefficients fit one size down
  if (lg - 1 < kBatchWrapMinLog) return cls * kWrapKey;
  const uint64_t N1 = 1ull << (lg - 1), wb = C - N1;
  if (wb > N1 / 4 || ta > N1 || tb > N1) return cls * kWrapKey;
  return (lg - 1) * kWrapKey + ntt_log_points(2 * wb);
}

// Classes that run on the batched pass kernels (a transform spectrum in HBM
// that two products can share).
bool class_pairs(int cls) { return cls >= kPassClass || cls > kSmallNttMaxLog; }

// Multi-pass NTT products of one transform length, all at once: pass
// kernels carry a product grid dimension; scratch is 4N doubles per product.
// pairs: products 2k and 2k+1 have the same operand a, whose forward
// transform runs once (count is even).
// Group size, scratch per product and carry chunks per product of a batched
// multi-pass launch at 2^logN points (wrapped: top products of 2^logM).
struct BatchGeom {
  uint64_t N, M, per;
  size_t group_max;
  uint32_t nch;
};
BatchGeom batch_geom(int logN, bool pairs, int logM) {
  BatchGeom g;
  g.N = 1ull << logN;
  g.M = logM ? 1ull << logM : 0;
  g.per = 4 * g.N + 4 * g.M;  // main spectra, then the top product's
  g.group_max = std::max<size_t>(1, ((size_t)1 << 28) / g.per);  // 2 GiB of scratch
  if (pairs) g.group_max = std::max<size_t>(2, g.group_max & ~(size_t)1);
  // chunk carry functions: (products per group) x (chunks per product); a
  // wrapped product has up to N + M / 2 + 1 limbs
  g.nch = (uint32_t)((g.N + g.M / 2 + 1 + kCarryBlock - 1) / kCarryBlock);
  return g;
}

// Grows the batched scratch to what such a launch needs (synchronising the
// stream first); nothing is allocated when it already suffices.
int batched_reserve(Ctx& c, size_t count, int logN, bool pairs, int logM) {
  if (batch_onepass13(logN)) {
    logN = kBigLogE;
    logM = 0;
  }
  const BatchGeom g = batch_geom(logN, pairs, logM);
  const size_t want = std::min(count, g.group_max) * g.per;
  const size_t carry_need = std::min(count, g.group_max) * (size_t)g.nch;
  if (carry_need > c.bcarry_cap) {
    ++c.scratch_gen;
    HIPCHK(hipStreamSynchronize(c.stream));
    if (c.bagg) HIPCHK(hipFree(c.bagg));
    if (c.bcin) HIPCHK(hipFree(c.bcin));
    c.bagg = nullptr;
    c.bcin = nullptr;
    c.bcarry_cap = 0;
    const size_t cap = std::max<size_t>(carry_need, (size_t)1 << 16);
    HIPCHK(hipMalloc(&c.bagg, cap * sizeof(CarryFn)));
    HIPCHK(hipMalloc(&c.bcin, cap * sizeof(uint64_t)));
    c.bcarry_cap = cap;
  }
  if (want > c.bspec_doubles) {
    ++c.scratch_gen;
    HIPCHK(hipStreamSynchronize(c.stream));
    if (c.bspec) HIPCHK(hipFree(c.bspec));
    c.bspec = nullptr;
    c.bspec_doubles = 0;
    if (hipMalloc(&c.bspec, want * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PCG_ERR_OOM, "cannot allocate batched transform scratch");
    }
    c.bspec_doubles = want;
  }
  if (logM && std::min(count, g.group_max) > c.wrap_desc_cap) {
    ++c.scratch_gen;
    HIPCHK(hipStreamSynchronize(c.stream));
    if (c.wrap_desc) HIPCHK(hipFree(c.wrap_desc));
    c.wrap_desc = nullptr;
    c.wrap_desc_cap = 0;
    const size_t cap = std::max<size_t>(std::min(count, g.group_max), 1024);
    HIPCHK(hipMalloc(&c.wrap_desc, cap * sizeof(BatchDesc)));
    c.wrap_desc_cap = cap;
  }
  return PCG_OK;
}

int batched_multipass(Ctx& c, const BatchDesc* d_desc, size_t count, int logN, bool pairs = false,
                      int logM = 0) {
  if (pairs && (count & 1)) return fail(PCG_ERR_INTERNAL, "odd count of paired products");
  const int rr = batched_reserve(c, count, logN, pairs, logM);
  if (rr) return rr;
  const bool onepass = batch_onepass13(logN);
  if (onepass) {
    logN = kBigLogE;
    logM = 0;
  }
  const BatchGeom g = batch_geom(logN, pairs, logM);
  const uint64_t N = g.N, M = g.M, per = g.per;
  const size_t group_max = g.group_max;
  const uint32_t nch = g.nch;
  const Plan pl = onepass ? make_onepass_plan() : make_batch_plan(logN);
  const size_t m = pl.passes.size();
  double* A = c.bspec;
  double* B = c.bspec + 2 * N;
  for (size_t g0 = 0; g0 < count; g0 += group_max) {
    const unsigned k = (unsigned)std::min(group_max, count - g0);
    const BatchDesc* dd = d_desc + g0;
    int rc;
    const int sh = pairs ? 1 : 0;
    for (size_t j = 0; j < a_own_passes(pl); ++j) {
      rc = launch_pass(c, kFwd, pl, pl.passes[j], j == 0 ? nullptr : A, nullptr, 0, A, nullptr,
                       dd, 0, per, k >> sh, sh, 0);
      if (rc) return rc;
    }
    for (size_t j = 0; j + 1 < m; ++j) {
      rc = launch_pass(c, kFwd, pl, pl.passes[j], j == 0 ? nullptr : B, nullptr, 0, B, nullptr,
                       dd, 1, per, k);
      if (rc) return rc;
    }
    rc = launch_pass(c, pl.big ? kFusedA : kFused, pl, pl.passes[m - 1], m == 1 ? nullptr : B, nullptr, 0,
                     B, A, dd, 1, per, k, 0, sh);
    if (rc) return rc;
    for (size_t j = m - 1; j-- > 0;) {
      rc = launch_pass(c, kInv, pl, pl.passes[j], B, nullptr, 0, B, nullptr, dd, 1, per, k);
      if (rc) return rc;
    }
    const double* top = nullptr;
    if (logM) {
      // the short products of the wrapped ones: top limbs of each operand
      hipLaunchKernelGGL(k_wrap_descs, dim3((k + 255) / 256), dim3(256), 0, c.stream, dd, (uint64_t)k, N,
                         c.wrap_desc);
      const Plan pm = make_plan(logM);
      const size_t mm = pm.passes.size();
      double* A2 = c.bspec + 4 * N;
      double* B2 = A2 + 2 * M;
      for (size_t j = 0; j < a_own_passes(pm); ++j) {
        rc = launch_pass(c, kFwd, pm, pm.passes[j], j == 0 ? nullptr : A2, nullptr, 0, A2, nullptr,
                         c.wrap_desc, 0, per, k);
        if (rc) return rc;
      }
      for (size_t j = 0; j + 1 < mm; ++j) {
        rc = launch_pass(c, kFwd, pm, pm.passes[j], j == 0 ? nullptr : B2, nullptr, 0, B2, nullptr,
                         c.wrap_desc, 1, per, k);
        if (rc) return rc;
      }
      rc = launch_pass(c, pm.big ? kFusedA : kFused, pm, pm.passes[mm - 1], mm == 1 ? nullptr : B2, nullptr,
                       0, B2, A2, c.wrap_desc, 1, per, k);
      if (rc) return rc;
      for (size_t j = mm - 1; j-- > 0;) {
        rc = launch_pass(c, kInv, pm, pm.passes[j], B2, nullptr, 0, B2, nullptr, c.wrap_desc, 1, per, k);
        if (rc) return rc;
      }
      top = B2;
    }
    if ((size_t)k * nch > c.bcarry_cap) return fail(PCG_ERR_INTERNAL, "batched carry blocks exceed the scratch");
    hipLaunchKernelGGL(k_batch_carry_local, dim3(nch, k), dim3(kCarryThreads), 0, c.stream, dd, B, per,
                       N, c.d_prime + logN * 2, c.crt, c.bagg, nch, top, M);
    if (nch > 64) {
      hipLaunchKernelGGL(k_batch_carry_scan_wide, dim3(k), dim3(256), 0, c.stream, dd, c.bagg, c.bcin, nch);
      hipLaunchKernelGGL(k_batch_carry_fix, dim3((unsigned)(((uint64_t)k * nch + 255) / 256)), dim3(256), 0,
                         c.stream, dd, c.bcin, nch, (uint64_t)k);
    } else {  // the scan adds the carry-ins itself
      hipLaunchKernelGGL(k_batch_carry_scan, dim3(k), dim3(64), 0, c.stream, dd, c.bagg, c.bcin, nch);
    }
  }
  HIPCHK(hipGetLastError());
  return PCG_OK;
}

// Launches one small class (0..3 schoolbook, 10..12 single-workgroup NTT,
// 110..112 batched passes)
// over n consecutive descriptors (already on the device).
int launch_small_class(Ctx& c, int cls, const BatchDesc* base, size_t n) {
  if (n == 0) return PCG_OK;
  switch (cls) {
    case 0:
      hipLaunchKernelGGL(k_batch_school, dim3((unsigned)n), dim3(256), 0, c.stream, base);
      break;
    case 1:
      hipLaunchKernelGGL((k_batch_school_g<16, 4>), dim3((unsigned)((n + 15) / 16)), dim3(256), 0, c.stream,
                         base, (uint32_t)n);
      break;
    case 2:
      hipLaunchKernelGGL((k_batch_school_g<64, 4>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0, c.stream,
                         base, (uint32_t)n);
      break;
    case 3:
      hipLaunchKernelGGL((k_batch_school_g<64, 8>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0, c.stream,
                         base, (uint32_t)n);
      break;
    case kKaraClass:
      hipLaunchKernelGGL(k_batch_kara, dim3((unsigned)n), dim3(256), 0, c.stream, base);
      break;
    case 10:
      hipLaunchKernelGGL(k_batch_ntt<10>, dim3((unsigned)n), dim3((2 << 10) / kEPT), 0, c.stream, base,
                         c.d_prime + 10 * 2, c.crt);
      break;
    case 11:
      hipLaunchKernelGGL(k_batch_ntt<11>, dim3((unsigned)n), dim3((2 << 11) / kEPT), 0, c.stream, base,
                         c.d_prime + 11 * 2, c.crt);
      break;
    case 12:
      hipLaunchKernelGGL(k_batch_ntt<12>, dim3((unsigned)n), dim3((2 << 12) / kEPT), 0, c.stream, base,
                         c.d_prime + 12 * 2, c.crt);
      break;
    default:
      if (cls >= kPassClass + kSmallNttMinLog && cls <= kPassClass + kSmallNttMaxLog)
        return batched_multipass(c, base, n, cls - kPassClass);
      return fail(PCG_ERR_INTERNAL, "bad small class %d", cls);
  }
  HIPCHK(hipGetLastError());
  return PCG_OK;
}

__global__ void k_fill_pattern(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)(i * 2654435761u) ^ 0x9e3779b9u;
}

// Autotuner of the batched pass geometry (batch_pass_geom): for each size
// 2^b, b = 13 .. 20, it times one batched launch of products of
// na = nb = 2^(b-1) - 1 limbs (2^23 operand limbs in all, at least 8
// products) on the 4096-point and on the large-tile passes (best of three
// launches after a first one, HIP events on the library stream) and keeps
// the faster.  Products are exact on both; only the time differs.
int tune_geom(Ctx& c) {
  for (int b = 0; b <= kGeomMaxLog; ++b) g_batch_geom[b] = batch_geom_default(b);
  constexpr size_t kLimbs = (size_t)1 << 23;
  constexpr size_t kMaxCount = kLimbs >> (kGeomMinLog + 1);
  uint32_t* in = nullptr;
  uint32_t* out = nullptr;
  BatchDesc* dd = nullptr;
  if (hipMalloc(&in, kLimbs * 4) != hipSuccess || hipMalloc(&out, kLimbs * 4) != hipSuccess ||
      hipMalloc(&dd, kMaxCount * sizeof(BatchDesc)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(in);
    (void)hipFree(out);
    return fail(PCG_ERR_OOM, "cannot allocate the geometry autotuner's test batch");
  }
  hipLaunchKernelGGL(k_fill_pattern, dim3((unsigned)((kLimbs + 255) / 256)), dim3(256), 0, c.stream, in,
                     (uint64_t)kLimbs);
  std::vector<BatchDesc> h(kMaxCount);
  int rc = PCG_OK;
  for (int b = kGeomMinLog; b <= kGeomMaxLog && rc == PCG_OK; ++b) {
    const uint32_t n = (1u << (b - 1)) - 1;
    const size_t count = std::min(kMaxCount, kLimbs / (2 * (size_t)n));
    for (size_t i = 0; i < count; ++i) {
      h[i].a = in + i * 2 * n;
      h[i].b = h[i].a + n;
      h[i].out = out + i * 2 * n;
      h[i].na = n;
      h[i].nb = n;
    }
    if (hipMemcpyAsync(dd, h.data(), count * sizeof(BatchDesc), hipMemcpyHostToDevice, c.stream) !=
        hipSuccess) {
      rc = fail(PCG_ERR_HIP, "geometry autotuner upload failed");
      break;
    }
    float best_ms = 1e30f;
    int best = batch_geom_default(b);
    for (int g = 0; g <= 1 && rc == PCG_OK; ++g) {
      g_batch_geom[b] = g;
      float t = 1e30f;
      for (int rep = 0; rep < 4 && rc == PCG_OK; ++rep) {
        if (hipEventRecord(c.tev[0], c.stream) != hipSuccess) rc = fail(PCG_ERR_HIP, "event record failed");
        if (!rc) rc = batched_multipass(c, dd, count, b);
        float ms = 0;
        if (!rc && (hipEventRecord(c.tev[1], c.stream) != hipSuccess ||
                    hipEventSynchronize(c.tev[1]) != hipSuccess ||
                    hipEventElapsedTime(&ms, c.tev[0], c.tev[1]) != hipSuccess))
          rc = fail(PCG_ERR_HIP, "geometry autotuner timing failed");
        if (!rc && rep > 0) t = std::min(t, ms);
      }
      if (!rc && t < best_ms) {
        best_ms = t;
        best = g;
      }
    }
    g_batch_geom[b] = best;
  }
  (void)hipStreamSynchronize(c.stream);
  (void)hipFree(in);
  (void)hipFree(out);
  (void)hipFree(dd);
  return rc;
}

// Autotuner of the small-product classes: for each bucket b = 4..12 it
// times a launch of 2048 to 16384 products (about the bucket's share of one
// level of a large subtree) of na = nb = 3 * 2^(b-3) limbs (zL = 3/4 of the
// bucket's top) through every class valid for the bucket (the schoolbook
// kernels, the single-workgroup NTT and the batched pass kernels at
// 2^max(b, 10) points; best of three launches, HIP events on the library
// stream) and keeps the fastest; smaller buckets take the smallest
// schoolbook.  Results are exact in every class: the choice only moves the
// crossovers.
int tune_small(Ctx& c) {
  for (int b = 0; b <= kSmallNttMaxLog; ++b) g_small_class[b] = small_class_default(b);
  // products per timed launch: about what one level of a 2^19-leaf subtree
  // holds in the bucket (4 d / 2^(b-3)), between 2048 and 16384
  auto count_of = [](int b) { return std::min(16384, std::max(2048, 1 << (23 - b))); };
  constexpr int kMaxCount = 16384;
  constexpr size_t kMaxLimbs = (size_t)2048 * 2 * (3 << 9);  // count x zL, equal for b >= 9
  uint32_t* in = nullptr;
  uint32_t* out = nullptr;
  BatchDesc* dd = nullptr;
  const size_t in_limbs = kMaxLimbs, out_limbs = kMaxLimbs;
  if (hipMalloc(&in, in_limbs * 4) != hipSuccess || hipMalloc(&out, out_limbs * 4) != hipSuccess ||
      hipMalloc(&dd, kMaxCount * sizeof(BatchDesc)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(in);
    (void)hipFree(out);
    return fail(PCG_ERR_OOM, "cannot allocate the autotuner's test batch");
  }
  hipLaunchKernelGGL(k_fill_pattern, dim3((unsigned)((in_limbs + 255) / 256)), dim3(256), 0, c.stream, in,
                     (uint64_t)in_limbs);
  std::vector<BatchDesc> h(kMaxCount);
  int rc = PCG_OK;
  for (int b = 4; b <= kSmallNttMaxLog && rc == PCG_OK; ++b) {
    const uint32_t n = 3u << (b - 3);
    const int kCount = count_of(b);
    if ((size_t)kCount * 2 * n > kMaxLimbs) {
      rc = fail(PCG_ERR_INTERNAL, "autotuner batch exceeds its buffers");
      break;
    }
    for (int i = 0; i < kCount; ++i) {
      h[i].a = in + (size_t)i * 2 * n;
      h[i].b = h[i].a + n;
      h[i].out = out + (size_t)i * 2 * n;
      h[i].na = n;
      h[i].nb = n;
    }
    if (hipMemcpyAsync(dd, h.data(), kCount * sizeof(BatchDesc), hipMemcpyHostToDevice, c.stream) !=
        hipSuccess) {
      rc = fail(PCG_ERR_HIP, "autotuner upload failed");
      break;
    }
    const int lg = b < kSmallNttMinLog ? kSmallNttMinLog : b;
    const int cls[7] = {1, 2, 3, 0, kKaraClass, lg, kPassClass + lg};
    // one launch of class m, timed
    auto time_class = [&](int m, float& ms) {
      if (hipEventRecord(c.tev[0], c.stream) != hipSuccess) return fail(PCG_ERR_HIP, "event record failed");
      int r = launch_small_class(c, cls[m], dd, kCount);
      if (!r && (hipEventRecord(c.tev[1], c.stream) != hipSuccess || hipEventSynchronize(c.tev[1]) != hipSuccess ||
                 hipEventElapsedTime(&ms, c.tev[0], c.tev[1]) != hipSuccess))
        r = fail(PCG_ERR_HIP, "autotuner timing failed");
      return r;
    };
    // every valid class: an untimed launch and a timed one; the two fastest
    // then get two more timed launches each and the best time decides
    float best[7];
    int order[7], nvalid = 0;
    for (int m = 0; m < 7 && rc == PCG_OK; ++m) {
      best[m] = 1e30f;
      if (!small_class_valid(b, cls[m])) continue;
      float ms = 0;
      rc = time_class(m, ms);
      if (!rc) rc = time_class(m, ms);
      if (rc) break;
      best[m] = ms;
      order[nvalid++] = m;
    }
    if (rc) break;
    std::sort(order, order + nvalid, [&](int x, int y) { return best[x] < best[y]; });
    for (int k = 0; k < std::min(nvalid, 2) && rc == PCG_OK; ++k)
      for (int rep = 0; rep < 2 && rc == PCG_OK; ++rep) {
        float ms = 0;
        rc = time_class(order[k], ms);
        if (!rc) best[order[k]] = std::min(best[order[k]], ms);
      }
    if (rc) break;
    int best_cls = small_class_default(b);
    float best_ms = 1e30f;
    for (int k = 0; k < nvalid; ++k)
      if (best[order[k]] < best_ms) {
        best_ms = best[order[k]];
        best_cls = cls[order[k]];
      }
    g_small_class[b] = best_cls;
  }
  (void)hipStreamSynchronize(c.stream);
  (void)hipFree(in);
  (void)hipFree(out);
  (void)hipFree(dd);
  if (rc) return rc;
  rc = tune_geom(c);
  if (rc) return rc;
  g_small_ready = true;
  ++g_table_gen;
  return PCG_OK;
}

int ensure_tuned(Ctx& c) { return g_small_ready ? PCG_OK : tune_small(c); }

int batch_locked(Ctx& c, const pcg_mul_desc* d, uint64_t count, uint32_t flags, bool dev) {
  // class slots: 0..3 schoolbook classes 0..3, 4..6 single-workgroup NTT
  // (2^10 .. 2^12), 7..9 batched passes at 2^10 .. 2^12, 10 Karatsuba,
  // 11..19 batched multi-pass (2^13 .. 2^21)
  constexpr int kSmallSlots = 11;
  constexpr int kSlots = kSmallSlots + (kBatchMaxLog - kSmallNttMaxLog);
  int rc0 = ensure_tuned(c);
  if (rc0) return rc0;
  std::vector<uint64_t> cls[kSlots], large;
  for (uint64_t i = 0; i < count; ++i) {
    int rc = check_sizes(d[i].na, d[i].nb);
    if (rc) return rc;
    if (d[i].na + d[i].nb == 0) continue;
    if ((d[i].na + d[i].nb) && !d[i].out) return fail(PCG_ERR_INVALID, "null output pointer");
    if (d[i].na == 0 || d[i].nb == 0) {
      if (dev) {
        const uint64_t zL = d[i].na + d[i].nb;
        hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((zL + 255) / 256)), dim3(256), 0, c.stream,
                           d[i].out, zL);
      } else {
        memset(d[i].out, 0, (d[i].na + d[i].nb) * 4);
      }
      continue;
    }
    if (!d[i].a || !d[i].b) return fail(PCG_ERR_INVALID, "null operand pointer");
    const int k = batch_class(d[i].na, d[i].nb);
    if (k < 0)
      large.push_back(i);
    else
      cls[k <= kSmallNttMaxLog || k >= kPassClass ? small_slot(k) : k - kSmallNttMaxLog - 1 + kSmallSlots]
          .push_back(i);
  }
  size_t nsmall = 0;
  for (int k = 0; k < kSlots; ++k) nsmall += cls[k].size();
  if (nsmall) {
    size_t in_limbs = 0, out_limbs = 0;
    for (int k = 0; k < kSlots; ++k)
      for (uint64_t i : cls[k]) {
        in_limbs += d[i].na + d[i].nb;
        out_limbs += d[i].na + d[i].nb;
      }
    int rc = ensure_pool(c, dev ? 0 : in_limbs + out_limbs, nsmall);
    if (rc) return rc;
    size_t pos = 0, opos = in_limbs, j = 0;
    for (int k = 0; k < kSlots; ++k)
      for (uint64_t i : cls[k]) {
        BatchDesc& bd = c.h_desc[j++];
        bd.na = (uint32_t)d[i].na;
        bd.nb = (uint32_t)d[i].nb;
        if (dev) {
          bd.a = d[i].a;
          bd.b = d[i].b;
          bd.out = d[i].out;
        } else {
          memcpy(c.h_pool + pos, d[i].a, d[i].na * 4);
          bd.a = c.d_pool + pos;
          pos += d[i].na;
          memcpy(c.h_pool + pos, d[i].b, d[i].nb * 4);
          bd.b = c.d_pool + pos;
          pos += d[i].nb;
          bd.out = c.d_pool + opos;
          opos += d[i].na + d[i].nb;
        }
      }
    if (!dev)
      HIPCHK(hipMemcpyAsync(c.d_pool, c.h_pool, in_limbs * 4, hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(c.d_desc, c.h_desc, nsmall * sizeof(BatchDesc), hipMemcpyHostToDevice,
                          c.stream));
    size_t off = 0;
    for (int k = 0; k < kSmallSlots; ++k) {
      rc = launch_small_class(c, small_slot_class(k), c.d_desc + off, cls[k].size());
      if (rc) return rc;
      off += cls[k].size();
    }
    for (int k = kSmallSlots; k < kSlots; ++k) {
      if (cls[k].empty()) continue;
      rc = batched_multipass(c, c.d_desc + off, cls[k].size(), kSmallNttMaxLog + 1 + k - kSmallSlots);
      if (rc) return rc;
      off += cls[k].size();
    }
    if (!dev)
      HIPCHK(hipMemcpyAsync(c.h_pool + in_limbs, c.d_pool + in_limbs, out_limbs * 4,
                            hipMemcpyDeviceToHost, c.stream));
    // descriptors and pools are reused by the next call: finish before that
    HIPCHK(hipStreamSynchronize(c.stream));
    if (!dev) {
      size_t o = in_limbs;
      for (int k = 0; k < kSlots; ++k)
        for (uint64_t i : cls[k]) {
          const size_t zL = d[i].na + d[i].nb;
          memcpy(d[i].out, c.h_pool + o, zL * 4);
          o += zL;
        }
    }
  }
  for (uint64_t i : large) {
    int rc;
    if (dev) {
      rc = mul_dev_locked(c, d[i].a, d[i].na, d[i].b, d[i].nb, d[i].out, flags & ~PCG_PROFILE);
    } else {
      const uint64_t na = d[i].na, nb = d[i].nb, zL = na + nb;
      rc = ensure_stage(c, 2 * zL);
      if (rc) return rc;
      HIPCHK(hipMemcpyAsync(c.d_stage, d[i].a, na * 4, hipMemcpyHostToDevice, c.stream));
      HIPCHK(hipMemcpyAsync(c.d_stage + na, d[i].b, nb * 4, hipMemcpyHostToDevice, c.stream));
      rc = mul_dev_locked(c, c.d_stage, na, c.d_stage + na, nb, c.d_stage + zL, flags & ~PCG_PROFILE);
      if (rc) return rc;
      HIPCHK(hipMemcpyAsync(d[i].out, c.d_stage + zL, zL * 4, hipMemcpyDeviceToHost, c.stream));
      HIPCHK(hipStreamSynchronize(c.stream));
    }
    if (rc) return rc;
  }
  return PCG_OK;
}


// ------------------------------------------------------------------------
// Device-resident binary splitting (see bsr_kernels.hpp)
struct BsrNode {
  uint32_t a, b;
  int32_t left, right, height;
};

struct BsrMulPair {
  BsrMul m[2];  // same operand a
};

int32_t bsr_build(std::vector<BsrNode>& nodes, uint32_t a, uint32_t b) {
  const int32_t id = (int32_t)nodes.size();
  nodes.push_back(BsrNode{a, b, -1, -1, 1});
  if (b - a > 1) {
    const uint32_t m = (a + b) / 2;  // PI.h:90
    const int32_t l = bsr_build(nodes, a, m);
    const int32_t r = bsr_build(nodes, m, b);
    nodes[id].left = l;
    nodes[id].right = r;
    nodes[id].height = std::max(nodes[l].height, nodes[r].height) + 1;
  }
  return id;
}

// Bit bounds of a leaf's P, Q, R for any b < 2^32 (PI.h:67-88, with the
// factors 2b-1, 6b-5, 6b-1 in uint32 as the reference computes them):
// |P| < (13591409 + 545140134 b) 2^96 < 2^158, Q = b^3 26726400 409297880
// < 2^150, R < 2^96.
constexpr uint64_t kLeafBits[3] = {158, 150, 96};
using LeafBits = std::array<uint64_t, 3>;

// Bit lengths of |P|, Q, R of the leaf b = bmax, which bound those of every
// leaf b <= bmax while no factor wraps (6 bmax - 1 < 2^32: the values then
// grow with b); kLeafBits beyond.
LeafBits leaf_bits_for(uint32_t bmax) {
  if (bmax == 0 || bmax > 715827882u) return LeafBits{kLeafBits[0], kLeafBits[1], kLeafBits[2]};
  auto bits_of = [](uint64_t lead, std::initializer_list<uint32_t> fs) {
    uint32_t z[8] = {(uint32_t)lead, (uint32_t)(lead >> 32), 0, 0, 0, 0, 0, 0};
    int n = 2;
    for (uint32_t f : fs) {
      uint64_t c = 0;
      for (int i = 0; i < n; ++i) {
        c += (uint64_t)z[i] * f;
        z[i] = (uint32_t)c;
        c >>= 32;
      }
      z[n++] = (uint32_t)c;
    }
    int top = n - 1;
    while (top > 0 && z[top] == 0) --top;
    uint64_t b = 32 * (uint64_t)top;
    for (uint32_t v = z[top]; v; v >>= 1) ++b;
    return b;
  };
  const uint64_t b = bmax;
  const uint32_t f1 = 2 * bmax - 1, f2 = 6 * bmax - 5, f3 = 6 * bmax - 1;
  return LeafBits{bits_of(13591409ull + 545140134ull * b, {f1, f2, f3}),
                  bits_of(b * b, {bmax, 26726400u, 409297880u}), bits_of(f1, {f2, f3})};
}

uint64_t limbs_of_bits(uint64_t bits) { return (bits + 31) / 32; }

// Bounds of every value (P Q R x y per node) of a subtree, bottom-up:
// bits (x = P_L Q_R and y = P_R R_L add their operands' bits, P = x + y one
// more, Q and R add), tlen = limbs of the finished value (ceil(bits / 32):
// a product's one-limb trim and add/sub's carry limb and strip leave exactly
// the value's length), slot = limbs written before that (a product's
// na + nb, an addition's longer operand + 1).  Returns the largest tlen.
uint64_t bsr_bounds(const std::vector<BsrNode>& nodes, const LeafBits& leaf, std::vector<uint64_t>& tlen,
                    std::vector<uint64_t>& slot) {
  std::vector<uint64_t> bits(nodes.size() * 5, 0);
  tlen.assign(nodes.size() * 5, 0);
  slot.assign(nodes.size() * 5, 0);
  uint64_t mx = 0;
  for (size_t i = nodes.size(); i-- > 0;) {
    const BsrNode& n = nodes[i];
    const size_t v = i * 5;  // P Q R x y
    if (n.left < 0) {
      for (int j = 0; j < 3; ++j) bits[v + j] = leaf[j];
    } else {
      const size_t l = (size_t)n.left * 5, r = (size_t)n.right * 5;
      bits[v + 3] = bits[l + 0] + bits[r + 1];
      bits[v + 4] = bits[r + 0] + bits[l + 2];
      bits[v + 0] = std::max(bits[v + 3], bits[v + 4]) + 1;
      bits[v + 1] = bits[l + 1] + bits[r + 1];
      bits[v + 2] = bits[l + 2] + bits[r + 2];
    }
    for (int j = 0; j < 5; ++j) tlen[v + j] = limbs_of_bits(bits[v + j]);
    if (n.left < 0) {
      for (int j = 0; j < 3; ++j) slot[v + j] = tlen[v + j];
    } else {
      const size_t l = (size_t)n.left * 5, r = (size_t)n.right * 5;
      slot[v + 3] = tlen[l + 0] + tlen[r + 1];
      slot[v + 4] = tlen[r + 0] + tlen[l + 2];
      slot[v + 0] = std::max(tlen[v + 3], tlen[v + 4]) + 1;
      slot[v + 1] = tlen[l + 1] + tlen[r + 1];
      slot[v + 2] = tlen[l + 2] + tlen[r + 2];
    }
    for (int j = 0; j < 5; ++j) mx = std::max(mx, tlen[v + j]);
  }
  return mx;
}

template <class T>
T* carve(char*& p, size_t n) {
  T* r = reinterpret_cast<T*>(p);
  p += ((n * sizeof(T) + 255) / 256) * 256;
  return r;
}

// Plan of a subtree of d leaves: tree shape, value bounds and arena offsets
// (they depend on d only, since m - a = floor(d / 2)), per-height product
// lists sorted by size class, and the device copies of all of it.
struct BsrHeight {
  size_t mul0, nmul, add0, nsmall, nlarge;
  uint32_t add_maxch;  // 1024-limb chunks of the largest large-add result
  struct Launch {
    int cls;
    size_t n;    // products
    bool pairs;  // products 2k, 2k+1 share operand a (one forward transform)
    int wrap;    // > 0: wrapped at 2^cls points, top products of 2^wrap
  };
  std::vector<Launch> classes;  // in launch order
};
struct BsrPlan {
  uint32_t d = 0;
  LeafBits leaf{};  // leaf bounds the plan was sized with
  uint64_t maxb = 0, arena_limbs = 0;
  uint64_t root_bound[3] = {0, 0, 0};
  uint64_t root_off[3] = {0, 0, 0};
  int32_t H = 0;
  std::vector<BsrHeight> hp;
  uint64_t nleaves = 0;
  // device (buffers kept for the next plan laid out in this slot)
  void* buf = nullptr;
  uint32_t* arena = nullptr;
  size_t buf_cap = 0;       // bytes
  uint64_t arena_cap = 0;   // limbs
  BsrVal* meta = nullptr;
  uint32_t* leaf_rel = nullptr;  // b - a of each leaf
  uint32_t* leaf_ids = nullptr;
  BsrMul* muls = nullptr;
  BsrAdd* adds = nullptr;
  BatchDesc* descs = nullptr;
  BsrAddSlot* add_slots = nullptr;  // per large add of a height
  size_t nslots = 0;
  CarryFn* add_agg = nullptr;       // per (large add, chunk)
  uint64_t* add_cin = nullptr;
  uint32_t* a_dev = nullptr;  // the subtree's first index a (read by k_bsr_base)
  // the level schedule captured as a graph (replayed while the plan and the
  // batched scratch it names stay the same)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t graph_gen = 0;  // Ctx::scratch_gen at capture
};

void bsr_plan_drop_graph(BsrPlan& P) {
  if (P.exec) (void)hipGraphExecDestroy(P.exec);
  if (P.graph) (void)hipGraphDestroy(P.graph);
  P.exec = nullptr;
  P.graph = nullptr;
}

void bsr_plan_free(BsrPlan& P) {
  bsr_plan_drop_graph(P);
  (void)hipFree(P.buf);
  (void)hipFree(P.arena);
  P = BsrPlan();
}

// A subtree plan built node by node on the host (CPU only): the tree of d
// leaves (PI.h:90 midpoints), value bounds, arena offsets, and per height
// the products (grouped by class with the given small-class table, paired
// around shared operands) and the adds.  Plans are laid out on the device
// (bsr_plan_build); this direct construction is what pcg_bsr_plan_check
// compares that layout with.
struct BsrHostPlan {
  BsrPlan P;  // sizes and heights; no device memory
  std::vector<BsrVal> meta;
  std::vector<uint32_t> leaf_rel, leaf_ids;
  std::vector<BsrMul> muls;
  std::vector<BsrAdd> adds;
  size_t max_large = 0, max_large_chunks = 0;
};

using SmallTable = std::array<int, kSmallNttMaxLog + 1>;

void bsr_host_plan(uint32_t d, const LeafBits& leaf, const SmallTable& tab, BsrHostPlan& HP) {
  BsrPlan& P = HP.P;
  std::vector<BsrNode> nodes;
  nodes.reserve(2 * (size_t)d);
  bsr_build(nodes, 0, d);
  std::vector<uint64_t> tlen, slot;
  P.d = d;
  P.maxb = bsr_bounds(nodes, leaf, tlen, slot);
  const size_t nv = nodes.size() * 5;
  std::vector<BsrVal>& meta = HP.meta;
  meta.resize(nv);
  uint64_t off = 0;
  for (size_t v = 0; v < nv; ++v) {
    meta[v].off = off;
    meta[v].L = 0;
    meta[v].sign = 1;
    off += slot[v] + 1;  // one spare limb
  }
  P.arena_limbs = off;
  for (int j = 0; j < 3; ++j) {
    P.root_bound[j] = tlen[j];
    P.root_off[j] = meta[j].off;
  }
  P.H = nodes[0].height;
  std::vector<uint32_t>& leaf_rel = HP.leaf_rel;
  std::vector<uint32_t>& leaf_ids = HP.leaf_ids;
  std::vector<std::vector<int32_t>> by_h(P.H + 1);
  std::vector<int32_t> parent(nodes.size(), -1);
  for (int32_t i = 0; i < (int32_t)nodes.size(); ++i)
    if (nodes[i].left >= 0) parent[nodes[i].left] = parent[nodes[i].right] = i;
  for (int32_t i = 0; i < (int32_t)nodes.size(); ++i) {
    by_h[nodes[i].height].push_back(i);
    if (nodes[i].left < 0) {
      // k_bsr_base entries: a node of height 3 through its first leaf, a
      // two-leaf node outside one through its left leaf, a lone root leaf
      const int32_t q = parent[i];
      const int32_t g = q >= 0 ? parent[q] : -1;
      const uint32_t qs = q >= 0 ? nodes[q].b - nodes[q].a : 0, gs = g >= 0 ? nodes[g].b - nodes[g].a : 0;
      int32_t owner = i;  // node whose values the entry computes, -1: none
      uint32_t rel = nodes[i].b;
      if (qs == 3) {
        owner = q;
        rel |= kBsrQuadBit;
      } else if (qs == 2 && (gs == 3 || gs == 4)) {
        owner = nodes[i].a == nodes[g].a ? g : -1;
        rel |= kBsrQuadBit | (gs == 4 ? kBsrFourBit : 0u);
      } else if (qs == 2) {
        owner = nodes[q].left == i ? q : -1;
        rel |= kBsrPairBit;
      }
      leaf_rel.push_back(owner < 0 ? nodes[i].b : rel);
      for (int j = 0; j < 3; ++j) leaf_ids.push_back(owner < 0 ? kBsrSkip : (uint32_t)(5 * owner + j));
    }
  }
  P.nleaves = leaf_rel.size();
  P.hp.assign(P.H + 1, BsrHeight());
  std::vector<BsrMul>& muls = HP.muls;
  std::vector<BsrAdd>& adds = HP.adds;
  size_t& max_large = HP.max_large;
  size_t& max_large_chunks = HP.max_large_chunks;
  for (int32_t h = 2; h <= P.H; ++h) {
    std::vector<std::pair<int, BsrMul>> ms;
    std::vector<std::pair<int, BsrMulPair>> mp;
    ms.reserve(4 * by_h[h].size());
    mp.reserve(2 * by_h[h].size());
    std::vector<BsrAdd> small, large;
    uint64_t large_max = 0;
    for (int32_t i : by_h[h]) {
      const BsrNode& n = nodes[i];
      const uint32_t I = 5 * i, Lf = 5 * n.left, Rt = 5 * n.right;
      // x = P0 Q1, Q = Q0 Q1, y = P1 R0, R = R0 R1 (PI.h:110-112), as two
      // pairs around the operands used twice (products of exact integers
      // commute, so the shared one is operand a of both)
      const BsrMul m4[4] = {{Rt + 1, Lf + 0, I + 3}, {Rt + 1, Lf + 1, I + 1},
                            {Lf + 2, Rt + 0, I + 4}, {Lf + 2, Rt + 2, I + 2}};
      for (int q = 0; q < 4; q += 2) {
        const int c0 = bsr_key(tab, tlen[m4[q].a], tlen[m4[q].b]);
        const int c1 = bsr_key(tab, tlen[m4[q + 1].a], tlen[m4[q + 1].b]);
        if (c0 == c1 && class_pairs(c0 / kWrapKey)) {
          mp.push_back({c0, BsrMulPair{m4[q], m4[q + 1]}});
        } else {
          ms.push_back({c0, m4[q]});
          ms.push_back({c1, m4[q + 1]});
        }
      }
      const BsrAdd ad{I + 3, I + 4, I + 0};  // P = x.add(y)
      if (slot[I] <= 1024) {
        small.push_back(ad);
      } else {
        large.push_back(ad);
        large_max = std::max(large_max, slot[I]);
      }
    }
    // group the height's products by class (a height has only a few
    // distinct classes): one counting pass per class, in class order, the
    // paired products of a class before its single ones
    std::vector<int> kinds;
    for (const auto& e : ms)
      if (std::find(kinds.begin(), kinds.end(), e.first) == kinds.end()) kinds.push_back(e.first);
    for (const auto& e : mp)
      if (std::find(kinds.begin(), kinds.end(), e.first) == kinds.end()) kinds.push_back(e.first);
    std::sort(kinds.begin(), kinds.end());
    BsrHeight& hh = P.hp[h];
    hh.mul0 = muls.size();
    hh.nmul = ms.size() + 2 * mp.size();
    for (int kind : kinds) {
      size_t cnt = 0;
      for (const auto& e : mp)
        if (e.first == kind) {
          muls.push_back(e.second.m[0]);
          muls.push_back(e.second.m[1]);
          cnt += 2;
        }
      if (cnt) hh.classes.push_back({kind / kWrapKey, cnt, true, kind % kWrapKey});
      cnt = 0;
      for (const auto& e : ms)
        if (e.first == kind) {
          muls.push_back(e.second);
          ++cnt;
        }
      if (cnt) hh.classes.push_back({kind / kWrapKey, cnt, false, kind % kWrapKey});
    }
    hh.add0 = adds.size();
    hh.nsmall = small.size();
    hh.nlarge = large.size();
    hh.add_maxch = (uint32_t)((large_max + kCarryBlock - 1) / kCarryBlock);
    max_large = std::max(max_large, hh.nlarge);
    max_large_chunks = std::max(max_large_chunks, hh.nlarge * hh.add_maxch);
    adds.insert(adds.end(), small.begin(), small.end());
    adds.insert(adds.end(), large.begin(), large.end());
  }
}

// Size table of a subtree plan (bsr_layout_node): one entry per distinct
// subtree size, children before parents, with the value bounds of
// bsr_bounds per size, and per height the product groups and adds.
struct BsrSizes {
  LeafBits leaf{};
  std::vector<BsrSizeInfo> info;
  std::vector<std::array<uint64_t, 5>> bits, tlen, slot;
  std::vector<int32_t> height;
  std::vector<uint32_t> cnt;  // ns x ns subtree counts
  std::map<uint64_t, int32_t> index;
  int32_t root = -1;
  size_t max_large = 0, max_large_chunks = 0;
};

int32_t bsr_size_index(BsrSizes& S, uint64_t s) {
  auto it = S.index.find(s);
  if (it != S.index.end()) return it->second;
  int32_t l = -1, r = -1;
  if (s > 1) {
    l = bsr_size_index(S, s / 2);
    r = bsr_size_index(S, s - s / 2);
  }
  std::array<uint64_t, 5> bits{}, tl{}, sl{};
  if (l < 0) {
    for (int j = 0; j < 3; ++j) bits[j] = S.leaf[j];
  } else {
    const auto &bl = S.bits[l], &br = S.bits[r];
    bits[3] = bl[0] + br[1];
    bits[4] = br[0] + bl[2];
    bits[0] = std::max(bits[3], bits[4]) + 1;
    bits[1] = bl[1] + br[1];
    bits[2] = bl[2] + br[2];
  }
  for (int j = 0; j < 5; ++j) tl[j] = limbs_of_bits(bits[j]);
  if (l < 0) {
    for (int j = 0; j < 3; ++j) sl[j] = tl[j];
  } else {
    const auto &tll = S.tlen[l], &tlr = S.tlen[r];
    sl[3] = tll[0] + tlr[1];
    sl[4] = tlr[0] + tll[2];
    sl[0] = std::max(tl[3], tl[4]) + 1;
    sl[1] = tll[1] + tlr[1];
    sl[2] = tll[2] + tlr[2];
  }
  BsrSizeInfo e;
  memset(&e, 0, sizeof e);
  e.s = s;
  e.l = l;
  e.r = r;
  for (int j = 0; j < 5; ++j) {
    e.pre[j] = e.own;
    e.own += sl[j] + 1;  // one spare limb
  }
  e.arena = e.own + (l < 0 ? 0 : S.info[l].arena + S.info[r].arena);
  const int32_t i = (int32_t)S.info.size();
  S.info.push_back(e);
  S.bits.push_back(bits);
  S.tlen.push_back(tl);
  S.slot.push_back(sl);
  S.height.push_back(l < 0 ? 1 : std::max(S.height[l], S.height[r]) + 1);
  S.index[s] = i;
  return i;
}

// The plan of a subtree of d leaves from its size table: P's sizes and
// heights (launch groups, adds) and S, whose info and cnt are what
// bsr_layout_node reads.  Product groups follow bsr_key and class_pairs
// with the class table tab, in the order of the direct construction.
int bsr_plan_sizes(uint32_t d, const LeafBits& leaf, const SmallTable& tab, BsrPlan& P, BsrSizes& S) {
  P = BsrPlan();
  S = BsrSizes();
  if (d == 0) return fail(PCG_ERR_INVALID, "empty range");
  if (d > kBsrRelMask) return fail(PCG_ERR_TOO_LARGE, "subtree of %u leaves", d);
  S.leaf = leaf;
  P.leaf = leaf;
  S.root = bsr_size_index(S, d);
  const size_t ns = S.info.size();
  S.cnt.assign(ns * ns, 0);
  for (size_t i = 0; i < ns; ++i) {  // children have smaller indices
    uint32_t* ci = &S.cnt[i * ns];
    ci[i] += 1;
    if (S.info[i].l >= 0)
      for (size_t q = 0; q < ns; ++q)
        ci[q] += S.cnt[S.info[i].l * ns + q] + S.cnt[S.info[i].r * ns + q];
  }
  const uint32_t* croot = &S.cnt[(size_t)S.root * ns];
  P.d = d;
  P.maxb = 0;
  for (size_t i = 0; i < ns; ++i)
    for (int j = 0; j < 5; ++j) P.maxb = std::max(P.maxb, S.tlen[i][j]);
  P.arena_limbs = S.info[S.root].arena;
  for (int j = 0; j < 3; ++j) {
    P.root_bound[j] = S.tlen[S.root][j];
    P.root_off[j] = S.info[S.root].pre[j];
  }
  P.H = S.height[S.root];
  P.nleaves = d;
  P.hp.assign(P.H + 1, BsrHeight());
  uint64_t mcur = 0, acur = 0;
  for (int32_t h = 2; h <= P.H; ++h) {
    std::vector<int32_t> peers;
    for (size_t i = 0; i < ns; ++i)
      if (S.height[i] == h) peers.push_back((int32_t)i);
    if (peers.size() > (size_t)kBsrPeers)
      return fail(PCG_ERR_INTERNAL, "%zu subtree sizes at height %d", peers.size(), (int)h);
    // group code of each product: kind * 2, + 1 for a single (unpaired) one
    std::vector<std::array<int, 4>> code(peers.size());
    for (size_t p = 0; p < peers.size(); ++p) {
      const BsrSizeInfo& e = S.info[peers[p]];
      const auto &tl = S.tlen[e.l], &tr = S.tlen[e.r];
      // operands as bsr_layout_node's m4: (Q1, P0), (Q1, Q0), (R0, P1), (R0, R1)
      const int k[4] = {bsr_key(tab, tr[1], tl[0]), bsr_key(tab, tr[1], tl[1]),
                        bsr_key(tab, tl[2], tr[0]), bsr_key(tab, tl[2], tr[2])};
      for (int q = 0; q < 4; q += 2) {
        const bool paired = k[q] == k[q + 1] && class_pairs(k[q] / kWrapKey);
        code[p][q] = 2 * k[q] + (paired ? 0 : 1);
        code[p][q + 1] = 2 * k[q + 1] + (paired ? 0 : 1);
      }
    }
    std::vector<int> groups;
    for (const auto& cp : code)
      for (int c : cp)
        if (std::find(groups.begin(), groups.end(), c) == groups.end()) groups.push_back(c);
    std::sort(groups.begin(), groups.end());
    BsrHeight& hh = P.hp[h];
    hh.mul0 = mcur;
    std::map<int, uint64_t> gbase;
    for (int g : groups) {
      uint64_t n = 0;
      for (size_t p = 0; p < peers.size(); ++p)
        for (int j = 0; j < 4; ++j)
          if (code[p][j] == g) n += croot[peers[p]];
      gbase[g] = mcur;
      hh.classes.push_back({(g / 2) / kWrapKey, (size_t)n, (g & 1) == 0, (g / 2) % kWrapKey});
      mcur += n;
    }
    hh.nmul = mcur - hh.mul0;
    // adds: those with a result slot of at most 1024 limbs first
    hh.add0 = acur;
    hh.nsmall = hh.nlarge = 0;
    uint64_t large_max = 0;
    for (int32_t q : peers) {
      if (S.slot[q][0] <= 1024) {
        hh.nsmall += croot[q];
      } else {
        hh.nlarge += croot[q];
        large_max = std::max(large_max, S.slot[q][0]);
      }
    }
    hh.add_maxch = (uint32_t)((large_max + kCarryBlock - 1) / kCarryBlock);
    S.max_large = std::max(S.max_large, hh.nlarge);
    S.max_large_chunks = std::max(S.max_large_chunks, hh.nlarge * hh.add_maxch);
    acur += hh.nsmall + hh.nlarge;
    for (size_t p = 0; p < peers.size(); ++p) {
      BsrSizeInfo& e = S.info[peers[p]];
      e.npeer = (int32_t)peers.size();
      for (size_t o = 0; o < peers.size(); ++o) e.peer[o] = peers[o];
      for (int j = 0; j < 4; ++j) {
        e.mbase[j] = gbase[code[p][j]];
        e.mloc[j] = 0;
        for (int i = 0; i < j; ++i) e.mloc[j] += code[p][i] == code[p][j];
        for (size_t o = 0; o < peers.size(); ++o) {
          e.mmul[j][o] = 0;
          for (int i = 0; i < 4; ++i) e.mmul[j][o] += code[o][i] == code[p][j];
        }
      }
      const bool small = S.slot[peers[p]][0] <= 1024;
      e.abase = hh.add0 + (small ? 0 : hh.nsmall);
      for (size_t o = 0; o < peers.size(); ++o) e.amul[o] = (S.slot[peers[o]][0] <= 1024) == small;
    }
  }
  if (mcur != 4 * (uint64_t)(d - 1) || acur != (uint64_t)(d - 1))
    return fail(PCG_ERR_INTERNAL, "subtree plan of %u leaves: %llu products, %llu adds", d,
                (unsigned long long)mcur, (unsigned long long)acur);
  return PCG_OK;
}

// The largest value length (tlen) of a subtree of d leaves whose leaves are
// bounded by leaf: over the size table, O(log d) sizes.
uint64_t bsr_max_limbs(uint32_t d, const LeafBits& leaf) {
  BsrSizes S;
  S.leaf = leaf;
  bsr_size_index(S, d);
  uint64_t mx = 0;
  for (const auto& t : S.tlen)
    for (uint64_t v : t) mx = std::max(mx, v);
  return mx;
}

BsrLayout bsr_layout_of(const BsrSizes& S, uint32_t d) {
  BsrLayout L;
  memset(&L, 0, sizeof L);
  L.ns = (int32_t)S.info.size();
  L.root = S.root;
  L.nodes = 2 * (uint64_t)d - 1;
  L.nmeta = 5 * L.nodes;
  L.nleaf = d;
  L.nmul = 4 * (uint64_t)(d - 1);
  L.nadd = d - 1;
  return L;
}

// Device side: lays the plan out with k_bsr_layout (one thread per node)
// into the slot's buffers, grown when too small; synchronises to check it.
int bsr_plan_build(Ctx& c, uint32_t d, const LeafBits& leaf, const SmallTable& tab, BsrPlan& P) {
  BsrSizes S;
  BsrPlan NP;
  int rc = bsr_plan_sizes(d, leaf, tab, NP, S);
  if (rc) return rc;
  BsrLayout L = bsr_layout_of(S, d);
  const size_t ns = S.info.size();
  const size_t need = L.nmeta * sizeof(BsrVal) + L.nleaf * 4 * 4 + L.nmul * sizeof(BsrMul) +
                      L.nadd * sizeof(BsrAdd) + L.nmul * sizeof(BatchDesc) +
                      S.max_large * sizeof(BsrAddSlot) +
                      S.max_large_chunks * (sizeof(CarryFn) + sizeof(uint64_t)) +
                      ns * sizeof(BsrSizeInfo) + ns * ns * 4 + 4 + 14 * 256;
  HIPCHK(hipStreamSynchronize(c.stream));  // the slot's old plan: last use
  bsr_plan_drop_graph(P);
  NP.buf = P.buf;
  NP.arena = P.arena;
  NP.buf_cap = P.buf_cap;
  NP.arena_cap = P.arena_cap;
  P = BsrPlan();  // the slot holds no plan until this one is laid out
  bool ok = true;
  if (NP.buf_cap < need) {
    (void)hipFree(NP.buf);
    NP.buf = nullptr;
    NP.buf_cap = 0;
    ok = hipMalloc(&NP.buf, need) == hipSuccess;
    if (ok) NP.buf_cap = need;
  }
  if (ok && NP.arena_cap < NP.arena_limbs) {
    (void)hipFree(NP.arena);
    NP.arena = nullptr;
    NP.arena_cap = 0;
    ok = hipMalloc(&NP.arena, NP.arena_limbs * 4) == hipSuccess;
    if (ok) NP.arena_cap = NP.arena_limbs;
  }
  if (!ok) {
    (void)hipGetLastError();
    const uint64_t al = NP.arena_limbs;
    bsr_plan_free(NP);
    return fail(PCG_ERR_OOM, "cannot allocate the binary-splitting plan (%llu arena limbs)",
                (unsigned long long)al);
  }
  P = std::move(NP);
  char* p = static_cast<char*>(P.buf);
  P.meta = carve<BsrVal>(p, L.nmeta);
  P.leaf_rel = carve<uint32_t>(p, L.nleaf);
  P.leaf_ids = carve<uint32_t>(p, 3 * L.nleaf);
  P.muls = carve<BsrMul>(p, L.nmul);
  P.adds = carve<BsrAdd>(p, L.nadd);
  P.descs = carve<BatchDesc>(p, L.nmul);
  P.add_slots = carve<BsrAddSlot>(p, S.max_large);
  P.add_agg = carve<CarryFn>(p, S.max_large_chunks);
  P.add_cin = carve<uint64_t>(p, S.max_large_chunks);
  P.a_dev = carve<uint32_t>(p, 1);
  BsrSizeInfo* d_info = carve<BsrSizeInfo>(p, ns);
  uint32_t* d_cnt = carve<uint32_t>(p, ns * ns);
  uint32_t* d_err = carve<uint32_t>(p, 1);
  L.info = d_info;
  L.cnt = d_cnt;
  L.meta = P.meta;
  L.leaf_rel = P.leaf_rel;
  L.leaf_ids = P.leaf_ids;
  L.muls = P.muls;
  L.adds = P.adds;
  L.err = d_err;
  HIPCHK(hipMemcpyAsync(d_info, S.info.data(), ns * sizeof(BsrSizeInfo), hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipMemcpyAsync(d_cnt, S.cnt.data(), ns * ns * 4, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipMemsetAsync(d_err, 0, 4, c.stream));
  P.nslots = S.max_large;
  const uint64_t blocks = std::min<uint64_t>((L.nodes + 255) / 256, 8192);
  hipLaunchKernelGGL(k_bsr_layout, dim3((unsigned)blocks), dim3(256), 0, c.stream, L);
  HIPCHK(hipGetLastError());
  uint32_t err = 0;
  HIPCHK(hipMemcpyAsync(&err, d_err, 4, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));  // S's tables are freed on return
  if (err) {
    bsr_plan_free(P);
    return fail(PCG_ERR_INTERNAL, "subtree plan of %u leaves: an entry fell outside its list", d);
  }
  return PCG_OK;
}

// The layout of bsr_plan_build run on the host, compared entry by entry with
// the direct construction (bsr_host_plan).  Host only.
int bsr_plan_check(uint32_t d, const LeafBits& leaf, const SmallTable& tab) {
  BsrPlan P;
  BsrSizes S;
  int rc = bsr_plan_sizes(d, leaf, tab, P, S);
  if (rc) return rc;
  BsrLayout L = bsr_layout_of(S, d);
  std::vector<BsrVal> meta(L.nmeta);
  std::vector<uint32_t> leaf_rel(L.nleaf), leaf_ids(3 * L.nleaf);
  std::vector<BsrMul> muls(L.nmul);
  std::vector<BsrAdd> adds(L.nadd);
  L.info = S.info.data();
  L.cnt = S.cnt.data();
  L.meta = meta.data();
  L.leaf_rel = leaf_rel.data();
  L.leaf_ids = leaf_ids.data();
  L.muls = muls.data();
  L.adds = adds.data();
  for (uint64_t t = 0; t < L.nodes; ++t)
    if (!bsr_layout_node(L, t)) return fail(PCG_ERR_INTERNAL, "d %u: node %llu outside its lists", d, (unsigned long long)t);
  BsrHostPlan H;
  bsr_host_plan(d, leaf, tab, H);
  const BsrPlan& Q = H.P;
  auto differ = [&](const char* what) { return fail(PCG_ERR_INTERNAL, "d %u: %s differs", d, what); };
  if (P.maxb != Q.maxb || P.arena_limbs != Q.arena_limbs || P.H != Q.H || P.nleaves != Q.nleaves)
    return differ("plan size");
  for (int j = 0; j < 3; ++j)
    if (P.root_bound[j] != Q.root_bound[j] || P.root_off[j] != Q.root_off[j]) return differ("root");
  if (S.max_large != H.max_large || S.max_large_chunks != H.max_large_chunks) return differ("add scratch");
  for (int32_t h = 2; h <= P.H; ++h) {
    const BsrHeight &x = P.hp[h], &y = Q.hp[h];
    if (x.mul0 != y.mul0 || x.nmul != y.nmul || x.add0 != y.add0 || x.nsmall != y.nsmall ||
        x.nlarge != y.nlarge || x.add_maxch != y.add_maxch || x.classes.size() != y.classes.size())
      return differ("height");
    for (size_t k = 0; k < x.classes.size(); ++k) {
      const auto &u = x.classes[k], &v = y.classes[k];
      if (u.cls != v.cls || u.n != v.n || u.pairs != v.pairs || u.wrap != v.wrap) return differ("launch group");
    }
  }
  if (meta.size() != H.meta.size() || muls.size() != H.muls.size() || adds.size() != H.adds.size() ||
      leaf_rel != H.leaf_rel || leaf_ids != H.leaf_ids)
    return differ("leaves or list sizes");
  for (size_t v = 0; v < meta.size(); ++v)
    if (meta[v].off != H.meta[v].off || meta[v].L != H.meta[v].L || meta[v].sign != H.meta[v].sign)
      return differ("arena offset");
  for (size_t k = 0; k < muls.size(); ++k)
    if (muls[k].a != H.muls[k].a || muls[k].b != H.muls[k].b || muls[k].out != H.muls[k].out)
      return differ("product list");
  for (size_t k = 0; k < adds.size(); ++k)
    if (adds[k].x != H.adds[k].x || adds[k].y != H.adds[k].y || adds[k].out != H.adds[k].out)
      return differ("add list");
  return PCG_OK;
}

// Two cached plans: the binary splitting produces at most two subtree sizes
// at a given depth.
BsrPlan g_bsr_plans[2];
int g_bsr_next = 0;

SmallTable small_table() {
  SmallTable t;
  for (int b = 0; b <= kSmallNttMaxLog; ++b) t[b] = g_small_class[b];
  return t;
}

// The cached plan of a subtree of d leaves with leaf bounds leaf, laid out
// if needed in the older of the two slots (reusing its buffers).
int bsr_plan_get(Ctx& c, uint32_t d, const LeafBits& leaf, BsrPlan** out) {
  for (BsrPlan& q : g_bsr_plans)
    if (q.d == d && q.leaf == leaf && q.meta) {
      *out = &q;
      return PCG_OK;
    }
  BsrPlan* P = &g_bsr_plans[g_bsr_next];
  g_bsr_next ^= 1;
  const int rc = bsr_plan_build(c, d, leaf, small_table(), *P);
  if (rc) return rc;
  *out = P;
  return PCG_OK;
}

// Grows the batched scratch to what every launch of the plan's level
// schedule needs, so that the schedule itself allocates nothing (it can be
// captured).
int bsr_reserve(Ctx& c, const BsrPlan& P) {
  for (int32_t h = 4; h <= P.H; ++h)
    for (const auto& cl : P.hp[h].classes) {
      int rc = PCG_OK;
      if (cl.pairs || cl.wrap)
        rc = batched_reserve(c, cl.n, cl.cls >= kPassClass ? cl.cls - kPassClass : cl.cls, cl.pairs, cl.wrap);
      else if (cl.cls >= kPassClass)
        rc = batched_reserve(c, cl.n, cl.cls - kPassClass, false, 0);
      else if (cl.cls > kSmallNttMaxLog)
        rc = batched_reserve(c, cl.n, cl.cls, false, cl.wrap);
      if (rc) return rc;
    }
  return PCG_OK;
}

// The level schedule of a subtree on the plan (the first index a in
// P.a_dev): leaves and nodes of up to four leaves, then per height the
// operand descriptors, the products by class, the trims and the adds.
// Launches only (after bsr_reserve: no allocation, no synchronisation).
int bsr_schedule(Ctx& c, const BsrPlan& Pl) {
  const BsrPlan* P = &Pl;
  // the large adds' slots start zeroed (k_bsr_add_finish clears each after use)
  if (P->nslots) {
    static_assert(sizeof(BsrAddSlot) % 4 == 0, "slots clear as words");
    const uint64_t words = P->nslots * sizeof(BsrAddSlot) / 4;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, c.stream,
                       reinterpret_cast<uint32_t*>(P->add_slots), words);
  }
  // leaves and every node of height 2 and 3 in one launch
  hipLaunchKernelGGL(k_bsr_base, dim3((unsigned)((P->nleaves + kBsrBaseThreads - 1) / kBsrBaseThreads)),
                     dim3(kBsrBaseThreads), 0, c.stream,
                     P->leaf_rel, P->a_dev, P->leaf_ids, P->nleaves, P->meta, P->arena);
  for (int32_t h = 4; h <= P->H; ++h) {
    const BsrHeight& hh = P->hp[h];
    const BsrMul* hm = P->muls + hh.mul0;
    BatchDesc* hd = P->descs + hh.mul0;
    hipLaunchKernelGGL(k_bsr_descs, dim3((unsigned)((hh.nmul + 255) / 256)), dim3(256), 0, c.stream, hm,
                       (uint64_t)hh.nmul, P->meta, P->arena, hd);
    size_t o = 0;
    for (const auto& cl : hh.classes) {
      int rc;
      if (cl.pairs || cl.wrap)
        rc = batched_multipass(c, hd + o, cl.n, cl.cls >= kPassClass ? cl.cls - kPassClass : cl.cls,
                               cl.pairs, cl.wrap);
      else if (cl.cls <= kSmallNttMaxLog || cl.cls >= kPassClass)
        rc = launch_small_class(c, cl.cls, hd + o, cl.n);
      else
        rc = batched_multipass(c, hd + o, cl.n, cl.cls, false, cl.wrap);
      if (rc) return rc;
      o += cl.n;
    }
    hipLaunchKernelGGL(k_bsr_trim, dim3((unsigned)((hh.nmul + 255) / 256)), dim3(256), 0, c.stream, hm,
                       (uint64_t)hh.nmul, P->meta, P->arena);
    if (hh.nsmall)
      hipLaunchKernelGGL(k_bsr_add<64>, dim3((unsigned)hh.nsmall), dim3(64), 0, c.stream,
                         P->adds + hh.add0, P->meta, P->arena);
    if (hh.nlarge) {
      const BsrAdd* la = P->adds + hh.add0 + hh.nsmall;
      const uint64_t n = hh.nlarge;
      const uint32_t mc = hh.add_maxch;
      const dim3 grid(mc, (unsigned)n);
      // the slots are zero here (zeroed before the leaves, cleared after use)
      hipLaunchKernelGGL(k_bsr_add_cmp, grid, dim3(256), 0, c.stream, la, P->meta, P->arena, P->add_slots);
      hipLaunchKernelGGL(k_bsr_add_local, grid, dim3(kCarryThreads), 0, c.stream, la, P->meta, P->arena,
                         P->add_slots, P->add_agg, mc);
      if (mc > 64) {
        hipLaunchKernelGGL(k_bsr_add_scan_wide, dim3((unsigned)n), dim3(256), 0, c.stream, la, P->meta,
                           P->arena, P->add_slots, P->add_agg, P->add_cin, mc);
        hipLaunchKernelGGL(k_bsr_add_fix, dim3((unsigned)((n * mc + 255) / 256)), dim3(256), 0, c.stream, la,
                           P->meta, P->arena, P->add_slots, P->add_cin, mc, n);
      } else {  // the scan adds the carry-ins itself
        hipLaunchKernelGGL(k_bsr_add_scan, dim3((unsigned)n), dim3(64), 0, c.stream, la, P->meta, P->arena,
                           P->add_slots, P->add_agg, P->add_cin, mc);
      }
      hipLaunchKernelGGL(k_bsr_add_finish, dim3((unsigned)n), dim3(kCarryThreads), 0, c.stream, la, P->meta,
                         P->arena, P->add_slots);
    }
  }
  HIPCHK(hipGetLastError());
  return PCG_OK;
}

__global__ void k_put_u32(uint32_t* p, uint32_t v) { *p = v; }

// Whether subtree evaluations replay their plan's captured level schedule
// (pcg_bsr_graphs).  Off by default: the pi computation's subtrees mostly
// have plans of their own (plans follow each subtree's leaf bounds), so
// nearly every evaluation pays a capture and an instantiation on the host
// while the GPU waits (pi to 1e9 digits 1.284-1.289 s with graphs against
// 1.261 s kernel by kernel on one box).
bool g_bsr_graphs = false;

// Runs the level schedule of plan P: eagerly, or as the plan's graph
// (captured on first use, and again after the batched scratch it names was
// reallocated).
int bsr_run_schedule(Ctx& c, BsrPlan& P) {
  if (!g_bsr_graphs) return bsr_schedule(c, P);
  if (P.exec && P.graph_gen != c.scratch_gen) bsr_plan_drop_graph(P);
  if (!P.exec) {
    HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    const int rc = bsr_schedule(c, P);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c.stream, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return rc ? rc : fail(PCG_ERR_HIP, "capturing the subtree schedule: %s", hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return fail(PCG_ERR_HIP, "instantiating the subtree schedule");
    }
    P.graph = g;
    P.exec = x;
    P.graph_gen = c.scratch_gen;
  }
  HIPCHK(hipGraphLaunch(P.exec, c.stream));
  return PCG_OK;
}

// Evaluates the subtree [a, b) into the plan's arena; rv receives the
// metadata of P, Q, R (their limbs stay in (*Pout)->arena until the next call).
int bsr_eval_locked(Ctx& c, uint32_t a, uint32_t b, uint64_t max_limbs, const uint64_t* caps,
                    BsrPlan** Pout, BsrVal rv[3]) {
  if (b <= a) return fail(PCG_ERR_INVALID, "empty range");
  int rct = ensure_tuned(c);  // plans record the small-product classes
  if (rct) return rct;
  const uint32_t d = b - a;
  BsrPlan* P = nullptr;
  int rc = bsr_plan_get(c, d, leaf_bits_for(b), &P);  // the leaves are a + 1 .. b
  if (rc) return rc;
  if (P->maxb > max_limbs)
    return fail(PCG_ERR_TOO_LARGE, "subtree values may reach %llu limbs > %llu",
                (unsigned long long)P->maxb, (unsigned long long)max_limbs);
  if (caps)
    for (int j = 0; j < 3; ++j)
      if (caps[j] < P->root_bound[j]) return fail(PCG_ERR_TOO_LARGE, "result buffer too small");
  rc = bsr_reserve(c, *P);
  if (rc) return rc;
  hipLaunchKernelGGL(k_put_u32, dim3(1), dim3(1), 0, c.stream, P->a_dev, a);
  rc = bsr_run_schedule(c, *P);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(rv, P->meta, 3 * sizeof(BsrVal), hipMemcpyDeviceToHost, c.stream));
  rc = wait_stream(c);
  if (rc) return rc;
  *Pout = P;
  return PCG_OK;
}

int bsr_exact_locked(Ctx& c, uint32_t a, uint32_t b, uint64_t max_limbs, pcg_bsr_value* out[3]) {
  const uint64_t caps[3] = {out[0]->cap, out[1]->cap, out[2]->cap};
  BsrPlan* P = nullptr;
  BsrVal rv[3];
  int rc = bsr_eval_locked(c, a, b, max_limbs, caps, &P, rv);
  if (rc) return rc;
  for (int j = 0; j < 3; ++j) {
    out[j]->L = rv[j].L;
    out[j]->sign = rv[j].sign;
    if (rv[j].L)
      HIPCHK(hipMemcpyAsync(out[j]->limbs, P->arena + rv[j].off, rv[j].L * 4, hipMemcpyDeviceToHost,
                            c.stream));
  }
  HIPCHK(hipStreamSynchronize(c.stream));
  return PCG_OK;
}

// Waits for the stream and checks the carry-overflow status, whose copy goes
// through the same wait as whatever readback the caller has enqueued.
int sync_locked(Ctx& c) {
  HIPCHK(hipMemcpyAsync(c.h_status, c.status, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  int rc = wait_stream(c);
  if (rc) return rc;
  const uint32_t st = *c.h_status;
  if (st) {
    uint32_t z = 0;
    HIPCHK(hipMemcpy(c.status, &z, sizeof z, hipMemcpyHostToDevice));
    return fail(PCG_ERR_INTERNAL, "carry propagation overflowed (status %u)", st);
  }
  return PCG_OK;
}

// ------------------------------------------------------------------------
// Device-resident BigFloat values (see df_kernels.hpp)
constexpr uint64_t kExtraPrecision = 2;  // BigFloat.h:6

// Pool of limb buffers: capacities are 2^k or 3 * 2^(k-1) limbs (at least
// 256), and released buffers are kept for reuse.  All work runs on the one
// library stream, so a buffer may be handed out again as soon as its last
// use is enqueued.
std::multimap<uint64_t, uint32_t*> g_df_free;

uint64_t df_round(uint64_t n) {
  uint64_t c = 256;
  while (c < n) {
    if (c + c / 2 >= n) return c + c / 2;
    c <<= 1;
  }
  return c;
}

int df_release_pool(Ctx& c) {
  if (g_df_free.empty()) return PCG_OK;
  HIPCHK(hipStreamSynchronize(c.stream));
  for (auto& kv : g_df_free) HIPCHK(hipFree(kv.second));
  g_df_free.clear();
  return PCG_OK;
}

int df_alloc(Ctx& c, uint64_t n, uint32_t** p, uint64_t* cap) {
  const uint64_t want = df_round(n);
  auto it = g_df_free.find(want);
  if (it != g_df_free.end()) {
    *p = it->second;
    *cap = want;
    g_df_free.erase(it);
    return PCG_OK;
  }
  if (hipMalloc(p, want * sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    int rc = df_release_pool(c);
    if (rc) return rc;
    if (hipMalloc(p, want * sizeof(uint32_t)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PCG_ERR_OOM, "cannot allocate a value of %llu limbs", (unsigned long long)want);
    }
  }
  *cap = want;
  return PCG_OK;
}

void df_put_back(uint32_t* p, uint64_t cap) {
  if (p) g_df_free.emplace(cap, p);
}

// z takes the new buffer and metadata; its old buffer goes back to the pool.
void df_assign(pcg_dfloat* z, uint32_t* buf, uint64_t cap, uint64_t L, int64_t exp, int32_t sign) {
  df_put_back(z->limbs, z->cap);
  z->limbs = buf;
  z->cap = cap;
  z->L = L;
  z->exp = exp;
  z->sign = sign;
}

void df_assign_zero(pcg_dfloat* z, int32_t sign = 1) {
  df_put_back(z->limbs, z->cap);
  z->limbs = nullptr;
  z->cap = 0;
  z->L = 0;
  z->exp = 0;
  z->sign = sign;
}

DfOperand df_opnd(const pcg_dfloat& x) {
  DfOperand o;
  o.t = x.limbs;
  o.exp = x.exp;
  o.L = x.L;
  return o;
}

int df_ensure_blocks(Ctx& c, uint64_t nblk) {
  if (nblk <= c.df_nblk_cap) return PCG_OK;
  HIPCHK(hipStreamSynchronize(c.stream));
  if (c.df_agg) HIPCHK(hipFree(c.df_agg));
  if (c.df_cin) HIPCHK(hipFree(c.df_cin));
  c.df_agg = nullptr;
  c.df_cin = nullptr;
  c.df_nblk_cap = 0;
  HIPCHK(hipMalloc(&c.df_agg, (nblk + nblk / kScanChunk + 1) * sizeof(CarryFn)));  // + scan totals
  HIPCHK(hipMalloc(&c.df_cin, nblk * sizeof(uint64_t)));
  c.df_nblk_cap = nblk;
  return PCG_OK;
}

// Copies `n` slots to the pinned host mirror and waits for the stream.
int df_read_slots(Ctx& c, int n) {
  HIPCHK(hipMemcpyAsync(c.h_slot, c.df_slot, n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        c.stream));
  return wait_stream(c);
}

int df_read_limb(Ctx& c, const uint32_t* p, uint32_t* v) {
  HIPCHK(hipMemcpyAsync(c.h_slot, p, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  const int rc = wait_stream(c);
  if (rc) return rc;
  memcpy(v, c.h_slot, sizeof(uint32_t));
  return PCG_OK;
}

// Block pass OP over a.n limbs into a.out, block-function scan, fix-up;
// the carry out of the top lands in df_slot[0].
template <int OP>
int df_carry_pass(Ctx& c, DfArgs a) {
  const uint64_t nblk = (a.n + kCarryBlock - 1) / kCarryBlock;
  if (nblk == 0) {
    hipLaunchKernelGGL(k_df_put_u64, dim3(1), dim3(1), 0, c.stream, c.df_slot, 0ull);
    HIPCHK(hipGetLastError());
    return PCG_OK;
  }
  int rc = df_ensure_blocks(c, nblk);
  if (rc) return rc;
  a.agg = c.df_agg;
  hipLaunchKernelGGL(k_df_local<OP>, dim3((unsigned)nblk), dim3(kCarryThreads), 0, c.stream, a);
  CarryArgs ca;
  memset(&ca, 0, sizeof ca);
  ca.zL = a.n;
  ca.out = a.out;
  ca.agg = c.df_agg;
  ca.cin = c.df_cin;
  ca.nblk = nblk;
  ca.tot = c.df_agg + c.df_nblk_cap;
  ca.carry_out = reinterpret_cast<uint64_t*>(c.df_slot);
  launch_carry_scan(c, ca);
  hipLaunchKernelGGL(k_carry_fix, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, c.stream, ca);
  HIPCHK(hipGetLastError());
  return PCG_OK;
}

// Grid of the top-down searches: at most 1024 workgroups of 256.
unsigned df_search_blocks(uint64_t n) {
  const uint64_t b = (n + 255) / 256;
  return (unsigned)(b < 1024 ? (b ? b : 1) : 1024);
}

// BigFloat::ucmp (BigFloat.cpp:123-147): magnitudes exp + L first, then the
// highest position where the aligned words differ.
int df_ucmp_locked(Ctx& c, const pcg_dfloat& A, const pcg_dfloat& B, int* cmp) {
  const int64_t magA = A.exp + (int64_t)A.L, magB = B.exp + (int64_t)B.L;
  if (magA != magB) {
    *cmp = magA > magB ? 1 : -1;
    return PCG_OK;
  }
  const int64_t lo = A.exp < B.exp ? A.exp : B.exp;
  if (magA <= lo) {
    *cmp = 0;
    return PCG_OK;
  }
  const uint64_t n = (uint64_t)(magA - lo);
  hipLaunchKernelGGL(k_df_put_u64, dim3(1), dim3(1), 0, c.stream, c.df_slot + 2, 0ull);
  hipLaunchKernelGGL(k_df_cmp, dim3(df_search_blocks(n)), dim3(256), 0, c.stream, df_opnd(A),
                     df_opnd(B), lo, n, c.df_slot + 2);
  hipLaunchKernelGGL(k_df_cmp_final, dim3(1), dim3(1), 0, c.stream, df_opnd(A), df_opnd(B), lo,
                     c.df_slot + 2, reinterpret_cast<int64_t*>(c.df_slot + 3));
  HIPCHK(hipGetLastError());
  int rc = df_read_slots(c, 4);
  if (rc) return rc;
  *cmp = (int)(int64_t)c.h_slot[3];
  return PCG_OK;
}

// BigFloat::uadd / usub (BigFloat.cpp:184-232, 233-287): positions
// [bot, top) aligned by exponent, the bottom cut so at most p + 2 limbs
// remain (p = 0: exact); uadd appends a carry limb, usub strips every
// leading zero limb.  usub requires |A| >= |B|.
int df_uaddsub_locked(Ctx& c, pcg_dfloat* z, const pcg_dfloat A, const pcg_dfloat B, uint64_t p,
                      bool sub, int32_t sign) {
  const int64_t magA = A.exp + (int64_t)A.L, magB = B.exp + (int64_t)B.L;
  const int64_t top = magA > magB ? magA : magB;
  int64_t bot = A.exp < B.exp ? A.exp : B.exp;
  int64_t TL = top - bot;
  const uint64_t pp = p == 0 ? (uint64_t)TL : p + kExtraPrecision;
  if (TL > (int64_t)pp) {
    bot = top - (int64_t)pp;
    TL = (int64_t)pp;
  }
  uint32_t* buf = nullptr;
  uint64_t cap = 0;
  int rc = df_alloc(c, (uint64_t)TL + 1, &buf, &cap);
  if (rc) return rc;
  DfArgs a;
  memset(&a, 0, sizeof a);
  a.a = df_opnd(A);
  a.b = df_opnd(B);
  a.bot = bot;
  a.n = (uint64_t)TL;
  a.out = buf;
  rc = sub ? df_carry_pass<kDfSub>(c, a) : df_carry_pass<kDfAdd>(c, a);
  if (rc) {
    df_put_back(buf, cap);
    return rc;
  }
  uint64_t L = (uint64_t)TL;
  int64_t exp = bot;
  if (!sub) {
    rc = df_read_slots(c, 1);
    if (rc) return rc;
    if (c.h_slot[0]) {
      hipLaunchKernelGGL(k_df_put_u32, dim3(1), dim3(1), 0, c.stream, buf, L, 1u);
      ++L;
    }
  } else {
    hipLaunchKernelGGL(k_df_put_u64, dim3(1), dim3(1), 0, c.stream, c.df_slot + 1, 0ull);
    if (L)
      hipLaunchKernelGGL(k_df_topnz, dim3(df_search_blocks(L)), dim3(256), 0, c.stream, buf, L,
                         c.df_slot + 1);
    rc = df_read_slots(c, 2);
    if (rc) return rc;
    L = c.h_slot[1];
    if (L == 0) {
      exp = 0;
      sign = 1;
    }
  }
  HIPCHK(hipGetLastError());
  df_assign(z, buf, cap, L, exp, sign);
  return PCG_OK;
}

// BigFloat::add / sub (BigFloat.cpp:288-325).
int df_add_locked(Ctx& c, pcg_dfloat* z, const pcg_dfloat* pa, const pcg_dfloat* pb, uint64_t p,
                  bool subtract) {
  const pcg_dfloat A = *pa, B = *pb;
  const bool same = (A.sign != 0) == (B.sign != 0);
  if (same != subtract) return df_uaddsub_locked(c, z, A, B, p, false, A.sign);
  int cmp = 0;
  int rc = df_ucmp_locked(c, A, B, &cmp);
  if (rc) return rc;
  if (cmp > 0) return df_uaddsub_locked(c, z, A, B, p, true, A.sign);
  rc = df_uaddsub_locked(c, z, B, A, p, true, B.sign);
  if (rc) return rc;
  if (subtract && z->L) z->sign = !z->sign;  // negate (BigFloat.cpp:150-155)
  return PCG_OK;
}

// Operand of BigFloat::mul after the precision cut (BigFloat.cpp:337-366):
// the top pp limbs, the exponent raised by the limbs dropped.
struct DfCut {
  const uint32_t* t;
  uint64_t L;
  int64_t exp;
};
DfCut df_cut(const pcg_dfloat& x, uint64_t pp) {
  DfCut o{x.limbs, x.L, x.exp};
  if (o.L > pp) {
    o.t += o.L - pp;
    o.exp += (int64_t)(o.L - pp);
    o.L = pp;
  }
  return o;
}

// Limbs kept by BigFloat::mul at precision p (p = 0: exact).
uint64_t df_mul_keep(const pcg_dfloat& A, const pcg_dfloat& B, uint64_t p) {
  return p == 0 ? A.L + B.L : p + kExtraPrecision;
}

// BigFloat::mul (BigFloat.cpp:326-444): operands cut to their top p + 2
// limbs (p = 0: exact), full product, one zero top limb trimmed.
int df_mul_locked(Ctx& c, pcg_dfloat* z, const pcg_dfloat* pa, const pcg_dfloat* pb, uint64_t p,
                  uint32_t flags) {
  const pcg_dfloat A = *pa, B = *pb;
  if (A.L == 0 || B.L == 0) {
    df_assign_zero(z);
    return PCG_OK;
  }
  const uint64_t pp = df_mul_keep(A, B, p);
  const DfCut a = df_cut(A, pp), b = df_cut(B, pp);
  uint32_t* buf = nullptr;
  uint64_t cap = 0;
  int rc = df_alloc(c, a.L + b.L, &buf, &cap);
  if (rc) return rc;
  if (flags & (PCG_KEEP_B | PCG_REUSE_B)) {
    // b first: its transform is the one kept or reused (exact products commute)
    const uint32_t keep = ((flags & PCG_KEEP_B) ? kKeepA : 0u) | ((flags & PCG_REUSE_B) ? kReuseA : 0u);
    rc = mul_dev_locked(c, b.t, b.L, a.t, a.L, buf, flags & PCG_ENGINE_MASK, keep);
  } else {
    rc = mul_dev_locked(c, a.t, a.L, b.t, b.L, buf, flags & PCG_ENGINE_MASK);
  }
  // the top limb comes back through the status check's wait
  if (!rc && hipMemcpyAsync(c.h_slot, buf + a.L + b.L - 1, sizeof(uint32_t), hipMemcpyDeviceToHost,
                            c.stream) != hipSuccess)
    rc = fail(PCG_ERR_HIP, "top limb readback failed");
  if (!rc) rc = sync_locked(c);
  if (rc) {
    df_put_back(buf, cap);
    return rc;
  }
  uint32_t topl = 0;
  memcpy(&topl, c.h_slot, sizeof topl);
  df_assign(z, buf, cap, a.L + b.L - (topl == 0 ? 1 : 0), a.exp + b.exp, (A.sign != 0) == (B.sign != 0));
  return PCG_OK;
}

// Two products with one shared operand, z1 = x1 s and z2 = x2 s, each with
// the rules of BigFloat::mul.  Both cut s to the same limbs (its top p + 2,
// or all of it when p = 0), so when the two products have the same
// transform length the forward transform of s runs once: the node combine of
// PI.h:110-112 multiplies Q1 and R0 twice each.  The results are assigned
// after both products are complete, so z1 and z2 may alias any operand.
int df_mul_pair_locked(Ctx& c, pcg_dfloat* z1, const pcg_dfloat* px1, pcg_dfloat* z2,
                       const pcg_dfloat* px2, const pcg_dfloat* ps, uint64_t p, uint32_t flags) {
  if (z1 == z2) return fail(PCG_ERR_INVALID, "the two results of a product pair must be distinct");
  const pcg_dfloat X[2] = {*px1, *px2}, S = *ps;
  const uint32_t eng = flags & PCG_ENGINE_MASK;
  if (eng == 3) return fail(PCG_ERR_INVALID, "unknown engine flag");
  DfCut x[2], sc[2];
  uint64_t zL[2] = {0, 0};
  for (int i = 0; i < 2; ++i) {
    if (X[i].L == 0 || S.L == 0) continue;
    const uint64_t pp = df_mul_keep(X[i], S, p);
    x[i] = df_cut(X[i], pp);
    sc[i] = df_cut(S, pp);
    zL[i] = x[i].L + sc[i].L;
    const int rc = check_sizes(x[i].L, sc[i].L);
    if (rc) return rc;
  }
  MulPlan mp[2];
  for (int i = 0; i < 2; ++i)
    if (zL[i]) mp[i] = mul_plan(sc[i].L, x[i].L);
  const bool shared = zL[0] && zL[1] && eng != PCG_ENGINE_FFT && sc[0].L == sc[1].L &&
                      mp[0].logN == mp[1].logN;
  uint32_t* buf[2] = {nullptr, nullptr};
  uint64_t cap[2] = {0, 0};
  int rc = PCG_OK;
  for (int i = 0; i < 2 && !rc; ++i)
    if (zL[i]) rc = df_alloc(c, zL[i], &buf[i], &cap[i]);
  if (!rc && shared) {
    // one forward transform of s (the main length of both products), then
    // each product's short top product if it wraps, its tail and carry
    ++c.spec_gen;
    for (int i = 0; i < 2 && !rc; ++i) rc = ensure_mul_workspace(c, mp[i], zL[i]);
    const Plan pl = make_plan(mp[0].logN);
    if (!rc) rc = ntt_forward_into(c, pl, sc[0].t, sc[0].L, c.spec);
    for (int i = 0; i < 2 && !rc; ++i)
      rc = ntt_mul_tail(c, pl, mp[i], sc[i].t, sc[i].L, x[i].t, x[i].L, buf[i], false);
  } else {
    for (int i = 0; i < 2 && !rc; ++i)
      if (zL[i]) rc = mul_dev_locked(c, x[i].t, x[i].L, sc[i].t, sc[i].L, buf[i], eng);
  }
  for (int i = 0; i < 2 && !rc; ++i)
    if (zL[i])
      HIPCHK(hipMemcpyAsync(reinterpret_cast<uint32_t*>(c.h_slot) + i, buf[i] + zL[i] - 1, sizeof(uint32_t),
                            hipMemcpyDeviceToHost, c.stream));
  if (!rc) rc = sync_locked(c);
  if (rc) {
    for (int i = 0; i < 2; ++i) df_put_back(buf[i], cap[i]);
    return rc;
  }
  uint32_t top[2];
  memcpy(top, c.h_slot, sizeof top);
  pcg_dfloat* z[2] = {z1, z2};
  for (int i = 0; i < 2; ++i) {
    if (!zL[i]) {
      df_assign_zero(z[i]);
      continue;
    }
    df_assign(z[i], buf[i], cap[i], zL[i] - (top[i] == 0 ? 1 : 0), x[i].exp + sc[i].exp,
              (X[i].sign != 0) == (S.sign != 0));
  }
  return PCG_OK;
}

// BigFloat::mul(uint32_t) (BigFloat.cpp:156-183).
int df_mul_u32_locked(Ctx& c, pcg_dfloat* z, const pcg_dfloat* pa, uint32_t x) {
  const pcg_dfloat A = *pa;
  if (A.L == 0 || x == 0) {
    df_assign_zero(z);
    return PCG_OK;
  }
  uint32_t* buf = nullptr;
  uint64_t cap = 0;
  int rc = df_alloc(c, A.L + 1, &buf, &cap);
  if (rc) return rc;
  DfArgs a;
  memset(&a, 0, sizeof a);
  a.a = df_opnd(A);
  a.n = A.L + 1;
  a.x = x;
  a.out = buf;
  rc = df_carry_pass<kDfMulU32>(c, a);
  uint32_t topl = 0;
  if (!rc) rc = df_read_limb(c, buf + A.L, &topl);
  if (rc) {
    df_put_back(buf, cap);
    return rc;
  }
  df_assign(z, buf, cap, A.L + (topl ? 1 : 0), A.exp, A.sign);
  return PCG_OK;
}

// The product of pcg_mul_dev on the NTT engine with every scratch buffer
// carved from the static array (c.noalloc): no allocation of any kind.
int mul_noalloc_locked(Ctx& c, const uint32_t* d_a, uint64_t na, const uint32_t* d_b, uint64_t nb,
                       uint32_t* d_out, uint32_t flags) {
  const uint32_t eng = flags & PCG_ENGINE_MASK;
  if (eng != PCG_ENGINE_AUTO && eng != PCG_ENGINE_NTT)
    return fail(PCG_ERR_INVALID, "pcg_mul_dev_noalloc runs the NTT engine only");
  const uint64_t zL = na + nb;
  if (na > kNoAllocMaxLimbs || nb > kNoAllocMaxLimbs || zL > kNoAllocMaxLimbs)
    return fail(PCG_ERR_TOO_LARGE,
                "pcg_mul_dev_noalloc: na + nb = %llu limbs exceeds its static scratch limit of %llu limbs",
                (unsigned long long)zL, (unsigned long long)kNoAllocMaxLimbs);
  if (zL == 0) return PCG_OK;
  if (na == 0 || nb == 0) {
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((zL + 255) / 256)), dim3(256), 0, c.stream, d_out, zL);
    HIPCHK(hipGetLastError());
    return PCG_OK;
  }
  const int logN = ntt_log_points(zL);
  if (logN > kNoAllocMaxLogN) return fail(PCG_ERR_INTERNAL, "noalloc transform of 2^%d points", logN);
  const Plan pl = make_plan(logN);
  const uint64_t N = 1ull << logN;
  double* A = reinterpret_cast<double*>(c.noalloc);
  double* B = A + 2 * N;
  CarryFn* agg = reinterpret_cast<CarryFn*>(c.noalloc + kNoAllocSpecBytes);
  uint64_t* cin = reinterpret_cast<uint64_t*>(c.noalloc + kNoAllocSpecBytes + kNoAllocAggBytes);
  // every point of A and B and every carry entry read below is written
  // first by this product, so nothing needs clearing between calls
  int rc = ntt_forward_into(c, pl, d_a, na, A);
  if (!rc) rc = ntt_tail_into(c, pl, d_b, nb, A, B);
  if (!rc) rc = ntt_carry(c, pl, B, zL, d_out, WrapPlan(), CarryScratch{agg, agg + kNoAllocBlocks, cin});
  return rc;
}

// Set-up of pcg_init; on failure the caller releases whatever was built.
int init_locked(Ctx& c, int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return fail(PCG_ERR_NO_DEVICE, "no HIP device available");
  }
  if (device < 0 || device >= n) return fail(PCG_ERR_NO_DEVICE, "device %d out of range (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  {
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    c.num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  }
  HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  int rc = build_tables(c);
  if (rc) return rc;
  HIPCHK(hipMalloc(&c.status, sizeof(uint32_t)));
  HIPCHK(hipMemset(c.status, 0, sizeof(uint32_t)));
  HIPCHK(hipMalloc(&c.df_slot, 4 * sizeof(unsigned long long)));
  HIPCHK(hipHostMalloc(&c.h_slot, 4 * sizeof(unsigned long long), hipHostMallocDefault));
  HIPCHK(hipHostMalloc(&c.h_status, sizeof(uint32_t), hipHostMallocDefault));
  HIPCHK(hipEventCreateWithFlags(&c.ev_wait, hipEventDisableTiming));
  for (int i = 0; i < 4; ++i) HIPCHK(hipEventCreate(&c.ev[i]));
  for (int i = 0; i < 2; ++i) HIPCHK(hipEventCreate(&c.tev[i]));
  for (int b = 0; b <= kSmallNttMaxLog; ++b) g_small_class[b] = small_class_default(b);
  for (int b = 0; b <= kGeomMaxLog; ++b) g_batch_geom[b] = batch_geom_default(b);
  g_small_ready = false;
  // this device's copy of the allocation-free product's scratch array
  HIPCHK(hipGetSymbolAddress(reinterpret_cast<void**>(&c.noalloc), HIP_SYMBOL(scratch)));
  if (!c.noalloc) return fail(PCG_ERR_HIP, "no device address for the static scratch array");
  c.device = device;
  c.ready = true;
  return PCG_OK;
}

// Frees every buffer, event and the stream (each only if it exists) and
// returns the context to its initial state.
void release_locked(Ctx& c) {
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  for (int pi = 0; pi < 2; ++pi)
    for (int t = 0; t < 3; ++t) {
      (void)hipFree(c.tw[pi][t]);
      c.tw[pi][t] = nullptr;
    }
  for (int t = 0; t < 3; ++t) {
    (void)hipFree(c.ftab[t]);
    c.ftab[t] = nullptr;
  }
  for (BsrPlan& q : g_bsr_plans) bsr_plan_free(q);
  if (c.stream) (void)df_release_pool(c);
  void* dev[] = {c.d_maxdist, c.d_prime, c.spec, c.top_spec, c.agg, c.cin, c.status, c.d_stage,
                 c.d_pool, c.d_desc, c.bspec, c.wrap_desc, c.bagg, c.bcin, c.df_agg, c.df_cin,
                 c.df_slot};
  for (void* q : dev) (void)hipFree(q);
  void* host[] = {c.h_pool, c.h_desc, c.h_slot, c.h_status};
  for (void* q : host) (void)hipHostFree(q);
  if (c.ev_wait) (void)hipEventDestroy(c.ev_wait);
  c.ev_wait = nullptr;
  for (int i = 0; i < 4; ++i) {
    if (c.ev[i]) (void)hipEventDestroy(c.ev[i]);
    c.ev[i] = nullptr;
  }
  for (int i = 0; i < 2; ++i) {
    if (c.tev[i]) (void)hipEventDestroy(c.tev[i]);
    c.tev[i] = nullptr;
  }
  if (c.stream) (void)hipStreamDestroy(c.stream);
  (void)hipGetLastError();
  c.ready = false;
  c.device = -1;
  c.stream = nullptr;
  c.d_maxdist = nullptr;
  c.d_prime = nullptr;
  c.spec = nullptr;
  c.spec_points = 0;
  c.top_spec = nullptr;
  c.top_points = 0;
  c.agg = nullptr;
  c.cin = nullptr;
  c.nblk_cap = 0;
  c.status = nullptr;
  c.d_stage = nullptr;
  c.stage_limbs = 0;
  c.h_pool = nullptr;
  c.d_pool = nullptr;
  c.pool_limbs = 0;
  c.h_desc = nullptr;
  c.d_desc = nullptr;
  c.desc_cap = 0;
  c.bspec = nullptr;
  c.bspec_doubles = 0;
  c.wrap_desc = nullptr;
  c.wrap_desc_cap = 0;
  c.bagg = nullptr;
  c.bcin = nullptr;
  c.bcarry_cap = 0;
  c.prof_valid = false;
  c.df_agg = nullptr;
  c.df_cin = nullptr;
  c.df_nblk_cap = 0;
  c.df_slot = nullptr;
  c.h_slot = nullptr;
  c.h_status = nullptr;
  c.noalloc = nullptr;
}

}  // namespace

extern "C" {

int pcg_init(int device) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (c.ready) {
    if (device != c.device) return fail(PCG_ERR_INVALID, "already initialised on device %d", c.device);
    return PCG_OK;
  }
  const int rc = init_locked(c, device);
  if (rc) release_locked(c);  // nothing half-built survives a failed init
  return rc;
}

void pcg_shutdown(void) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return;
  release_locked(c);
}

const char* pcg_last_error_string(void) { return g_err.c_str(); }

int pcg_reserve(uint64_t total_limbs) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called");
  if (total_limbs > PCG_MAX_TOTAL_LIMBS) return fail(PCG_ERR_TOO_LARGE, "FFT length is too small.");
  int logN = ceil_log2(total_limbs);
  if (logN < kMinLogN) logN = kMinLogN;
  return ensure_workspace(c, logN, total_limbs);
}

int pcg_mul_dev(const uint32_t* d_a, uint64_t na, const uint32_t* d_b, uint64_t nb, uint32_t* d_out,
                uint32_t flags) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called");
  if ((na && !d_a) || (nb && !d_b) || (na + nb && !d_out)) return fail(PCG_ERR_INVALID, "null pointer");
  int rc = mul_dev_locked(c, d_a, na, d_b, nb, d_out, flags);
  if (rc) return rc;
  if (flags & PCG_SYNC) return sync_locked(c);
  return PCG_OK;
}

int pcg_mul_dev_noalloc(const uint32_t* d_a, uint64_t na, const uint32_t* d_b, uint64_t nb,
                        uint32_t* d_out, uint32_t flags) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called");
  if ((na && !d_a) || (nb && !d_b) || (na + nb && !d_out)) return fail(PCG_ERR_INVALID, "null pointer");
  int rc = mul_noalloc_locked(c, d_a, na, d_b, nb, d_out, flags);
  if (rc) return rc;
  if (flags & PCG_SYNC) return sync_locked(c);
  return PCG_OK;
}

int pcg_mul(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb, uint32_t* out,
            uint32_t flags) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called");
  if ((na && !a) || (nb && !b) || (na + nb && !out)) return fail(PCG_ERR_INVALID, "null pointer");
  int rc = check_sizes(na, nb);
  if (rc) return rc;
  const uint64_t zL = na + nb;
  if (zL == 0) return PCG_OK;
  if (na == 0 || nb == 0) {
    memset(out, 0, zL * sizeof(uint32_t));
    return PCG_OK;
  }
  rc = ensure_stage(c, 2 * zL);
  if (rc) return rc;
  uint32_t* da = c.d_stage;
  uint32_t* db = c.d_stage + na;
  uint32_t* dout = c.d_stage + zL;
  HIPCHK(hipMemcpyAsync(da, a, na * 4, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipMemcpyAsync(db, b, nb * 4, hipMemcpyHostToDevice, c.stream));
  rc = mul_dev_locked(c, da, na, db, nb, dout, flags);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(out, dout, zL * 4, hipMemcpyDeviceToHost, c.stream));
  return sync_locked(c);
}

int pcg_mul_batch_dev(const pcg_mul_desc* d, uint64_t count, uint32_t flags) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called");
  if (count && !d) return fail(PCG_ERR_INVALID, "null descriptor array");
  if (flags & PCG_BATCH_PLAIN) {
    for (uint64_t i = 0; i < count; ++i) {
      int rc = mul_dev_locked(c, d[i].a, d[i].na, d[i].b, d[i].nb, d[i].out, flags & ~PCG_PROFILE);
      if (rc) return rc;
    }
    return sync_locked(c);
  }
  int rc = batch_locked(c, d, count, flags, true);
  if (rc) return rc;
  return sync_locked(c);
}

int pcg_mul_batch(const pcg_mul_desc* d, uint64_t count, uint32_t flags) {
  if (count && !d) return fail(PCG_ERR_INVALID, "null descriptor array");
  if (flags & PCG_BATCH_PLAIN) {
    for (uint64_t i = 0; i < count; ++i) {
      int rc = pcg_mul(d[i].a, d[i].na, d[i].b, d[i].nb, d[i].out, flags & ~PCG_PROFILE);
      if (rc) return rc;
    }
    return PCG_OK;
  }
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called");
  int rc = batch_locked(c, d, count, flags, false);
  if (rc) return rc;
  return sync_locked(c);
}

int pcg_sync(void) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called");
  return sync_locked(c);
}

int pcg_dev_alloc(void** ptr, uint64_t bytes) {
  if (!ptr) return fail(PCG_ERR_INVALID, "null pointer");
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PCG_ERR_OOM, "hipMalloc(%llu) failed", (unsigned long long)bytes);
  }
  return PCG_OK;
}
int pcg_dev_free(void* ptr) {
  HIPCHK(hipFree(ptr));
  return PCG_OK;
}
int pcg_host_alloc(void** ptr, uint64_t bytes) {
  if (!ptr) return fail(PCG_ERR_INVALID, "null pointer");
  if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(PCG_ERR_OOM, "hipHostMalloc(%llu) failed", (unsigned long long)bytes);
  }
  return PCG_OK;
}
int pcg_host_free(void* ptr) {
  HIPCHK(hipHostFree(ptr));
  return PCG_OK;
}
int pcg_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return PCG_OK;
}
int pcg_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return PCG_OK;
}

int pcg_fft_last_roundoff(double* max_dist, int* recomputed) {
  Ctx& c = g_ctx;
  if (max_dist) *max_dist = c.fft_maxdist;
  if (recomputed) *recomputed = c.fft_recomputed;
  return PCG_OK;
}

int pcg_last_profile(double* stage_ms, int n_stage, double* total_ms) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.prof_valid) return fail(PCG_ERR_INVALID, "no profiled product yet");
  for (int i = 0; i < n_stage && i < 3; ++i) stage_ms[i] = c.prof_ms[i];
  if (total_ms) *total_ms = c.prof_total;
  return PCG_OK;
}

int pcg_bsr_exact(uint32_t a, uint32_t b, uint64_t max_limbs, pcg_bsr_value* P, pcg_bsr_value* Q,
                  pcg_bsr_value* R) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called");
  if (!P || !Q || !R || !P->limbs || !Q->limbs || !R->limbs) return fail(PCG_ERR_INVALID, "null pointer");
  pcg_bsr_value* out[3] = {P, Q, R};
  return bsr_exact_locked(c, a, b, max_limbs, out);
}

int pcg_bsr_bound(uint32_t a, uint32_t b, uint64_t* bound) {
  if (!bound || b <= a) return fail(PCG_ERR_INVALID, "bad arguments");
  *bound = bsr_max_limbs(b - a, leaf_bits_for(b));
  return PCG_OK;
}

int pcg_timer_start(void) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called");
  HIPCHK(hipEventRecord(c.tev[0], c.stream));
  return PCG_OK;
}

int pcg_timer_stop(double* ms) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called");
  if (!ms) return fail(PCG_ERR_INVALID, "null pointer");
  HIPCHK(hipEventRecord(c.tev[1], c.stream));
  HIPCHK(hipEventSynchronize(c.tev[1]));
  float f = 0;
  HIPCHK(hipEventElapsedTime(&f, c.tev[0], c.tev[1]));
  *ms = f;
  return PCG_OK;
}

int pcg_ntt_plan(uint64_t total_limbs, int* log2_points, int* passes) {
  if (total_limbs > PCG_MAX_TOTAL_LIMBS) return fail(PCG_ERR_TOO_LARGE, "FFT length is too small.");
  int logN = ceil_log2(total_limbs);
  if (logN < kMinLogN) logN = kMinLogN;
  const Plan pl = make_plan(logN);
  if (!plan_supported(pl)) return fail(PCG_ERR_INTERNAL, "no kernel instance for 2^%d points", logN);
  if (log2_points) *log2_points = logN;
  if (passes) *passes = (int)pl.passes.size();
  return PCG_OK;
}

// ---------------------------------------------------------------- device values
#define DF_ENTER()                                                        \
  Ctx& c = g_ctx;                                                         \
  std::lock_guard<std::mutex> lk(c.mu);                                   \
  if (!c.ready) return fail(PCG_ERR_NOT_INIT, "pcg_init has not been called")

int pcg_df_free(pcg_dfloat* x) {
  if (!x) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  df_assign_zero(x);
  return PCG_OK;
}

int pcg_df_set(pcg_dfloat* x, const uint32_t* limbs, uint64_t L, int64_t exp, int32_t sign) {
  if (!x || (L && !limbs)) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  if (L == 0) {
    df_assign_zero(x, sign ? 1 : 0);
    x->exp = exp;
    return PCG_OK;
  }
  uint32_t* buf = nullptr;
  uint64_t cap = 0;
  int rc = df_alloc(c, L, &buf, &cap);
  if (rc) return rc;
  if (hipMemcpyAsync(buf, limbs, L * 4, hipMemcpyHostToDevice, c.stream) != hipSuccess ||
      hipStreamSynchronize(c.stream) != hipSuccess) {
    df_put_back(buf, cap);
    return fail(PCG_ERR_HIP, "value upload failed: %s", hipGetErrorString(hipGetLastError()));
  }
  df_assign(x, buf, cap, L, exp, sign ? 1 : 0);
  return PCG_OK;
}

int pcg_df_read(const pcg_dfloat* x, uint64_t first, uint64_t count, uint32_t* out) {
  if (!x || (count && !out)) return fail(PCG_ERR_INVALID, "null pointer");
  if (first > x->L || count > x->L - first) return fail(PCG_ERR_INVALID, "range beyond the value");
  DF_ENTER();
  if (count == 0) return PCG_OK;
  HIPCHK(hipMemcpyAsync(out, x->limbs + first, count * 4, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return PCG_OK;
}

int pcg_df_copy(pcg_dfloat* z, const pcg_dfloat* x) {
  if (!z || !x) return fail(PCG_ERR_INVALID, "null pointer");
  if (z == x) return PCG_OK;
  DF_ENTER();
  const pcg_dfloat X = *x;
  if (X.L == 0) {
    df_assign_zero(z, X.sign);
    z->exp = X.exp;
    return PCG_OK;
  }
  uint32_t* buf = nullptr;
  uint64_t cap = 0;
  int rc = df_alloc(c, X.L, &buf, &cap);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(buf, X.limbs, X.L * 4, hipMemcpyDeviceToDevice, c.stream));
  df_assign(z, buf, cap, X.L, X.exp, X.sign);
  return PCG_OK;
}

int pcg_df_ucmp(const pcg_dfloat* a, const pcg_dfloat* b, int* cmp) {
  if (!a || !b || !cmp) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  return df_ucmp_locked(c, *a, *b, cmp);
}

int pcg_df_add(pcg_dfloat* z, const pcg_dfloat* a, const pcg_dfloat* b, uint64_t p) {
  if (!z || !a || !b) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  return df_add_locked(c, z, a, b, p, false);
}

int pcg_df_sub(pcg_dfloat* z, const pcg_dfloat* a, const pcg_dfloat* b, uint64_t p) {
  if (!z || !a || !b) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  return df_add_locked(c, z, a, b, p, true);
}

int pcg_df_mul(pcg_dfloat* z, const pcg_dfloat* a, const pcg_dfloat* b, uint64_t p, uint32_t flags) {
  if (!z || !a || !b) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  return df_mul_locked(c, z, a, b, p, flags);
}

int pcg_df_mul_pair(pcg_dfloat* z1, const pcg_dfloat* x1, pcg_dfloat* z2, const pcg_dfloat* x2,
                    const pcg_dfloat* s, uint64_t p, uint32_t flags) {
  if (!z1 || !x1 || !z2 || !x2 || !s) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  return df_mul_pair_locked(c, z1, x1, z2, x2, s, p, flags);
}

int pcg_df_mul_u32(pcg_dfloat* z, const pcg_dfloat* a, uint32_t x) {
  if (!z || !a) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  return df_mul_u32_locked(c, z, a, x);
}

int pcg_df_to_hex(const pcg_dfloat* x, uint64_t digits, char* out, uint64_t cap, uint64_t* written) {
  if (!x || !out || !written) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  const pcg_dfloat X = *x;
  const uint64_t nl = (digits + 7) / 8;
  if (X.L == 0 || nl > X.L - 1) return fail(PCG_ERR_INVALID, "Limit Exceeded");
  uint32_t topl = 0;
  int rc = df_read_limb(c, X.limbs + X.L - 1, &topl);
  if (rc) return rc;
  char head[16];
  const int hn = snprintf(head, sizeof head, "%u.", topl);
  const uint64_t total = (uint64_t)hn + 8 * nl;
  const uint64_t want = 2 + digits < total ? 2 + digits : total;
  if (cap < want + 1) return fail(PCG_ERR_TOO_LARGE, "output buffer too small");
  const uint64_t nhex = want > (uint64_t)hn ? want - (uint64_t)hn : 0;
  memcpy(out, head, want < (uint64_t)hn ? want : (uint64_t)hn);
  if (nhex) {
    uint32_t* buf = nullptr;
    uint64_t bcap = 0;
    rc = df_alloc(c, (nhex + 3) / 4, &buf, &bcap);
    if (rc) return rc;
    char* dch = reinterpret_cast<char*>(buf);
    hipLaunchKernelGGL(k_df_hex, dim3((unsigned)((nhex + 255) / 256)), dim3(256), 0, c.stream, X.limbs,
                       X.L, nhex, dch);
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpyAsync(out + hn, dch, nhex, hipMemcpyDeviceToHost, c.stream) == hipSuccess &&
                    hipStreamSynchronize(c.stream) == hipSuccess;
    df_put_back(buf, bcap);
    if (!ok) return fail(PCG_ERR_HIP, "hex conversion failed");
  }
  out[want] = 0;
  *written = want;
  return PCG_OK;
}

int pcg_df_bsr_exact(uint32_t a, uint32_t b, uint64_t max_limbs, pcg_dfloat* P, pcg_dfloat* Q,
                     pcg_dfloat* R) {
  if (!P || !Q || !R) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  BsrPlan* plan = nullptr;
  BsrVal rv[3];
  int rc = bsr_eval_locked(c, a, b, max_limbs, nullptr, &plan, rv);
  if (rc) return rc;
  pcg_dfloat* out[3] = {P, Q, R};
  for (int j = 0; j < 3; ++j) {
    if (rv[j].L == 0) {
      df_assign_zero(out[j]);
      continue;
    }
    uint32_t* buf = nullptr;
    uint64_t cap = 0;
    rc = df_alloc(c, rv[j].L, &buf, &cap);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(buf, plan->arena + rv[j].off, rv[j].L * 4, hipMemcpyDeviceToDevice,
                          c.stream));
    df_assign(out[j], buf, cap, rv[j].L, 0, rv[j].sign);
  }
  return PCG_OK;
}

int pcg_bsr_graphs(int enable) {
  DF_ENTER();
  g_bsr_graphs = enable != 0;
  if (!g_bsr_graphs)
    for (BsrPlan& q : g_bsr_plans) bsr_plan_drop_graph(q);
  return PCG_OK;
}

int pcg_bsr_prepare(const uint32_t* sizes, uint64_t n) {
  if (!sizes && n) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  int rc = ensure_tuned(c);  // plans record the small-product classes
  if (rc) return rc;
  // the plan cache holds two sizes: the first two distinct ones
  std::vector<uint32_t> want;
  for (uint64_t i = 0; i < n && want.size() < 2; ++i)
    if (sizes[i] && std::find(want.begin(), want.end(), sizes[i]) == want.end()) want.push_back(sizes[i]);
  // laid out with the bounds of any leaf, so their buffers hold every later
  // plan of these sizes
  for (uint32_t d : want) {
    BsrPlan* P = nullptr;
    rc = bsr_plan_get(c, d, leaf_bits_for(0), &P);
    if (rc) return rc;
  }
  return PCG_OK;
}

int pcg_bsr_plan_check(uint32_t leaves, uint32_t bmax, const int* table, int n) {
  if (!table || n != kSmallNttMaxLog + 1) return fail(PCG_ERR_INVALID, "the class table has %d entries", kSmallNttMaxLog + 1);
  SmallTable tab;
  for (int b = 0; b < n; ++b) {
    if (!small_class_valid(b, table[b])) return fail(PCG_ERR_INVALID, "class %d is not valid for bucket %d", table[b], b);
    tab[b] = table[b];
  }
  return bsr_plan_check(leaves, leaf_bits_for(bmax), tab);
}

int pcg_autotune(int* table, int n) {
  DF_ENTER();
  HIPCHK(hipStreamSynchronize(c.stream));
  for (BsrPlan& q : g_bsr_plans) bsr_plan_free(q);  // plans record the classes
  int rc = tune_small(c);
  if (rc) return rc;
  for (int b = 0; b < n && b <= kGeomMaxLog; ++b)
    table[b] = b <= kSmallNttMaxLog ? g_small_class[b] : g_batch_geom[b];
  return PCG_OK;
}

int pcg_batch_table(int* table, int n, int set) {
  if (!table && n > 0) return fail(PCG_ERR_INVALID, "null pointer");
  DF_ENTER();
  if (set) {
    for (int b = 0; b < n && b <= kGeomMaxLog; ++b)
      if (b <= kSmallNttMaxLog ? !small_class_valid(b, table[b]) : !batch_geom_valid(b, table[b]))
        return fail(PCG_ERR_INVALID, "class %d is not valid for bucket %d", table[b], b);
    HIPCHK(hipStreamSynchronize(c.stream));
    for (BsrPlan& q : g_bsr_plans) bsr_plan_free(q);
    for (int b = 0; b < n && b <= kGeomMaxLog; ++b) {
      if (b <= kSmallNttMaxLog)
        g_small_class[b] = table[b];
      else
        g_batch_geom[b] = table[b];
    }
    g_small_ready = true;
    ++g_table_gen;
    return PCG_OK;
  }
  for (int b = 0; b < n && b <= kGeomMaxLog; ++b)
    table[b] = b <= kSmallNttMaxLog ? g_small_class[b] : g_batch_geom[b];
  return PCG_OK;
}

int pcg_df_release_pool(void) {
  DF_ENTER();
  return df_release_pool(c);
}

}  // extern "C"
