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

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/picugpu.h"
#include "carry_kernels.hpp"
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

struct Plan {
  int logN;
  int logE;
  std::vector<Pass> passes;
};

Plan make_plan(int logN) {
  Plan pl;
  pl.logN = logN;
  pl.logE = logN < kLogE ? logN : kLogE;
  int m = (logN + kLogE - 1) / kLogE;
  if (m < 1) m = 1;
  const int base = logN / m, extra = logN % m;
  int rem = logN;
  for (int i = 0; i < m; ++i) {
    Pass ps;
    ps.logL = base + (i < extra ? 1 : 0);
    rem -= ps.logL;
    ps.logS = rem;
    ps.logLines = pl.logE - ps.logL;
    if (ps.logS >= ps.logLines) {
      ps.logG = ps.logLines;
      ps.logB = 0;
    } else {
      ps.logG = ps.logS;
      ps.logB = ps.logLines - ps.logS;
    }
    pl.passes.push_back(ps);
  }
  return pl;
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
  CarryFn* agg = nullptr;
  uint64_t* cin = nullptr;
  size_t nblk_cap = 0;
  uint32_t* status = nullptr;
  // staging for host-pointer calls
  uint32_t* d_stage = nullptr;
  size_t stage_limbs = 0;
  // profiling
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double prof_ms[3] = {0, 0, 0};
  double prof_total = 0;
  bool prof_valid = false;
  double fft_maxdist = 0;
  int fft_recomputed = 0;
};

Ctx g_ctx;

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
      }
      const uint64_t ninv = powmod_u64(powmod_u64(2, logN, p), p - 2, p);
      P.ninv = shoup_pair(ninv, p);
      P.twL = c.tw[pi][0];
      P.twHi = c.tw[pi][1];
      P.twLo = c.tw[pi][2];
    }
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
    if (c.spec) HIPCHK(hipFree(c.spec));
    c.spec = nullptr;
    c.spec_points = 0;
    if (hipMalloc(&c.spec, pts * 4 * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PCG_ERR_OOM, "cannot allocate %zu bytes of transform scratch", pts * 32);
    }
    c.spec_points = pts;
  }
  const size_t nblk = (zL + kCarryBlock - 1) / kCarryBlock;
  if (nblk > c.nblk_cap) {
    if (c.agg) HIPCHK(hipFree(c.agg));
    if (c.cin) HIPCHK(hipFree(c.cin));
    c.agg = nullptr;
    c.cin = nullptr;
    c.nblk_cap = 0;
    HIPCHK(hipMalloc(&c.agg, nblk * sizeof(CarryFn)));
    HIPCHK(hipMalloc(&c.cin, nblk * sizeof(uint64_t)));
    c.nblk_cap = nblk;
  }
  return PCG_OK;
}

int ensure_stage(Ctx& c, size_t limbs) {
  if (limbs <= c.stage_limbs) return PCG_OK;
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

void launch_pass(Ctx& c, int mode, const Plan& pl, const Pass& ps, const double* src,
                 const uint32_t* limbs, uint64_t nlimbs, double* dst, const double* spec) {
  PassArgs a;
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
  const int E = 1 << pl.logE;
  dim3 grid((unsigned)(a.N / E), 2), block(E / kEPT);
  if (mode == kFwd)
    hipLaunchKernelGGL(k_ntt_pass<kFwd>, grid, block, 0, c.stream, a);
  else if (mode == kInv)
    hipLaunchKernelGGL(k_ntt_pass<kInv>, grid, block, 0, c.stream, a);
  else
    hipLaunchKernelGGL(k_ntt_pass<kFused>, grid, block, 0, c.stream, a);
}

// Enqueue the whole NTT product on the library stream (device pointers).
int ntt_mul_enqueue(Ctx& c, const uint32_t* d_a, uint64_t na, const uint32_t* d_b, uint64_t nb,
                    uint32_t* d_out, bool profile) {
  const uint64_t zL = na + nb;
  int logN = ceil_log2(zL);
  if (logN < kMinLogN) logN = kMinLogN;
  int rc = ensure_workspace(c, logN, zL);
  if (rc) return rc;
  const Plan pl = make_plan(logN);
  const uint64_t N = 1ull << logN;
  double* A = c.spec;
  double* B = c.spec + 2 * N;
  const size_t m = pl.passes.size();
  if (profile) HIPCHK(hipEventRecord(c.ev[0], c.stream));
  for (size_t j = 0; j < m; ++j)
    launch_pass(c, kFwd, pl, pl.passes[j], j == 0 ? nullptr : A, d_a, na, A, nullptr);
  if (profile) HIPCHK(hipEventRecord(c.ev[1], c.stream));
  for (size_t j = 0; j + 1 < m; ++j)
    launch_pass(c, kFwd, pl, pl.passes[j], j == 0 ? nullptr : B, d_b, nb, B, nullptr);
  launch_pass(c, kFused, pl, pl.passes[m - 1], m == 1 ? nullptr : B, d_b, nb, B, A);
  for (size_t j = m - 1; j-- > 0;) launch_pass(c, kInv, pl, pl.passes[j], B, nullptr, 0, B, nullptr);
  if (profile) HIPCHK(hipEventRecord(c.ev[2], c.stream));
  CarryArgs ca;
  ca.prm = c.d_prime + logN * 2;
  ca.r = B;
  ca.N = N;
  ca.zL = zL;
  ca.out = d_out;
  ca.agg = c.agg;
  ca.cin = c.cin;
  ca.nblk = (zL + kCarryBlock - 1) / kCarryBlock;
  ca.crt = c.crt;
  ca.status = c.status;
  hipLaunchKernelGGL(k_carry_local, dim3((unsigned)ca.nblk), dim3(kCarryThreads), 0, c.stream, ca);
  hipLaunchKernelGGL(k_carry_scan, dim3(1), dim3(1024), 0, c.stream, ca);
  hipLaunchKernelGGL(k_carry_fix, dim3((unsigned)((ca.nblk + 255) / 256)), dim3(256), 0, c.stream,
                     ca);
  if (profile) HIPCHK(hipEventRecord(c.ev[3], c.stream));
  HIPCHK(hipGetLastError());
  return PCG_OK;
}

int mul_dev_locked(Ctx& c, const uint32_t* d_a, uint64_t na, const uint32_t* d_b, uint64_t nb,
                   uint32_t* d_out, uint32_t flags) {
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
  if (eng == PCG_ENGINE_FFT) return fail(PCG_ERR_INVALID, "FFT engine not available in this build");
  const bool prof = (flags & PCG_PROFILE) != 0;
  rc = ntt_mul_enqueue(c, d_a, na, d_b, nb, d_out, prof);
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

int sync_locked(Ctx& c) {
  HIPCHK(hipStreamSynchronize(c.stream));
  uint32_t st = 0;
  HIPCHK(hipMemcpy(&st, c.status, sizeof st, hipMemcpyDeviceToHost));
  if (st) {
    uint32_t z = 0;
    HIPCHK(hipMemcpy(c.status, &z, sizeof z, hipMemcpyHostToDevice));
    return fail(PCG_ERR_INTERNAL, "carry propagation overflowed (status %u)", st);
  }
  return PCG_OK;
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
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return fail(PCG_ERR_NO_DEVICE, "no HIP device available");
  }
  if (device < 0 || device >= n) return fail(PCG_ERR_NO_DEVICE, "device %d out of range (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  int rc = build_tables(c);
  if (rc) return rc;
  HIPCHK(hipMalloc(&c.status, sizeof(uint32_t)));
  HIPCHK(hipMemset(c.status, 0, sizeof(uint32_t)));
  for (int i = 0; i < 4; ++i) HIPCHK(hipEventCreate(&c.ev[i]));
  c.device = device;
  c.ready = true;
  return PCG_OK;
}

void pcg_shutdown(void) {
  Ctx& c = g_ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready) return;
  (void)hipStreamSynchronize(c.stream);
  for (int pi = 0; pi < 2; ++pi)
    for (int t = 0; t < 3; ++t) (void)hipFree(c.tw[pi][t]);
  (void)hipFree(c.d_prime);
  (void)hipFree(c.spec);
  (void)hipFree(c.agg);
  (void)hipFree(c.cin);
  (void)hipFree(c.status);
  (void)hipFree(c.d_stage);
  for (int i = 0; i < 4; ++i) (void)hipEventDestroy(c.ev[i]);
  (void)hipStreamDestroy(c.stream);
  Ctx fresh_defaults;
  c.ready = false;
  c.device = -1;
  c.stream = nullptr;
  c.spec = nullptr;
  c.spec_points = 0;
  c.agg = nullptr;
  c.cin = nullptr;
  c.nblk_cap = 0;
  c.status = nullptr;
  c.d_stage = nullptr;
  c.stage_limbs = 0;
  c.d_prime = nullptr;
  c.prof_valid = false;
  (void)fresh_defaults;
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
  for (uint64_t i = 0; i < count; ++i) {
    int rc = mul_dev_locked(c, d[i].a, d[i].na, d[i].b, d[i].nb, d[i].out, flags & ~PCG_PROFILE);
    if (rc) return rc;
  }
  return sync_locked(c);
}

int pcg_mul_batch(const pcg_mul_desc* d, uint64_t count, uint32_t flags) {
  if (count && !d) return fail(PCG_ERR_INVALID, "null descriptor array");
  for (uint64_t i = 0; i < count; ++i) {
    int rc = pcg_mul(d[i].a, d[i].na, d[i].b, d[i].nb, d[i].out, flags & ~PCG_PROFILE);
    if (rc) return rc;
  }
  return PCG_OK;
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

int pcg_ntt_plan(uint64_t total_limbs, int* log2_points, int* passes) {
  if (total_limbs > PCG_MAX_TOTAL_LIMBS) return fail(PCG_ERR_TOO_LARGE, "FFT length is too small.");
  int logN = ceil_log2(total_limbs);
  if (logN < kMinLogN) logN = kMinLogN;
  if (log2_points) *log2_points = logN;
  if (passes) *passes = (int)make_plan(logN).passes.size();
  return PCG_OK;
}

}  // extern "C"
