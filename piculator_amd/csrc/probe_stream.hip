// HBM access-pattern probe for the large-tile NTT passes (measurement only,
// not on the product path).  It moves a 2^25-point transform of two primes
// (2 x 2^25 doubles, 512 MiB) from one buffer to another in tiles of 16384
// doubles, with the passes' thread shape: 1024 threads x 16 doubles, one
// workgroup per compute unit, the next tile's loads issued before the
// current tile's stores, tiles ranked XCD by XCD.  A side of the copy is
// either contiguous (a row tile: 16 coalesced 8 KB rows) or strided (a
// column tile: C adjacent columns of a 2^25-point matrix with rows of
// 2^25 / (16384 / C) doubles, i.e. runs of 8 C bytes).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kTile = 16384;
constexpr int kThreads = 1024;

template <int COLS>
__device__ __forceinline__ uint64_t strided_at(uint64_t T, int t, int r) {
  constexpr uint64_t R = kTile / COLS;       // rows of a column tile
  constexpr uint64_t C = (1ull << 25) / R;   // columns of the matrix
  const uint64_t pr = T >> 11, tc = T & 2047;  // 2048 tiles per prime
  const int lam = t % COLS, k = t / COLS + (kThreads / COLS) * r;
  return (pr << 25) + (uint64_t)k * C + tc * COLS + lam;
}
__device__ __forceinline__ uint64_t contiguous_at(uint64_t T, int t, int r) {
  return T * kTile + (uint64_t)r * kThreads + t;
}
// COLS 0: the contiguous side; otherwise the strided one
template <int COLS>
__device__ __forceinline__ uint64_t side_at(uint64_t T, int t, int r) {
  if constexpr (COLS == 0)
    return contiguous_at(T, t, r);
  else
    return strided_at<COLS>(T, t, r);
}

template <int RCOLS, int WCOLS>
__global__ __launch_bounds__(kThreads) void k_stream(const double* __restrict__ src, double* __restrict__ dst,
                                                     uint64_t ntiles) {
  const uint64_t g = gridDim.x;
  uint64_t w = blockIdx.x;
  if ((g & 7) == 0) w = (uint64_t)(blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
  if (w >= ntiles) return;
  const int t = threadIdx.x;
  double x[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) x[r] = src[side_at<RCOLS>(w, t, r)];
  for (;;) {
    double v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = x[r];
    const uint64_t wn = w + g;
    if (wn < ntiles) {
#pragma unroll
      for (int r = 0; r < 16; ++r) x[r] = src[side_at<RCOLS>(wn, t, r)];
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) dst[side_at<WCOLS>(w, t, r)] = v[r];
    w = wn;
    if (w >= ntiles) return;
  }
}

using StreamFn = void (*)(const double*, double*, uint64_t);
StreamFn stream_fn(int rcols, int wcols) {
#define PCG_PAT(R, W) \
  if (rcols == R && wcols == W) return &k_stream<R, W>;
  PCG_PAT(0, 0) PCG_PAT(0, 8) PCG_PAT(8, 0) PCG_PAT(8, 8) PCG_PAT(0, 16) PCG_PAT(16, 0) PCG_PAT(16, 16)
  PCG_PAT(32, 32) PCG_PAT(4, 4)
#undef PCG_PAT
  return nullptr;
}

}  // namespace

extern "C" {

// Best of five timed copies (after one warm-up) of the pattern with read
// side rcols and write side wcols (0 = contiguous, else adjacent columns per
// tile: 4, 8, 16 or 32; see stream_fn).  Returns 0 and the time in
// microseconds and the read + write rate in bytes per second; nonzero on a
// HIP error or an unknown pattern.
int pcg_probe_stream(int rcols, int wcols, double* us, double* bytes_per_s) {
  StreamFn fn = stream_fn(rcols, wcols);
  if (!fn) return 2;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
  const uint64_t n = 1ull << 26;  // two primes x 2^25 points
  const uint64_t ntiles = n / kTile;
  double *s = nullptr, *d = nullptr;
  if (hipMalloc(&s, n * sizeof(double)) != hipSuccess) return 1;
  if (hipMalloc(&d, n * sizeof(double)) != hipSuccess) {
    (void)hipFree(s);
    return 1;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess;
  float best = 1e30f;
  for (int rep = 0; rep < 6 && !rc; ++rep) {
    rc |= hipEventRecord(e0, nullptr) != hipSuccess;
    hipLaunchKernelGGL(fn, dim3((unsigned)prop.multiProcessorCount), dim3(kThreads), 0, nullptr, s, d, ntiles);
    rc |= hipGetLastError() != hipSuccess;
    rc |= hipEventRecord(e1, nullptr) != hipSuccess;
    rc |= hipEventSynchronize(e1) != hipSuccess;
    float ms = 0;
    rc |= hipEventElapsedTime(&ms, e0, e1) != hipSuccess;
    if (rep > 0 && ms < best) best = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(s);
  (void)hipFree(d);
  if (rc) return 1;
  *us = best * 1e3;
  *bytes_per_s = 2.0 * (double)n * sizeof(double) / (best * 1e-3);
  return 0;
}

}  // extern "C"
