// Compute-peak probes for the NTT roofline (measurement only, not on the
// product path): the rate of the engine's FP64 modular multiply (mm_shoup of
// modfp64.hpp, six FP64 operations) and of plain FP64 fused multiply-adds,
// both register-resident, on every compute unit.  bench.py reports the NTT
// multiply's mulmods per second as a fraction of the first.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "modfp64.hpp"

#pragma clang fp contract(off)

using namespace pcg;

namespace {

constexpr int kChains = 8;   // independent dependency chains per thread
constexpr int kThreads = 256;

// Each thread runs kChains independent chains of `iters` mulmods; the
// operands stay below 2^47 (|r| < p), as inside the NTT stages.
__global__ __launch_bounds__(kThreads) void k_mulmod_rate(double* sink, int iters, double w, double wp,
                                                          double p) {
  double x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = (double)(threadIdx.x * kChains + c + 1 + blockIdx.x);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = mm_shoup(x[c], w, wp, p);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  if (s == 12345.678) sink[blockIdx.x] = s;  // keeps the chains live
}

__global__ __launch_bounds__(kThreads) void k_fma_rate(double* sink, int iters, double a, double b) {
  double x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = (double)(threadIdx.x + c);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = __builtin_fma(x[c], a, b);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  if (s == 12345.678) sink[blockIdx.x] = s;
}

template <class F>
int time_kernel(F launch, double* ms_out) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return 1;
  if (hipEventCreate(&e1) != hipSuccess) return 1;
  launch();  // warm-up
  int rc = hipDeviceSynchronize() != hipSuccess;
  float best = 1e30f;
  for (int r = 0; r < 3 && !rc; ++r) {
    float ms = 0;
    rc |= hipEventRecord(e0, nullptr) != hipSuccess;
    launch();
    rc |= hipEventRecord(e1, nullptr) != hipSuccess;
    rc |= hipEventSynchronize(e1) != hipSuccess;
    rc |= hipEventElapsedTime(&ms, e0, e1) != hipSuccess;
    if (ms < best) best = ms;
  }
  rc |= hipEventDestroy(e0) != hipSuccess;
  rc |= hipEventDestroy(e1) != hipSuccess;
  *ms_out = best;
  return rc;
}

}  // namespace

extern "C" {

// mulmods per second (mm_shoup) and FP64 FMA per second over the whole GPU
// (device 0, 16 workgroups of 256 threads per compute unit).  0 on success.
int pcg_probe_fp64_rates(double* mulmods_per_s, double* fma_per_s, int* compute_units) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 1;
  const int cu = prop.multiProcessorCount;
  const unsigned blocks = (unsigned)cu * 16u;
  double* sink = nullptr;
  if (hipMalloc(&sink, blocks * sizeof(double)) != hipSuccess) return 1;
  const uint64_t p = kP0;
  const uint64_t w = 123456789012345ull % p;
  const double wd = (double)w, wpd = (double)w / (double)p, pd = (double)p;
  const int iters = 4096;
  double ms_mm = 0, ms_fma = 0;
  int rc = time_kernel([&] { hipLaunchKernelGGL(k_mulmod_rate, dim3(blocks), dim3(kThreads), 0, nullptr, sink, iters, wd, wpd, pd); },
                       &ms_mm);
  if (!rc)
    rc = time_kernel([&] { hipLaunchKernelGGL(k_fma_rate, dim3(blocks), dim3(kThreads), 0, nullptr, sink, iters, 1.0000001, 1e-9); },
                     &ms_fma);
  (void)hipFree(sink);
  if (rc) return rc;
  const double ops = (double)blocks * kThreads * kChains * iters;
  *mulmods_per_s = ops / (ms_mm * 1e-3);
  *fma_per_s = ops / (ms_fma * 1e-3);
  *compute_units = cu;
  return 0;
}

}  // extern "C"
