// Range-checked loads of 32-bit limbs through a raw buffer resource
// (device side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcg {

// Word 3 of a raw buffer resource on gfx9-family GPUs: 32-bit data format,
// no swizzle, no thread-id offset.
constexpr int kBufWord3 = 0x00020000;

// Raw buffer view of n 32-bit limbs at a wave-uniform address.  A load past
// the n limbs returns 0 (the buffer unit's range check, no memory access),
// so a ragged operand loads without a select or a branch per limb and a
// thread's loads all stay in flight together (a select after each flat
// load had the compiler wait for every load in turn).  n < 2^29.
struct LimbBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ LimbBuf(const uint32_t* p, uint32_t n)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(p), (short)0, (int)(n * 4u), kBufWord3)) {}
  __device__ __forceinline__ uint32_t operator[](uint32_t i) const {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)(i * 4u), 0, 0);
  }
};

}  // namespace pcg
