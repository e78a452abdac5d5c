// Buffer-resource loads and stores for gfx950 (raw buffers, no swizzle).
//
// A plane (one channel or kernel spectrum) is addressed through a buffer resource built from
// its uniform base pointer and size: the per-thread part of an address is a 32-bit VGPR
// offset and the per-element part a compile-time SGPR offset, so an unrolled transform's
// loads and stores need no 64-bit address arithmetic at all.  Accesses at or beyond the
// resource's size are dropped (loads return 0) instead of reaching other memory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cn {

#ifdef __HIPCC__
using BufRes = __amdgpu_buffer_rsrc_t;
constexpr int kBufWord3 = 0x00020000;  // gfx9 family: 32-bit data format, raw addressing

__device__ __forceinline__ BufRes make_res(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kBufWord3);
}

__device__ __forceinline__ float buf_load_f32(BufRes r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

__device__ __forceinline__ float2 buf_load_c64(BufRes r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

__device__ __forceinline__ void buf_store_c64(float2 v, BufRes r, uint32_t voff, uint32_t soff) {
  using u2 = __attribute__((ext_vector_type(2))) unsigned int;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
}
#endif

}  // namespace cn
