// wave.hpp -- wave-level helpers of the persistent kernels (64-lane waves).
// The CPU SIMT emulator used by the tests defines N2_EMULATED_WAVE and
// supplies its own versions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace n2 {

#ifndef N2_EMULATED_WAVE
__device__ __forceinline__ int lane_id() {
  return static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)));
}

// number of set bits of `mask` below this lane
__device__ __forceinline__ int rank_in(uint64_t mask) {
  return static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u)));
}

// v of lane src mod 64 of this wave, one ds_bpermute: the whole wave
// executes it, and that lane is active
__device__ __forceinline__ uint32_t lane_read(uint32_t v, int src) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_ds_bpermute(static_cast<int>((static_cast<uint32_t>(src) & 63u) << 2), static_cast<int>(v)));
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// This wave's global stores so far are complete (in L2) before its later
// loads; with load_l2, which reads past the CU's vector L1, those loads see them.
__device__ __forceinline__ void wave_global_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ float load_l2(const float* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The XCD (accelerator complex die) this wave runs on: hardware register
// XCC_ID (20) on gfx94x / gfx950, 0-7 on MI355X.
__device__ __forceinline__ int xcd_id() {
  return static_cast<int>(__builtin_amdgcn_s_getreg((31 << 11) | (0 << 6) | 20));
}

// The hardware square root (V_SQRT_F32): one instruction, not correctly
// rounded (within 1 ulp for normal inputs), against about 15 for sqrtf.
// Only for bounds that carry a margin, never for values the results depend on.
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// The same value, through a copy the compiler cannot see through: a use
// after it cannot be hoisted out of a loop, so a loop-invariant value that is
// needed only on a rare path is not held in registers for the whole loop.
__device__ __forceinline__ uint32_t opaque_u32(uint32_t v) {
  asm volatile("" : "+s"(v));
  return v;
}
__device__ __forceinline__ int opaque_lane_int(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// A kernel's by-value argument block, read in place from the kernel-argument
// segment (it starts there) through an opaque pointer: each use is a scalar
// load where it stands, for arguments needed only on rare paths.
template <class T>
using kernarg_ptr = const T __attribute__((address_space(4)))*;
template <class T>
__device__ __forceinline__ kernarg_ptr<T> kernel_args() {
  kernarg_ptr<T> p = (kernarg_ptr<T>)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return p;
}
#endif

__device__ __forceinline__ float lane_read_f(float v, int src) {
  return __uint_as_float(lane_read(__float_as_uint(v), src));
}

}  // namespace n2
