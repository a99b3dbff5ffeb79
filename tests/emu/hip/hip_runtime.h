// Host stand-in for <hip/hip_runtime.h> used ONLY by the CPU SIMT emulator of
// the walker kernel (tests/emu/walker_emu.cpp).  It provides the qualifiers,
// vector types and atomics the device code uses, as plain C++.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#define __host__
#define __device__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static

struct float4 {
  float x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
struct uint4 {
  uint32_t x, y, z, w;
};
struct uint2 {
  uint32_t x, y;
};
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }

inline int __float_as_int(float f) {
  int i;
  std::memcpy(&i, &f, 4);
  return i;
}
inline uint32_t __float_as_uint(float f) {
  uint32_t i;
  std::memcpy(&i, &f, 4);
  return i;
}
inline float __uint_as_float(uint32_t i) {
  float f;
  std::memcpy(&f, &i, 4);
  return f;
}
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll(static_cast<long long>(x)); }

template <class T>
inline T atomicAdd(T* p, T v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
template <class T>
inline T atomicSub(T* p, T v) {
  return __atomic_fetch_sub(p, v, __ATOMIC_SEQ_CST);
}

using std::isfinite;
using std::isnan;
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
typedef void* hipStream_t;
typedef int hipError_t;

struct EmuDim3 {
  unsigned x;
};
extern EmuDim3 threadIdx;
extern EmuDim3 blockIdx;
