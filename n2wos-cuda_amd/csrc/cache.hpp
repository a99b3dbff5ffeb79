// cache.hpp -- learned cache: multiresolution hash grid + ReLU MLP.
//
// Encoding (Mueller, Evans, Schied, Keller 2022, "Instant Neural Graphics
// Primitives with a Multiresolution Hash Encoding"): L levels of F = 2
// features; level l has N_l = floor(N_min * b^l) cells per axis over the unit
// cube, i.e. (N_l + 1)^3 vertices; levels whose vertex count fits in T entries
// are indexed densely (x + y (N_l+1) + z (N_l+1)^2), the rest by the spatial
// hash (x * 1 ^ y * 2654435761 ^ z * 805459861) mod T.  Features are
// trilinearly interpolated from the 8 cell corners.  Tables are FP32.
//
// MLP: 2L = 32 inputs -> 4 hidden layers of 64 (ReLU) -> 1 output, no biases
// (the layout of a fully fused MLP).  Weights are FP32 masters with an FP16
// copy for the MFMA kernels (FP32 accumulation).  Layout of the weight buffer
// (row-major, out x in): A1 [64 x 32], A2..A4 [64 x 64], A5 [16 x 64] whose
// row 0 is the output layer and rows 1..15 are zero (MFMA tile padding).
#pragma once

#include <cstdint>

namespace n2 {

constexpr int kCacheLevels = 16;
constexpr int kCacheFeat = 2;
constexpr int kCacheIn = kCacheLevels * kCacheFeat;  // 32
constexpr int kCacheHidden = 64;
constexpr int kCacheW1 = kCacheHidden * kCacheIn;            // 2048
constexpr int kCacheWh = kCacheHidden * kCacheHidden;        // 4096
constexpr int kCacheW5 = 16 * kCacheHidden;                  // 1024 (row 0 live)
constexpr int kCacheWeights = kCacheW1 + 3 * kCacheWh + kCacheW5;  // 15360
constexpr int kCacheOffA2 = kCacheW1;
constexpr int kCacheOffA3 = kCacheW1 + kCacheWh;
constexpr int kCacheOffA4 = kCacheW1 + 2 * kCacheWh;
constexpr int kCacheOffA5 = kCacheW1 + 3 * kCacheWh;

struct GridLayout {
  int n_levels;
  uint32_t table_size;        // T (power of two)
  uint32_t offset[kCacheLevels];  // first entry of each level
  uint32_t size[kCacheLevels];    // entries of each level
  int res[kCacheLevels];          // N_l
  int dense[kCacheLevels];
  uint32_t total_entries;
  float in_min[3];                // world -> unit cube: (p - in_min) * in_scale, clamped
  float in_scale[3];
};

}  // namespace n2
