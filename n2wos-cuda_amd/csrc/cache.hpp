// cache.hpp -- learned cache: multiresolution hash grid + ReLU MLP.
//
// Encoding (Mueller, Evans, Schied, Keller 2022, "Instant Neural Graphics
// Primitives with a Multiresolution Hash Encoding"): L levels of F = 2
// features; level l has N_l = floor(N_min * b^l) cells per axis over the unit
// cube, i.e. (N_l + 1)^3 vertices.  The vertices are grouped in bricks of
// 2 x 2 x 2 (brick (x/2, y/2, z/2), 8 entries = 64 bytes, one memory request
// for a gradient scatter and at most one cache line per brick a lookup
// touches): vertex (x, y, z) is entry 8 brick + (x & 1 | (y & 1) << 1 |
// (z & 1) << 2).  Levels whose B^3 bricks (B = N_l/2 + 1) fit in T entries
// index bricks densely (bx + B by + B^2 bz), the rest by the spatial hash of
// the brick, (bx * 1 ^ by * 2654435761 ^ bz * 805459861) mod T/8.  Features
// are trilinearly interpolated from the 8 cell corners.  Tables are FP32.
//
// MLP: 2L = 32 inputs -> `layers` hidden layers of width H (ReLU) -> 1 output,
// no biases (the layout of a fully fused MLP); H in {16, 32, 64}, layers in
// 1..4.  Weights are FP32 masters with an FP16 copy for the MFMA kernels
// (FP32 accumulation).  Layout of the weight buffer (row-major, out x in):
// A1 [H x 32] at 0, hidden layer k = 1 .. layers-1 [H x H] at
// 32 H + (k - 1) H^2, the output block [16 x H] at 32 H + (layers - 1) H^2,
// whose row 0 is the output layer and rows 1..15 are zero (MFMA tile
// padding).  The default 4 x 64 network has 15,360 weights.
#pragma once

#include <cstdint>

namespace n2 {

constexpr int kCacheLevels = 16;
constexpr int kCacheFeat = 2;
constexpr int kCacheIn = kCacheLevels * kCacheFeat;  // 32
constexpr int kCacheMaxHidden = 64;
constexpr int kCacheMaxLayers = 4;

constexpr int cache_weights(int h, int layers) { return kCacheIn * h + (layers - 1) * h * h + 16 * h; }
// first weight of hidden layer k (1 .. layers-1; its input is activation k)
constexpr int cache_off_hidden(int h, int k) { return kCacheIn * h + (k - 1) * h * h; }
constexpr int cache_off_out(int h, int layers) { return kCacheIn * h + (layers - 1) * h * h; }

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
