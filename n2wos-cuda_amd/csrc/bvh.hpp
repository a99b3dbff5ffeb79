// bvh.hpp -- the flattened 4-wide BVH used by every kernel.
//
// Layout in HBM (all arrays 256-byte aligned, built once on the host):
//   nodes:  Bvh4Node[node_count], 128 B each = one cache line.  The six
//           child-box planes are stored as float4 rows (lox, loy, loz, hix,
//           hiy, hiz), so one node is seven 16-byte loads per lane.
//   tris:   TriRec[triangle_count], 48 B each, in leaf order: a leaf is a
//           contiguous run [first, first + count).  Each record keeps the
//           three vertices and the triangle's original mesh index.
// Child references: bit 31 clear = inner node index; bit 31 set = leaf with
// first = (ref >> 5) & 0x3ffffff and count = ref & 31 (count 0 = empty slot,
// whose box is inverted so it is never visited).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace n2 {

struct alignas(16) Bvh4Node {
  float lox[4], loy[4], loz[4];
  float hix[4], hiy[4], hiz[4];
  uint32_t child[4];
  uint32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "node must be one 128-byte line");

struct alignas(16) TriRec {
  float ax, ay, az, bx;
  float by, bz, cx, cy;
  float cz;
  int32_t id;
  int32_t pad0, pad1;
};
static_assert(sizeof(TriRec) == 48, "triangle record must be 48 bytes");

constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr int kMaxLeafSize = 31;

struct HostBvh4 {
  std::vector<Bvh4Node> nodes;
  std::vector<TriRec> tris;
  int leaf_size = 8;
  int binary_depth = 0;   // depth of the median-split binary tree
  int depth = 0;          // depth of the 4-wide tree (root = 0)
  int binary_nodes = 0;
  float delta = 0.0f;     // absolute pruning margin, 2^-18 * max(1, max |coord|)
};

// Builds the binary median-split tree (largest centroid-extent axis, median by
// nth_element, leaves of at most leaf_size triangles -- the reference builder
// src/cuda/bvh_build_host.cpp:74-120) and collapses it into 4-wide nodes by
// repeatedly opening the child with the largest surface area.
HostBvh4 build_bvh4(const float* verts, int nv, const uint32_t* tris, int nt, int leaf_size, bool sah = false);

}  // namespace n2
