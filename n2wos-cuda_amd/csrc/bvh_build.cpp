// bvh_build.cpp -- host construction of the 4-wide BVH (see bvh.hpp): a
// binary tree split at the centroid median of the largest extent (the
// reference's build, src/cuda/bvh_build_host.cpp:74-120) or by a binned
// surface-area heuristic, collapsed to 4-wide nodes.
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

#include "bvh.hpp"

namespace n2 {
namespace {

struct Box {
  float lo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                 std::numeric_limits<float>::infinity()};
  float hi[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                 -std::numeric_limits<float>::infinity()};
  void grow(const float* p) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  void grow(const Box& b) {
    grow(b.lo);
    grow(b.hi);
  }
  float area() const {
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    return ex * ey + ey * ez + ez * ex;
  }
};

struct Prim {
  Box box;
  float cen[3];
};

struct BinNode {
  Box box;
  int left = -1, right = -1;
  int first = 0, count = 0;
};

struct Binary {
  std::vector<BinNode> nodes;
  std::vector<int> order;  // leaf order of original triangle ids
  int depth = 0;
};

// Binned surface-area-heuristic split (32 bins per axis over the centroid
// box): minimises area(L) n(L) + area(R) n(R); partitions order[begin, end)
// and returns the split index, or -1 when no bin boundary separates the
// centroids (the caller then splits at the median).
int sah_partition(std::vector<int>& order, const std::vector<Prim>& prims, int begin, int end, const Box& cbox) {
  constexpr int kBins = 32;
  float best_cost = std::numeric_limits<float>::infinity();
  int best_axis = -1, best_bin = -1;
  for (int a = 0; a < 3; ++a) {
    const float lo = cbox.lo[a], ext = cbox.hi[a] - cbox.lo[a];
    if (!(ext > 0.0f)) continue;
    Box bb[kBins];
    int bn[kBins] = {0};
    const float k = kBins / ext;
    for (int i = begin; i < end; ++i) {
      const Prim& p = prims[static_cast<size_t>(order[static_cast<size_t>(i)])];
      const int b = std::min(kBins - 1, static_cast<int>((p.cen[a] - lo) * k));
      bn[b] += 1;
      bb[b].grow(p.box);
    }
    float right_area[kBins];
    int right_n[kBins];
    Box acc;
    int n = 0;
    for (int b = kBins - 1; b > 0; --b) {
      if (bn[b]) acc.grow(bb[b]);
      n += bn[b];
      right_area[b] = n ? acc.area() : 0.0f;
      right_n[b] = n;
    }
    Box lacc;
    int ln = 0;
    for (int b = 0; b < kBins - 1; ++b) {
      if (bn[b]) lacc.grow(bb[b]);
      ln += bn[b];
      const int rn = right_n[b + 1];
      if (ln == 0 || rn == 0) continue;
      const float cost = lacc.area() * static_cast<float>(ln) + right_area[b + 1] * static_cast<float>(rn);
      if (cost < best_cost) {
        best_cost = cost;
        best_axis = a;
        best_bin = b;
      }
    }
  }
  if (best_axis < 0) return -1;
  const float lo = cbox.lo[best_axis], k = kBins / (cbox.hi[best_axis] - cbox.lo[best_axis]);
  auto it = std::partition(order.begin() + begin, order.begin() + end, [&](int id) {
    const Prim& p = prims[static_cast<size_t>(id)];
    return std::min(kBins - 1, static_cast<int>((p.cen[best_axis] - lo) * k)) <= best_bin;
  });
  const int mid = static_cast<int>(it - order.begin());
  return (mid > begin && mid < end) ? mid : -1;
}

int build_binary(Binary& out, const std::vector<Prim>& prims, int begin, int end, int depth, int leaf_size,
                 bool sah) {
  out.depth = std::max(out.depth, depth);
  Box box, cbox;
  for (int i = begin; i < end; ++i) {
    const Prim& p = prims[static_cast<size_t>(out.order[static_cast<size_t>(i)])];
    box.grow(p.box);
    cbox.grow(p.cen);
  }
  const int idx = static_cast<int>(out.nodes.size());
  out.nodes.push_back(BinNode{});
  out.nodes[static_cast<size_t>(idx)].box = box;
  const int count = end - begin;
  if (count <= leaf_size) {
    out.nodes[static_cast<size_t>(idx)].first = begin;
    out.nodes[static_cast<size_t>(idx)].count = count;
    return idx;
  }
  int mid = -1;
  if (sah) mid = sah_partition(out.order, prims, begin, end, cbox);
  if (mid < 0) {
    // median split on the largest centroid extent
    const float ex = cbox.hi[0] - cbox.lo[0], ey = cbox.hi[1] - cbox.lo[1], ez = cbox.hi[2] - cbox.lo[2];
    const int axis = (ex >= ey && ex >= ez) ? 0 : ((ey >= ex && ey >= ez) ? 1 : 2);
    mid = begin + count / 2;
    auto first = out.order.begin() + begin;
    std::nth_element(first, out.order.begin() + mid, out.order.begin() + end, [&](int a, int b) {
      return prims[static_cast<size_t>(a)].cen[axis] < prims[static_cast<size_t>(b)].cen[axis];
    });
  }
  const int l = build_binary(out, prims, begin, mid, depth + 1, leaf_size, sah);
  const int r = build_binary(out, prims, mid, end, depth + 1, leaf_size, sah);
  out.nodes[static_cast<size_t>(idx)].left = l;
  out.nodes[static_cast<size_t>(idx)].right = r;
  return idx;
}

uint32_t leaf_ref(const BinNode& b) {
  return kLeafFlag | (static_cast<uint32_t>(b.first) << 5) | static_cast<uint32_t>(b.count);
}

// Emits a 4-wide node for binary inner node `bi` (pre-order) and returns its index.
int collapse(HostBvh4& out, const Binary& bin, int bi, int depth) {
  out.depth = std::max(out.depth, depth);
  std::vector<int> kids;
  const BinNode& root = bin.nodes[static_cast<size_t>(bi)];
  if (root.count > 0) {
    kids.push_back(bi);  // a leaf root gets a node with one leaf child
  } else {
    kids.push_back(root.left);
    kids.push_back(root.right);
    while (kids.size() < 4) {
      int best = -1;
      float best_area = -1.0f;
      for (size_t k = 0; k < kids.size(); ++k) {
        const BinNode& n = bin.nodes[static_cast<size_t>(kids[k])];
        if (n.count == 0 && n.box.area() > best_area) {
          best_area = n.box.area();
          best = static_cast<int>(k);
        }
      }
      if (best < 0) break;
      const BinNode& n = bin.nodes[static_cast<size_t>(kids[static_cast<size_t>(best)])];
      kids[static_cast<size_t>(best)] = n.left;
      kids.insert(kids.begin() + best + 1, n.right);
    }
  }
  const int idx = static_cast<int>(out.nodes.size());
  out.nodes.push_back(Bvh4Node{});
  for (int k = 0; k < 4; ++k) {
    Bvh4Node& nd = out.nodes[static_cast<size_t>(idx)];
    if (k < static_cast<int>(kids.size())) {
      const BinNode& c = bin.nodes[static_cast<size_t>(kids[static_cast<size_t>(k)])];
      nd.lox[k] = c.box.lo[0]; nd.loy[k] = c.box.lo[1]; nd.loz[k] = c.box.lo[2];
      nd.hix[k] = c.box.hi[0]; nd.hiy[k] = c.box.hi[1]; nd.hiz[k] = c.box.hi[2];
      nd.child[k] = c.count > 0 ? leaf_ref(c) : 0u;  // inner refs patched below
    } else {
      const float inf = std::numeric_limits<float>::infinity();
      nd.lox[k] = nd.loy[k] = nd.loz[k] = inf;
      nd.hix[k] = nd.hiy[k] = nd.hiz[k] = -inf;
      nd.child[k] = kLeafFlag;  // empty
    }
    nd.pad[k] = 0;
  }
  for (int k = 0; k < static_cast<int>(kids.size()); ++k) {
    const BinNode& c = bin.nodes[static_cast<size_t>(kids[static_cast<size_t>(k)])];
    if (c.count == 0) {
      const int ci = collapse(out, bin, kids[static_cast<size_t>(k)], depth + 1);
      out.nodes[static_cast<size_t>(idx)].child[k] = static_cast<uint32_t>(ci);
    }
  }
  return idx;
}

}  // namespace

HostBvh4 build_bvh4(const float* verts, int nv, const uint32_t* tris, int nt, int leaf_size, bool sah) {
  if (nv <= 0 || nt <= 0) throw std::invalid_argument("mesh has no vertices or no triangles");
  if (leaf_size < 1 || leaf_size > kMaxLeafSize) throw std::invalid_argument("leaf size must be in [1, 31]");
  if (nt >= (1 << 26)) throw std::invalid_argument("at most 2^26 - 1 triangles are supported");
  std::vector<Prim> prims(static_cast<size_t>(nt));
  float scale = 1.0f;
  for (int i = 0; i < nt; ++i) {
    float c[3] = {0.0f, 0.0f, 0.0f};
    for (int k = 0; k < 3; ++k) {
      const uint32_t v = tris[3 * i + k];
      if (v >= static_cast<uint32_t>(nv)) throw std::invalid_argument("triangle references an out-of-range vertex");
      const float* p = verts + 3 * static_cast<size_t>(v);
      prims[static_cast<size_t>(i)].box.grow(p);
      for (int a = 0; a < 3; ++a) {
        c[a] = c[a] + p[a];
        scale = std::max(scale, std::fabs(p[a]));
      }
    }
    for (int a = 0; a < 3; ++a) prims[static_cast<size_t>(i)].cen[a] = c[a] / 3.0f;
  }
  for (int i = 0; i < 3 * nv; ++i) {
    if (!std::isfinite(verts[i])) throw std::invalid_argument("vertex coordinates must be finite");
  }
  Binary bin;
  bin.order.resize(static_cast<size_t>(nt));
  std::iota(bin.order.begin(), bin.order.end(), 0);
  bin.nodes.reserve(static_cast<size_t>(2 * (nt / std::max(1, leaf_size)) + 3));
  build_binary(bin, prims, 0, nt, 0, leaf_size, sah);

  HostBvh4 out;
  out.leaf_size = leaf_size;
  out.binary_depth = bin.depth;
  out.binary_nodes = static_cast<int>(bin.nodes.size());
  out.delta = 3.814697265625e-06f * scale;  // 2^-18 * scale
  out.nodes.reserve(bin.nodes.size() / 2 + 2);
  collapse(out, bin, 0, 0);

  out.tris.resize(static_cast<size_t>(nt));
  for (int i = 0; i < nt; ++i) {
    const int id = bin.order[static_cast<size_t>(i)];
    const float* a = verts + 3 * static_cast<size_t>(tris[3 * id]);
    const float* b = verts + 3 * static_cast<size_t>(tris[3 * id + 1]);
    const float* c = verts + 3 * static_cast<size_t>(tris[3 * id + 2]);
    TriRec& r = out.tris[static_cast<size_t>(i)];
    r.ax = a[0]; r.ay = a[1]; r.az = a[2];
    r.bx = b[0]; r.by = b[1]; r.bz = b[2];
    r.cx = c[0]; r.cy = c[1]; r.cz = c[2];
    r.id = id;
    r.pad0 = r.pad1 = 0;
  }
  return out;
}

}  // namespace n2
