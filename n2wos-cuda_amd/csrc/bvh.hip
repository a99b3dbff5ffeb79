// bvh.hip -- BVH handle and the closest-point query kernels behind
// n2wos_bvh_query_device[_masked|_counted] and n2wos_bvh_query_host.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <vector>

#include "api_internal.hpp"
#include "traverse.hpp"
#include "wave.hpp"
#include "wos.hpp"

namespace n2 {

// LDS ring capacity (entries per lane) of the stand-alone query kernels.
constexpr int kQueryCap = 16;

template <bool MASKED, bool COUNT>
__global__ void __launch_bounds__(1024) k_closest_point(const Bvh4Node* __restrict__ nodes,
                                                        const TriRec* __restrict__ tris, float delta,
                                                        const float* __restrict__ points,
                                                        const uint8_t* __restrict__ active, int n,
                                                        float* __restrict__ out_d2, float* __restrict__ out_cp,
                                                        int* __restrict__ out_tri, int* __restrict__ out_overflow,
                                                        uint32_t* __restrict__ out_nodes,
                                                        uint32_t* __restrict__ out_tris) {
  extern __shared__ uint2 lds[];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const V3 p = v3(points[3 * q], points[3 * q + 1], points[3 * q + 2]);
  if (MASKED && !active[q]) {
    // inactive slot: neutral outputs (src/cuda/cubql_bvh.cu:198-204)
    out_d2[q] = 0.0f;
    out_cp[3 * q] = p.x; out_cp[3 * q + 1] = p.y; out_cp[3 * q + 2] = p.z;
    out_tri[q] = -1;
    out_overflow[q] = 0;
    if (COUNT) { out_nodes[q] = 0; out_tris[q] = 0; }
    return;
  }
  Hit h;
  if (finite3(p)) {
    h = bvh4_closest<kQueryCap, COUNT>(nodes, tris, p, delta, FLT_MAX, lds + threadIdx.x, blockDim.x);
  } else {
    h.d2 = FLT_MAX; h.cp = v3(0.0f, 0.0f, 0.0f); h.id = -1; h.overflow = 0; h.nodes = 0; h.tris = 0;
  }
  out_d2[q] = h.d2;
  out_cp[3 * q] = h.cp.x; out_cp[3 * q + 1] = h.cp.y; out_cp[3 * q + 2] = h.cp.z;
  out_tri[q] = h.id;
  out_overflow[q] = h.overflow;
  if (COUNT) { out_nodes[q] = h.nodes; out_tris[q] = h.tris; }
}

// Four lanes per query (see bvh4_closest_quad): block of kQuadBlock lanes =
// kQuadBlock / 4 queries, each with an LDS ring of kQuadCap entries and one
// spare slot laid out [slot][query], so the quads of a wave hit consecutive
// banks.
constexpr int kQuadBlock = 256;
constexpr int kQuadQueries = kQuadBlock / 4;
constexpr int kQuadCap = 32;  // 16.5 KB per block with the spare slot

// Marker a capped query leaves in its overflow slot for the second pass.
constexpr int kQueryPending = 2;

template <bool MASKED, bool COUNT>
__global__ void __launch_bounds__(kQuadBlock) k_closest_point_quad(
    const Bvh4Node* __restrict__ nodes, const TriRec* __restrict__ tris, float delta,
    const float* __restrict__ points, const uint8_t* __restrict__ active, int n, float* __restrict__ out_d2,
    float* __restrict__ out_cp, int* __restrict__ out_tri, int* __restrict__ out_overflow,
    uint32_t* __restrict__ out_nodes, uint32_t* __restrict__ out_tris, int max_iters) {
  __shared__ uint2 s_ring[(kQuadCap + 1) * kQuadQueries];
  const int t = static_cast<int>(threadIdx.x);
  const int k = t & 3;
  const int q = static_cast<int>((static_cast<int64_t>(blockIdx.x) * kQuadBlock + t) >> 2);
  if (q >= n) return;  // all four lanes of a quad share q
  const V3 p = v3(points[3 * q], points[3 * q + 1], points[3 * q + 2]);
  Hit h;
  if (MASKED && !active[q]) {
    // inactive slot: neutral outputs (src/cuda/cubql_bvh.cu:198-204)
    h.d2 = 0.0f; h.cp = p; h.id = -1; h.overflow = 0; h.nodes = 0; h.tris = 0; h.pending = 0;
  } else if (finite3(p)) {
    h = bvh4_closest_quad<kQuadCap, COUNT>(nodes, tris, p, delta, s_ring + (t >> 2), kQuadQueries, k, max_iters);
  } else {
    h.d2 = FLT_MAX; h.cp = v3(0.0f, 0.0f, 0.0f); h.id = -1; h.overflow = 0; h.nodes = 0; h.tris = 0; h.pending = 0;
  }
  if (k == 0) {
    out_d2[q] = h.d2;
    out_cp[3 * q] = h.cp.x; out_cp[3 * q + 1] = h.cp.y; out_cp[3 * q + 2] = h.cp.z;
    out_tri[q] = h.id;
    out_overflow[q] = h.pending ? kQueryPending : h.overflow;
    if (COUNT) { out_nodes[q] = h.nodes; out_tris[q] = h.tris; }
  }
}

// Second pass: every query the first pass stopped at its iteration cap
// (overflow slot = kQueryPending) is finished by a whole wave.  The wave
// restarts the search from the root with the query's best so far: each round
// it pops up to 16 entries of a shared LDS stack, one per quad; in one memory
// phase a quad at an inner node loads child k's box in lane k and a quad at a
// leaf its first eight records.  Inner quads push the children within the
// threshold (ballot + rank, no order), leaf quads test their records four at a
// time (each behind its own box bound), and when some lane holds a (d2, id)
// below the best, the round's minimum over the wave updates the best and the
// threshold.  Every entry within the threshold is expanded and every record of
// an expanded leaf that can still win is tested, so the result is the exact
// minimum with the lowest index among ties, as in the one-quad search.  Stack
// keys keep the high 16 bits of their float: a lower bound of the key, so the
// pruning at a pop stays exact.  The stack holds `cap` entries (dynamic LDS,
// 6 bytes each), sized from the tree's depth by finish_stack_entries; a push
// beyond it sets the query's overflow flag.  A block is one wave over `slots`
// (1..64) consecutive query slots, whose marked queries it finishes one after
// another; slots with no marked query cost one load.
//
// Stack bound: popping up to 16 entries a round and pushing every child
// within the threshold, the stack of a full traversal with no pruning peaks
// at 240-276 entries on 65k-triangle trees of depth 8-11 and 352-404 on
// 1M-triangle trees of depth 13-15 (median and SAH builds, leaves of 3-8),
// under 28 per level; 64 per level, at least 512 and at most 2048 entries,
// leaves more than twice that.
__host__ __device__ constexpr int finish_stack_entries(int depth) {
  return 64 * (depth + 2) < 512 ? 512 : (64 * (depth + 2) > 2048 ? 2048 : 64 * (depth + 2));
}

__global__ void __launch_bounds__(64) k_closest_point_finish(const Bvh4Node* __restrict__ nodes,
                                                             const TriRec* __restrict__ tris, float delta,
                                                             const float* __restrict__ points, int n,
                                                             float* __restrict__ out_d2, float* __restrict__ out_cp,
                                                             int* __restrict__ out_tri,
                                                             int* __restrict__ out_overflow, int slots, int cap) {
  extern __shared__ uint32_t s_finish[];
  uint32_t* const s_ref = s_finish;
  uint16_t* const s_key = reinterpret_cast<uint16_t*>(s_finish + cap);
  const int lane = static_cast<int>(threadIdx.x);
  const int j = lane >> 2, k = lane & 3;
  const int q0 = static_cast<int>(blockIdx.x) * slots;
  const bool marked = lane < slots && q0 + lane < n && out_overflow[q0 + lane] == kQueryPending;
  uint64_t todo = __ballot(marked);
  while (todo != 0) {
    const int qq = q0 + __ffsll(static_cast<unsigned long long>(todo)) - 1;
    todo &= todo - 1;
    const V3 p = v3(points[3 * qq], points[3 * qq + 1], points[3 * qq + 2]);
    float bd2 = out_d2[qq];
    int bid = out_tri[qq];
    V3 bcp = v3(out_cp[3 * qq], out_cp[3 * qq + 1], out_cp[3 * qq + 2]);
    float thr = prune_threshold(bd2, delta);
    int ovf = 0;
    if (lane == 0) {
      s_ref[0] = 0u;  // the root
      s_key[0] = 0;
    }
    int sp = 1;  // wave-uniform
    wave_sync();
    while (sp > 0) {
      const int take = sp < 16 ? sp : 16;
      const int base = sp - take;
      sp = base;
      const bool has = j < take;
      const uint32_t ref = has ? s_ref[base + j] : 0u;
      const float key = has ? __uint_as_float(static_cast<uint32_t>(s_key[base + j]) << 16) : INFINITY;
      wave_sync();  // the popped slots may be overwritten by this round's pushes
      const bool live = has && key <= thr;
      const bool inner = live && !(ref & kLeafFlag);
      const int first = static_cast<int>((ref >> 5) & 0x3ffffffu);
      const int count = live && (ref & kLeafFlag) ? static_cast<int>(ref & 31u) : 0;
      // the round's one memory phase
      float lx = 0.0f, ly = 0.0f, lz = 0.0f, hx = 0.0f, hy = 0.0f, hz = 0.0f;
      uint32_t cr = 0u;
      if (inner) {
        const Bvh4Node* nd = nodes + ref;
        lx = nd->lox[k]; ly = nd->loy[k]; lz = nd->loz[k];
        hx = nd->hix[k]; hy = nd->hiy[k]; hz = nd->hiz[k];
        cr = nd->child[k];
      }
      float4 t0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), t1 = t0, t2 = t0, u0 = t0, u1 = t0, u2 = t0;
      if (k < count) {
        const float4* tp = reinterpret_cast<const float4*>(tris + first + k);
        t0 = tp[0]; t1 = tp[1]; t2 = tp[2];
      }
      if (4 + k < count) {
        const float4* tp = reinterpret_cast<const float4*>(tris + first + 4 + k);
        u0 = tp[0]; u1 = tp[1]; u2 = tp[2];
      }
      const float ck = inner ? box_dist2(p, lx, ly, lz, hx, hy, hz) : INFINITY;
      const bool push = inner && ck <= thr;
      const uint64_t pm = __ballot(push);
      const int npush = __popcll(pm);
      if (sp + npush > cap) {
        ovf = 1;  // the stack is full: the query reports overflow
      } else {
        if (push) {
          const int slot = sp + rank_in(pm);
          s_ref[slot] = cr;
          s_key[slot] = static_cast<uint16_t>(__float_as_uint(ck) >> 16);
        }
        sp += npush;
      }
      // leaves: records first + r0 + k, four per quad at a time
      float td2 = INFINITY;
      int tid = 0x7fffffff;
      int tlane_rec = -1;
      for (int r0 = 0;; r0 += 4) {
        const bool test = r0 + k < count;
        if (__ballot(test) == 0) break;
        if (test) {
          if (r0 == 4) {
            t0 = u0; t1 = u1; t2 = u2;
          } else if (r0 > 4) {
            const float4* tp = reinterpret_cast<const float4*>(tris + first + r0 + k);
            t0 = tp[0]; t1 = tp[1]; t2 = tp[2];
          }
          // the triangle's own box bounds its distance from below
          const float bxl = fminf(fminf(t0.x, t0.w), t1.z), bxh = fmaxf(fmaxf(t0.x, t0.w), t1.z);
          const float byl = fminf(fminf(t0.y, t1.x), t1.w), byh = fmaxf(fmaxf(t0.y, t1.x), t1.w);
          const float bzl = fminf(fminf(t0.z, t1.y), t2.x), bzh = fmaxf(fmaxf(t0.z, t1.y), t2.x);
          if (box_dist2(p, bxl, byl, bzl, bxh, byh, bzh) <= thr) {
            const V3 a = v3(t0.x, t0.y, t0.z), b = v3(t0.w, t1.x, t1.y), c = v3(t1.z, t1.w, t2.x);
            const V3 cq = closest_on_triangle_flat(p, a, b, c);
            const V3 e = cq - p;
            const float d2 = dot(e, e);
            const int id = __float_as_int(t2.y);
            if (d2 < td2 || (d2 == td2 && id < tid)) {
              td2 = d2;
              tid = id;
              tlane_rec = first + r0 + k;
            }
          }
        }
      }
      if (__ballot(td2 < bd2 || (td2 == bd2 && tid < bid)) != 0) {
        // the round's (d2, id) minimum over the wave: d2 >= 0, so its bits
        // order like its values, and the id breaks ties
        uint64_t key64 = (static_cast<uint64_t>(__float_as_uint(td2)) << 32) | static_cast<uint32_t>(tid);
        for (int off = 32; off > 0; off >>= 1) {
          const uint64_t o =
              (static_cast<uint64_t>(lane_read(static_cast<uint32_t>(key64 >> 32), lane ^ off)) << 32) |
              lane_read(static_cast<uint32_t>(key64), lane ^ off);
          key64 = o < key64 ? o : key64;
        }
        const float rd2 = __uint_as_float(static_cast<uint32_t>(key64 >> 32));
        const int rid = static_cast<int>(static_cast<uint32_t>(key64));
        // the winner's lane recomputes its closest point (the same operations)
        const uint64_t wm = __ballot(tlane_rec >= 0 && td2 == rd2 && tid == rid);
        const int wl = __ffsll(static_cast<unsigned long long>(wm)) - 1;
        const int rec = static_cast<int>(lane_read(static_cast<uint32_t>(tlane_rec), wl));
        const TriRec& tr = tris[rec];
        const V3 a = v3(tr.ax, tr.ay, tr.az), b = v3(tr.bx, tr.by, tr.bz), c = v3(tr.cx, tr.cy, tr.cz);
        bcp = closest_on_triangle_flat(p, a, b, c);
        bd2 = rd2;
        bid = rid;
        thr = fminf(thr, prune_threshold(bd2, delta));
      }
      wave_sync();
    }
    if (lane == 0) {
      out_d2[qq] = bd2;
      out_cp[3 * qq] = bcp.x;
      out_cp[3 * qq + 1] = bcp.y;
      out_cp[3 * qq + 2] = bcp.z;
      out_tri[qq] = bid;
      out_overflow[qq] = ovf;
    }
  }
}

// Ray-parity inside test (Moller-Trumbore hits with t > 0 along three fixed
// generic directions; majority vote), one lane per point.
__device__ int ray_crossings(const Bvh4Node* __restrict__ nodes, const TriRec* __restrict__ tris, V3 o, V3 dir) {
  const V3 inv = v3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  uint32_t stack[64];
  int sp = 0, hits = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const uint32_t ref = stack[--sp];
    if (!(ref & kLeafFlag)) {
      const Bvh4Node& nd = nodes[ref];
      for (int k = 0; k < 4; ++k) {
        const float tx0 = (nd.lox[k] - o.x) * inv.x, tx1 = (nd.hix[k] - o.x) * inv.x;
        const float ty0 = (nd.loy[k] - o.y) * inv.y, ty1 = (nd.hiy[k] - o.y) * inv.y;
        const float tz0 = (nd.loz[k] - o.z) * inv.z, tz1 = (nd.hiz[k] - o.z) * inv.z;
        const float tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
        const float tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
        const uint32_t c = nd.child[k];
        if (tmin <= tmax && !((c & kLeafFlag) && (c & 31u) == 0u) && sp < 64) stack[sp++] = c;
      }
    } else {
      const int first = static_cast<int>((ref >> 5) & 0x3ffffffu), count = static_cast<int>(ref & 31u);
      for (int i = 0; i < count; ++i) {
        const TriRec& t = tris[first + i];
        const V3 a = v3(t.ax, t.ay, t.az);
        const V3 e1 = v3(t.bx, t.by, t.bz) - a, e2 = v3(t.cx, t.cy, t.cz) - a;
        const V3 h = v3(dir.y * e2.z - dir.z * e2.y, dir.z * e2.x - dir.x * e2.z, dir.x * e2.y - dir.y * e2.x);
        const float det = dot(e1, h);
        if (fabsf(det) < 1e-20f) continue;
        const float inv_det = 1.0f / det;
        const V3 s = o - a;
        const float u = dot(s, h) * inv_det;
        if (u < 0.0f || u > 1.0f) continue;
        const V3 q = v3(s.y * e1.z - s.z * e1.y, s.z * e1.x - s.x * e1.z, s.x * e1.y - s.y * e1.x);
        const float v = dot(dir, q) * inv_det;
        if (v < 0.0f || u + v > 1.0f) continue;
        if (dot(e2, q) * inv_det > 0.0f) ++hits;
      }
    }
  }
  return hits;
}

__global__ void k_inside(const Bvh4Node* __restrict__ nodes, const TriRec* __restrict__ tris,
                         const float* __restrict__ points, int n, uint8_t* __restrict__ inside) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const V3 p = v3(points[3 * q], points[3 * q + 1], points[3 * q + 2]);
  if (!finite3(p)) {
    inside[q] = 0;
    return;
  }
  const V3 dirs[3] = {v3(0.9234801f, 0.2901198f, 0.2510278f), v3(-0.3376619f, 0.8152146f, 0.4705419f),
                      v3(0.1666621f, -0.2151597f, 0.9622526f)};
  int votes = 0;
  for (int k = 0; k < 3; ++k) votes += ray_crossings(nodes, tris, p, dirs[k]) & 1;
  inside[q] = votes >= 2 ? 1 : 0;
}

namespace {

template <bool MASKED, bool COUNT>
int launch_query(const n2wos_bvh* bvh, const float* d_points, const uint8_t* d_active, int n, float* d_d2,
                 float* d_cp, int* d_tri, int* d_overflow, uint32_t* d_nodes, uint32_t* d_tris, int block,
                 void* stream) {
  N2_REQUIRE(bvh != nullptr, N2WOS_ERR_NULL, "bvh handle is NULL");
  N2_REQUIRE(bvh->d_nodes && bvh->d_tris, N2WOS_ERR_EMPTY, "query_device called on an empty BVH");
  N2_REQUIRE(n >= 0, N2WOS_ERR_INVALID, "query_count must be non-negative");
  if (n == 0) return N2WOS_OK;
  N2_REQUIRE(d_points && d_d2 && d_cp && d_tri && d_overflow, N2WOS_ERR_NULL,
             "query_device received a null device pointer");
  N2_REQUIRE(!MASKED || d_active, N2WOS_ERR_NULL, "query_device_masked received a null mask");
  N2_REQUIRE(!COUNT || (d_nodes && d_tris), N2WOS_ERR_NULL, "counted query received a null counter pointer");
  N2_REQUIRE(block >= 1 && block <= 1024, N2WOS_ERR_BLOCK, "block size must be in [1, 1024]");
  if (tuning_value("query_quad", 1)) {
    // four lanes per query with its own launch shape; block_size, validated
    // above as the reference does, changes no result
    const int64_t lanes = 4 * static_cast<int64_t>(n);
    const int grid = static_cast<int>((lanes + kQuadBlock - 1) / kQuadBlock);
    // queries past the cap are finished by the wave-wide second pass (the
    // counting build runs every query to its end in one pass)
    const int cap = COUNT ? 0 : tuning_value("query_cap", 16);
    const int fslots = tuning_value("query_finish_slots", 8);
    hipLaunchKernelGGL((k_closest_point_quad<MASKED, COUNT>), dim3(grid), dim3(kQuadBlock), 0,
                       static_cast<hipStream_t>(stream), bvh->d_nodes, bvh->d_tris, bvh->delta, d_points, d_active, n,
                       d_d2, d_cp, d_tri, d_overflow, d_nodes, d_tris, cap);
    if (cap > 0) {
      const int scap = finish_stack_entries(bvh->depth);
      hipLaunchKernelGGL(k_closest_point_finish, dim3((n + fslots - 1) / fslots), dim3(64),
                         static_cast<size_t>(scap) * (sizeof(uint32_t) + sizeof(uint16_t)),
                         static_cast<hipStream_t>(stream), bvh->d_nodes, bvh->d_tris, bvh->delta, d_points, n, d_d2,
                         d_cp, d_tri, d_overflow, fslots, scap);
    }
    N2_HIP(hipGetLastError());
    return N2WOS_OK;
  }
  const size_t lds = static_cast<size_t>(kQueryCap) * static_cast<size_t>(block) * sizeof(uint2);
  auto kern = k_closest_point<MASKED, COUNT>;
  if (lds > 65536) {
    N2_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(lds)));
  }
  const int grid = (n + block - 1) / block;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, static_cast<hipStream_t>(stream), bvh->d_nodes,
                     bvh->d_tris, bvh->delta, d_points, d_active, n, d_d2, d_cp, d_tri, d_overflow, d_nodes,
                     d_tris);
  N2_HIP(hipGetLastError());
  return N2WOS_OK;
}

}  // namespace

void launch_closest_point(const n2wos_bvh* bvh, const float* d_points, int n, float* d_d2, float* d_cp, int* d_tri,
                          int* d_ovf, hipStream_t stream) {
  launch_query<false, false>(bvh, d_points, nullptr, n, d_d2, d_cp, d_tri, d_ovf, nullptr, nullptr, 128, stream);
}

}  // namespace n2

using namespace n2;

namespace {
// Build-method names: the median split ("median"; the reference's own host
// builder is "custom") and the binned SAH ("sah", the reference's cuBQL
// naming, src/cuda/cubql_bvh.cu:114-122 for the aliases).
std::string canonical_method(const char* build_method) {
  std::string m = build_method ? build_method : "median";
  if (m == "custom" || m == "custom_cuda_bvh") m = "median";
  if (m == "surface_area_heuristic" || m == "cubql_cuda_sah") m = "sah";
  N2_REQUIRE(m == "median" || m == "sah", N2WOS_ERR_UNSUPPORTED, "unsupported BVH build method: " + m);
  return m;
}
}  // namespace

extern "C" int n2wos_bvh_create(const float* verts, int nv, const uint32_t* tris, int nt, int leaf_size,
                                const char* build_method, n2wos_bvh** out) {
  return guarded([&]() -> int {
    N2_REQUIRE(out != nullptr, N2WOS_ERR_NULL, "out handle pointer is NULL");
    *out = nullptr;
    N2_REQUIRE(verts && tris, N2WOS_ERR_NULL, "mesh arrays must not be NULL");
    N2_REQUIRE(nv > 0 && nt > 0, N2WOS_ERR_EMPTY, "mesh has no vertices or no triangles");
    const std::string method = canonical_method(build_method);
    const auto t0 = std::chrono::steady_clock::now();
    HostBvh4 host = build_bvh4(verts, nv, tris, nt, leaf_size, method == "sah");
    const auto t1 = std::chrono::steady_clock::now();
    auto* h = new n2wos_bvh();
    try {
      N2_HIP(hipGetDevice(&h->device));
      N2_HIP(hipMalloc(&h->d_nodes, host.nodes.size() * sizeof(Bvh4Node)));
      N2_HIP(hipMalloc(&h->d_tris, host.tris.size() * sizeof(TriRec)));
      N2_HIP(hipMemcpy(h->d_nodes, host.nodes.data(), host.nodes.size() * sizeof(Bvh4Node), hipMemcpyHostToDevice));
      N2_HIP(hipMemcpy(h->d_tris, host.tris.data(), host.tris.size() * sizeof(TriRec), hipMemcpyHostToDevice));
    } catch (...) {
      if (h->d_nodes) (void)hipFree(h->d_nodes);
      if (h->d_tris) (void)hipFree(h->d_tris);
      delete h;
      throw;
    }
    h->triangle_count = nt;
    h->node_count = static_cast<int64_t>(host.nodes.size());
    h->binary_nodes = host.binary_nodes;
    h->leaf_size = host.leaf_size;
    h->max_depth = host.binary_depth;
    h->depth = host.depth;
    h->delta = host.delta;
    h->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    h->build_method = method;
    const Bvh4Node& r = host.nodes[0];
    for (int a = 0; a < 3; ++a) {
      const float* lo = a == 0 ? r.lox : (a == 1 ? r.loy : r.loz);
      const float* hi = a == 0 ? r.hix : (a == 1 ? r.hiy : r.hiz);
      float l = lo[0], u = hi[0];
      for (int k = 1; k < 4; ++k) {
        l = std::min(l, lo[k]);
        u = std::max(u, hi[k]);
      }
      h->bounds_lo[a] = l;
      h->bounds_hi[a] = u;
    }
    *out = h;
    return N2WOS_OK;
  });
}

extern "C" int n2wos_bvh_destroy(n2wos_bvh* bvh) {
  return guarded([&]() -> int {
    if (!bvh) return N2WOS_OK;
    if (bvh->d_nodes) (void)hipFree(bvh->d_nodes);
    if (bvh->d_tris) (void)hipFree(bvh->d_tris);
    delete bvh;
    return N2WOS_OK;
  });
}

extern "C" int n2wos_bvh_get_stats(const n2wos_bvh* bvh, n2wos_bvh_stats* out) {
  return guarded([&]() -> int {
    N2_REQUIRE(bvh && out, N2WOS_ERR_NULL, "NULL argument");
    std::memset(out, 0, sizeof(*out));
    out->triangle_count = bvh->triangle_count;
    out->node_count = bvh->node_count;
    out->index_count = bvh->triangle_count;
    out->binary_nodes = bvh->binary_nodes;
    out->leaf_size = bvh->leaf_size;
    out->max_depth = bvh->max_depth;
    out->depth4 = bvh->depth;
    out->prune_margin = bvh->delta;
    out->build_milliseconds = bvh->build_ms;
    out->device_bytes = bvh->node_count * static_cast<int64_t>(sizeof(Bvh4Node)) +
                        bvh->triangle_count * static_cast<int64_t>(sizeof(TriRec));
    for (int a = 0; a < 3; ++a) {
      out->bounds_lo[a] = bvh->bounds_lo[a];
      out->bounds_hi[a] = bvh->bounds_hi[a];
    }
    std::strncpy(out->build_method, bvh->build_method.c_str(), sizeof(out->build_method) - 1);
    return N2WOS_OK;
  });
}

extern "C" int n2wos_bvh_build_host(const float* verts, int nv, const uint32_t* tris, int nt, int leaf_size,
                                    const char* build_method, void* nodes_out, int64_t* node_count, void* tris_out,
                                    float* prune_margin) {
  return guarded([&]() -> int {
    N2_REQUIRE(verts && tris && node_count, N2WOS_ERR_NULL, "NULL argument");
    N2_REQUIRE(nv > 0 && nt > 0, N2WOS_ERR_EMPTY, "mesh has no vertices or no triangles");
    HostBvh4 host = build_bvh4(verts, nv, tris, nt, leaf_size, canonical_method(build_method) == "sah");
    *node_count = static_cast<int64_t>(host.nodes.size());
    if (prune_margin) *prune_margin = host.delta;
    if (nodes_out) std::memcpy(nodes_out, host.nodes.data(), host.nodes.size() * sizeof(Bvh4Node));
    if (tris_out) std::memcpy(tris_out, host.tris.data(), host.tris.size() * sizeof(TriRec));
    return N2WOS_OK;
  });
}

extern "C" int n2wos_bvh_export(const n2wos_bvh* bvh, void* nodes_out, void* tris_out) {
  return guarded([&]() -> int {
    N2_REQUIRE(bvh != nullptr, N2WOS_ERR_NULL, "bvh handle is NULL");
    if (nodes_out)
      N2_HIP(hipMemcpy(nodes_out, bvh->d_nodes, bvh->node_count * sizeof(Bvh4Node), hipMemcpyDeviceToHost));
    if (tris_out)
      N2_HIP(hipMemcpy(tris_out, bvh->d_tris, bvh->triangle_count * sizeof(TriRec), hipMemcpyDeviceToHost));
    return N2WOS_OK;
  });
}

extern "C" int n2wos_bvh_query_device(const n2wos_bvh* bvh, const float* d_points, int query_count,
                                      float* d_distance2, float* d_closest, int* d_triangle_id, int* d_overflow,
                                      int block_size, void* stream) {
  return guarded([&]() -> int {
    return launch_query<false, false>(bvh, d_points, nullptr, query_count, d_distance2, d_closest, d_triangle_id,
                                      d_overflow, nullptr, nullptr, block_size, stream);
  });
}

extern "C" int n2wos_bvh_query_host(const n2wos_bvh* bvh, const float* points, int query_count, float* distance2,
                                    float* closest, int* triangle_id, int* overflow_count,
                                    float* kernel_milliseconds, int block_size) {
  return guarded([&]() -> int {
    N2_REQUIRE(bvh != nullptr, N2WOS_ERR_NULL, "bvh handle is NULL");
    N2_REQUIRE(query_count >= 0, N2WOS_ERR_INVALID, "query_count must be non-negative");
    if (overflow_count) *overflow_count = 0;
    if (kernel_milliseconds) *kernel_milliseconds = 0.0f;
    if (query_count == 0) return N2WOS_OK;
    N2_REQUIRE(points && distance2 && closest && triangle_id, N2WOS_ERR_NULL, "NULL host pointer");
    const size_t n = static_cast<size_t>(query_count);
    // one allocation: points, closest (3 floats each), d2, tri, overflow
    void* buf = nullptr;
    N2_HIP(hipMalloc(&buf, n * 36));
    struct Free {
      void* p;
      ~Free() { (void)hipFree(p); }
    } guard{buf};
    float* d_pts = static_cast<float*>(buf);
    float* d_cp = d_pts + 3 * n;
    float* d_d2 = d_cp + 3 * n;
    int* d_tri = reinterpret_cast<int*>(d_d2 + n);
    int* d_ovf = d_tri + n;
    N2_HIP(hipMemcpy(d_pts, points, n * 12, hipMemcpyHostToDevice));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    N2_HIP(hipEventCreate(&e0));
    N2_HIP(hipEventCreate(&e1));
    struct Ev {
      hipEvent_t a, b;
      ~Ev() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
      }
    } evs{e0, e1};
    N2_HIP(hipEventRecord(e0, nullptr));
    const int rc = launch_query<false, false>(bvh, d_pts, nullptr, query_count, d_d2, d_cp, d_tri, d_ovf, nullptr,
                                              nullptr, block_size, nullptr);
    if (rc != N2WOS_OK) return rc;
    N2_HIP(hipEventRecord(e1, nullptr));
    N2_HIP(hipEventSynchronize(e1));
    if (kernel_milliseconds) N2_HIP(hipEventElapsedTime(kernel_milliseconds, e0, e1));
    N2_HIP(hipMemcpy(distance2, d_d2, n * 4, hipMemcpyDeviceToHost));
    N2_HIP(hipMemcpy(closest, d_cp, n * 12, hipMemcpyDeviceToHost));
    N2_HIP(hipMemcpy(triangle_id, d_tri, n * 4, hipMemcpyDeviceToHost));
    if (overflow_count) {
      std::vector<int> ovf(n);
      N2_HIP(hipMemcpy(ovf.data(), d_ovf, n * 4, hipMemcpyDeviceToHost));
      int c = 0;
      for (int v : ovf) c += v != 0;
      *overflow_count = c;
    }
    return N2WOS_OK;
  });
}

extern "C" int n2wos_bvh_query_device_masked(const n2wos_bvh* bvh, const float* d_points, const uint8_t* d_active,
                                             int query_count, float* d_distance2, float* d_closest,
                                             int* d_triangle_id, int* d_overflow, int block_size, void* stream) {
  return guarded([&]() -> int {
    return launch_query<true, false>(bvh, d_points, d_active, query_count, d_distance2, d_closest, d_triangle_id,
                                     d_overflow, nullptr, nullptr, block_size, stream);
  });
}

extern "C" int n2wos_bvh_inside_device(const n2wos_bvh* bvh, const float* d_points, int n, uint8_t* d_inside,
                                       void* stream) {
  return guarded([&]() -> int {
    N2_REQUIRE(bvh != nullptr, N2WOS_ERR_NULL, "bvh handle is NULL");
    N2_REQUIRE(bvh->d_nodes && bvh->d_tris, N2WOS_ERR_EMPTY, "inside test on an empty BVH");
    N2_REQUIRE(n >= 0, N2WOS_ERR_INVALID, "n must be non-negative");
    if (n == 0) return N2WOS_OK;
    N2_REQUIRE(d_points && d_inside, N2WOS_ERR_NULL, "NULL device pointer");
    const int block = 128;
    hipLaunchKernelGGL(k_inside, dim3((n + block - 1) / block), dim3(block), 0, static_cast<hipStream_t>(stream),
                       bvh->d_nodes, bvh->d_tris, d_points, n, d_inside);
    N2_HIP(hipGetLastError());
    return N2WOS_OK;
  });
}

extern "C" int n2wos_bvh_query_device_counted(const n2wos_bvh* bvh, const float* d_points, int query_count,
                                              float* d_distance2, float* d_closest, int* d_triangle_id,
                                              int* d_overflow, uint32_t* d_node_visits, uint32_t* d_tri_visits,
                                              int block_size, void* stream) {
  return guarded([&]() -> int {
    return launch_query<false, true>(bvh, d_points, nullptr, query_count, d_distance2, d_closest, d_triangle_id,
                                     d_overflow, d_node_visits, d_tri_visits, block_size, stream);
  });
}
