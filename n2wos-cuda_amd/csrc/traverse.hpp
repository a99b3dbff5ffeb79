// traverse.hpp -- exact closest-point traversal of the 4-wide BVH on gfx950.
//
// One lane = one query.  The traversal stack holds (node ref, squared box
// distance) pairs, one 8-byte entry each, so a push or pop is one LDS or
// scratch access.  Its top CAP entries live in an LDS ring laid out
// [slot][lane], so the lanes of a wave always hit consecutive banks; when a
// push meets a full ring, the ring's oldest entry moves to a per-lane spill
// array (scratch memory, touched only by deep traversals).  Pops take the
// ring first, then the spill array, so the visit order is exactly that of one
// stack of CAP + kSpill entries.  A node visit leaves at most 3 entries
// behind, so a tree of up to (CAP + kSpill) / 3 levels (18 for a ring of 8,
// 21 for a ring of 16) never overflows; the benchmark
// meshes build 10 to 17 levels.  A deeper tree whose traversal fills the spill
// array sets the overflow flag, like the reference's bounded stack.  Children
// are visited nearest first; a popped entry whose stored distance already
// exceeds the pruning threshold is discarded without touching its node.  Ties
// in squared distance go to the lowest original triangle index, which
// reproduces the reference's brute-force scan
// (src/geometry/closest_point.cpp:50-68) exactly.
#pragma once

#include "bvh.hpp"
#include "geom.hpp"

#include <climits>

namespace n2 {

struct Hit {
  float d2;
  V3 cp;
  int id;
  int overflow;
  uint32_t nodes;  // node lines loaded (counting variant only)
  uint32_t tris;   // triangle records tested (counting variant only)
  int pending;     // stopped at the iteration cap with work left (d2, cp, id: the best so far)
};

__device__ __forceinline__ void cswap(float& ka, uint32_t& ra, int& sa, float& kb, uint32_t& rb, int& sb) {
  const bool sw = (kb < ka) || (kb == ka && sb < sa);
  if (sw) {
    const float tk = ka; ka = kb; kb = tk;
    const uint32_t tr = ra; ra = rb; rb = tr;
    const int ts = sa; sa = sb; sb = ts;
  }
}

constexpr int kSpill = 48;

// sent points at this lane's column of the LDS ring (entry = node ref, bits
// of the squared distance); consecutive slots are `stride` entries apart.  thr_init bounds the search from above (a
// squared distance known to be >= the answer); FLT_MAX for an unbounded query.
template <int CAP, bool COUNT>
__device__ __forceinline__ Hit bvh4_closest(const Bvh4Node* __restrict__ nodes, const TriRec* __restrict__ tris,
                                            V3 p, float delta, float thr_init, uint2* sent, int stride) {
  static_assert((CAP & (CAP - 1)) == 0, "ring capacity must be a power of two");
  Hit h;
  h.d2 = FLT_MAX;
  h.cp = v3(0.0f, 0.0f, 0.0f);
  h.id = -1;
  h.overflow = 0;
  h.nodes = 0;
  h.tris = 0;
  h.pending = 0;
  uint64_t spill_e[kSpill];  // ref | distance bits << 32
  float thr = thr_init;
  uint32_t cur = 0;
  bool have = true;
  int top = 0, cnt = 0, scnt = 0;
  while (true) {
    if (have) {
      if (!(cur & kLeafFlag)) {
        const float4* np = reinterpret_cast<const float4*>(nodes + cur);
        const float4 lox = np[0], loy = np[1], loz = np[2];
        const float4 hix = np[3], hiy = np[4], hiz = np[5];
        const uint4 ch = reinterpret_cast<const uint4*>(np)[6];
        if (COUNT) ++h.nodes;
        float k0 = box_dist2(p, lox.x, loy.x, loz.x, hix.x, hiy.x, hiz.x);
        float k1 = box_dist2(p, lox.y, loy.y, loz.y, hix.y, hiy.y, hiz.y);
        float k2 = box_dist2(p, lox.z, loy.z, loz.z, hix.z, hiy.z, hiz.z);
        float k3 = box_dist2(p, lox.w, loy.w, loz.w, hix.w, hiy.w, hiz.w);
        const int nvalid = (k0 <= thr) + (k1 <= thr) + (k2 <= thr) + (k3 <= thr);
        if (nvalid == 0) {
          have = false;
        } else {
          uint32_t r0 = ch.x, r1 = ch.y, r2 = ch.z, r3 = ch.w;
          int s0 = 0, s1 = 1, s2 = 2, s3 = 3;
          k0 = k0 <= thr ? k0 : INFINITY;
          k1 = k1 <= thr ? k1 : INFINITY;
          k2 = k2 <= thr ? k2 : INFINITY;
          k3 = k3 <= thr ? k3 : INFINITY;
          cswap(k0, r0, s0, k1, r1, s1);
          cswap(k2, r2, s2, k3, r3, s3);
          cswap(k0, r0, s0, k2, r2, s2);
          cswap(k1, r1, s1, k3, r3, s3);
          cswap(k1, r1, s1, k2, r2, s2);
          // push farthest first, keep descending into the nearest
#define N2_PUSH(R, K)                                  \
  do {                                                 \
    if (cnt == CAP) {                                  \
      if (scnt < kSpill) {                             \
        const uint2 ve_ = sent[top * stride];          \
        spill_e[scnt] = (static_cast<uint64_t>(ve_.y) << 32) | ve_.x; \
        ++scnt;                                        \
      } else {                                         \
        h.overflow = 1;                                \
      }                                                \
    } else {                                           \
      ++cnt;                                           \
    }                                                  \
    sent[top * stride] = make_uint2((R), __float_as_uint(K)); \
    top = (top + 1) & (CAP - 1);                       \
  } while (0)
          if (nvalid >= 4) N2_PUSH(r3, k3);
          if (nvalid >= 3) N2_PUSH(r2, k2);
          if (nvalid >= 2) N2_PUSH(r1, k1);
#undef N2_PUSH
          cur = r0;
          continue;
        }
      } else {
        const int first = static_cast<int>((cur >> 5) & 0x3ffffffu);
        const int count = static_cast<int>(cur & 31u);
        // the next record's loads are issued before this record's test, so
        // the lane waits for memory once per leaf rather than per record
        const float4* lp = reinterpret_cast<const float4*>(tris + first);
        float4 n0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), n1 = n0, n2 = n0;
        if (count > 0) {
          n0 = lp[0];
          n1 = lp[1];
          n2 = lp[2];
        }
        for (int i = 0; i < count; ++i) {
          const float4 t0 = n0, t1 = n1, t2 = n2;
          if (i + 1 < count) {
            n0 = lp[3 * (i + 1)];
            n1 = lp[3 * (i + 1) + 1];
            n2 = lp[3 * (i + 1) + 2];
          }
          if (COUNT) ++h.tris;
          // the triangle's own box bounds its distance from below
          const float bxl = fminf(fminf(t0.x, t0.w), t1.z), bxh = fmaxf(fmaxf(t0.x, t0.w), t1.z);
          const float byl = fminf(fminf(t0.y, t1.x), t1.w), byh = fmaxf(fmaxf(t0.y, t1.x), t1.w);
          const float bzl = fminf(fminf(t0.z, t1.y), t2.x), bzh = fmaxf(fmaxf(t0.z, t1.y), t2.x);
          if (box_dist2(p, bxl, byl, bzl, bxh, byh, bzh) > thr) continue;
          const V3 a = v3(t0.x, t0.y, t0.z), b = v3(t0.w, t1.x, t1.y), c = v3(t1.z, t1.w, t2.x);
          const int id = __float_as_int(t2.y);
          const V3 q = closest_on_triangle_flat(p, a, b, c);
          const V3 e = q - p;
          const float d2 = dot(e, e);
          if (d2 < h.d2 || (d2 == h.d2 && id < h.id)) {
            h.d2 = d2;
            h.cp = q;
            h.id = id;
            thr = fminf(thr, prune_threshold(d2, delta));
          }
        }
        have = false;
      }
    }
    // pop the next entry that can still hold a closer triangle
    while (cnt > 0 || scnt > 0) {
      float k;
      uint32_t r;
      if (cnt > 0) {
        top = (top - 1) & (CAP - 1);
        --cnt;
        const uint2 e = sent[top * stride];
        k = __uint_as_float(e.y);
        r = e.x;
      } else {
        --scnt;
        const uint64_t e = spill_e[scnt];
        k = __uint_as_float(static_cast<uint32_t>(e >> 32));
        r = static_cast<uint32_t>(e);
      }
      if (k <= thr) {
        cur = r;
        have = true;
        break;
      }
    }
    if (!have) break;
  }
  return h;
}

// ---------------------------------------------------------------- quads
// Four lanes (one quad: lanes 4j .. 4j+3, which DPP quad permutations reach
// in one instruction) share one query.  Lane k of the quad loads and tests
// child k of each node and triangle k, k+4, ... of each leaf; every lane
// holds the same copy of the query's state (threshold, best, stack
// counters), so the quad's control flow is uniform and the visit order, the
// nodes and triangle records visited and the result are exactly those of
// bvh4_closest.  The LDS ring is one per quad (`sent` is the quad's column
// of CAP + 1 slots, the last a spare that takes the writes of absent
// entries; written by lane 0, read by all four); the scratch spill array is each
// lane's own copy.  A wave serves 16 queries, so a far query keeps 4 lanes
// busy instead of holding up 63 idle ones.
#ifndef N2_EMULATED_WAVE
template <int J>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), J * 0x55, 0xF, 0xF, false));
}
template <int J>
__device__ __forceinline__ float quad_bcast(float v) {
  return __uint_as_float(quad_bcast<J>(__float_as_uint(v)));
}
// value of the quad lane whose index is this lane's XOR X (X = 1 or 2)
template <int X>
__device__ __forceinline__ uint32_t quad_xor(uint32_t v) {
  constexpr int ctrl = X == 1 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] / [2,3,0,1]
  return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), ctrl, 0xF, 0xF, false));
}

// (d2, id) lexicographic minimum over the quad, carrying the closest point
template <int X>
__device__ __forceinline__ void quad_min_step(float& d2, int& id, V3& q) {
  const float od2 = __uint_as_float(quad_xor<X>(__float_as_uint(d2)));
  const int oid = static_cast<int>(quad_xor<X>(static_cast<uint32_t>(id)));
  const float ox = __uint_as_float(quad_xor<X>(__float_as_uint(q.x)));
  const float oy = __uint_as_float(quad_xor<X>(__float_as_uint(q.y)));
  const float oz = __uint_as_float(quad_xor<X>(__float_as_uint(q.z)));
  const bool take = od2 < d2 || (od2 == d2 && oid < id);
  d2 = take ? od2 : d2;
  id = take ? oid : id;
  q = v3(take ? ox : q.x, take ? oy : q.y, take ? oz : q.z);
}

// max_iters > 0 stops the query at its first pop after at least that many
// node visits and leaves (at most one descent past it) with work left
// (h.pending = 1, d2 / cp / id the best so far); <= 0: no limit.  The cap is
// tested only where the loop already decides to end, so uncapped queries run
// the same loop.
template <int CAP, bool COUNT>
__device__ __forceinline__ Hit bvh4_closest_quad(const Bvh4Node* __restrict__ nodes,
                                                 const TriRec* __restrict__ tris, V3 p, float delta, uint2* sent,
                                                 int stride, int k, int max_iters) {
  static_assert((CAP & (CAP - 1)) == 0, "ring capacity must be a power of two");
  Hit h;
  h.d2 = FLT_MAX;
  h.cp = v3(0.0f, 0.0f, 0.0f);
  h.id = -1;
  h.overflow = 0;
  h.nodes = 0;
  h.tris = 0;
  h.pending = 0;
  uint64_t spill_e[kSpill];  // ref | distance bits << 32
  float thr = FLT_MAX;
  uint32_t cur = 0;
  bool have = true;
  int top = 0, cnt = 0, scnt = 0;
  int iters = 0;
  const int lim = max_iters > 0 ? max_iters : INT_MAX;
  while (true) {
    if (have) {
      ++iters;
      if (!(cur & kLeafFlag)) {
        const Bvh4Node* nd = nodes + cur;
        if (COUNT) ++h.nodes;
        const float kk = box_dist2(p, nd->lox[k], nd->loy[k], nd->loz[k], nd->hix[k], nd->hiy[k], nd->hiz[k]);
        const uint32_t rk = nd->child[k];
        float k0 = quad_bcast<0>(kk), k1 = quad_bcast<1>(kk), k2 = quad_bcast<2>(kk), k3 = quad_bcast<3>(kk);
        uint32_t r0 = quad_bcast<0>(rk), r1 = quad_bcast<1>(rk), r2 = quad_bcast<2>(rk), r3 = quad_bcast<3>(rk);
        const int nvalid = (k0 <= thr) + (k1 <= thr) + (k2 <= thr) + (k3 <= thr);
        if (nvalid == 0) {
          have = false;
        } else {
          int s0 = 0, s1 = 1, s2 = 2, s3 = 3;
          k0 = k0 <= thr ? k0 : INFINITY;
          k1 = k1 <= thr ? k1 : INFINITY;
          k2 = k2 <= thr ? k2 : INFINITY;
          k3 = k3 <= thr ? k3 : INFINITY;
          cswap(k0, r0, s0, k1, r1, s1);
          cswap(k2, r2, s2, k3, r3, s3);
          cswap(k0, r0, s0, k2, r2, s2);
          cswap(k1, r1, s1, k3, r3, s3);
          cswap(k1, r1, s1, k2, r2, s2);
          // push r3, r2, r1 (the last nvalid - 1 of them, farthest first) in
          // one step: the oldest ring entries move to the spill stack first
          // when the ring overflows; absent entries are written to the spare
          // slot CAP, so lane 0 writes without branches
          const int m = nvalid - 1;
          if (cnt + m > CAP) {
            for (int j = CAP - cnt; j < m; ++j) {
              if (scnt < kSpill) {
                const uint2 ve = sent[((top + j) & (CAP - 1)) * stride];
                spill_e[scnt] = (static_cast<uint64_t>(ve.y) << 32) | ve.x;
                ++scnt;
              } else {
                h.overflow = 1;
              }
            }
          }
          if (k == 0) {
            const int q1 = m >= 1 ? ((top + m - 1) & (CAP - 1)) : CAP;
            const int q2 = m >= 2 ? ((top + m - 2) & (CAP - 1)) : CAP;
            const int q3 = m >= 3 ? top : CAP;
            sent[q3 * stride] = make_uint2(r3, __float_as_uint(k3));
            sent[q2 * stride] = make_uint2(r2, __float_as_uint(k2));
            sent[q1 * stride] = make_uint2(r1, __float_as_uint(k1));
          }
          top = (top + m) & (CAP - 1);
          cnt = cnt + m < CAP ? cnt + m : CAP;
          cur = r0;
          continue;
        }
      } else {
        const int first = static_cast<int>((cur >> 5) & 0x3ffffffu);
        const int count = static_cast<int>(cur & 31u);
        if (COUNT) h.tris += static_cast<uint32_t>(count);
        for (int i0 = 0; i0 < count; i0 += 4) {
          float d2 = INFINITY;
          int id = 0x7fffffff;
          V3 q = v3(0.0f, 0.0f, 0.0f);
          if (i0 + k < count) {
            const float4* tp = reinterpret_cast<const float4*>(tris + first + i0 + k);
            const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
            // the triangle's own box bounds its distance from below
            const float bxl = fminf(fminf(t0.x, t0.w), t1.z), bxh = fmaxf(fmaxf(t0.x, t0.w), t1.z);
            const float byl = fminf(fminf(t0.y, t1.x), t1.w), byh = fmaxf(fmaxf(t0.y, t1.x), t1.w);
            const float bzl = fminf(fminf(t0.z, t1.y), t2.x), bzh = fmaxf(fmaxf(t0.z, t1.y), t2.x);
            if (box_dist2(p, bxl, byl, bzl, bxh, byh, bzh) <= thr) {
              const V3 a = v3(t0.x, t0.y, t0.z), b = v3(t0.w, t1.x, t1.y), c = v3(t1.z, t1.w, t2.x);
              q = closest_on_triangle_flat(p, a, b, c);
              const V3 e = q - p;
              d2 = dot(e, e);
              id = __float_as_int(t2.y);
            }
          }
          quad_min_step<1>(d2, id, q);
          quad_min_step<2>(d2, id, q);
          if (d2 < h.d2 || (d2 == h.d2 && id < h.id)) {
            h.d2 = d2;
            h.cp = q;
            h.id = id;
            thr = fminf(thr, prune_threshold(d2, delta));
          }
        }
        have = false;
      }
    }
    // pop the next entry that can still hold a closer triangle
    while (cnt > 0 || scnt > 0) {
      float kp;
      uint32_t r;
      if (cnt > 0) {
        top = (top - 1) & (CAP - 1);
        --cnt;
        const uint2 e = sent[top * stride];
        kp = __uint_as_float(e.y);
        r = e.x;
      } else {
        --scnt;
        const uint64_t e = spill_e[scnt];
        kp = __uint_as_float(static_cast<uint32_t>(e >> 32));
        r = static_cast<uint32_t>(e);
      }
      if (kp <= thr) {
        cur = r;
        have = true;
        break;
      }
    }
    const bool stop = have && iters >= lim;
    if (stop) h.pending = 1;
    if (!have || stop) break;
  }
  return h;
}
#endif

}  // namespace n2
