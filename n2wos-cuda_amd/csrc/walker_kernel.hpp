// walker_kernel.hpp -- device side of the persistent Walk-on-Spheres walker
// (included by wos.hip; also compiled on the host by the SIMT emulator used in
// the CPU tests, which supplies its own definitions of the wave intrinsics).
#pragma once

#include "cvfield.hpp"
#include "traverse.hpp"
#include "wave.hpp"
#include "wos.hpp"

namespace n2 {

#ifndef N2_EMU_LOOP_HOOK
#define N2_EMU_LOOP_HOOK()
#define N2_EMU_TRACE(sN, sT, sU)
#define N2_EMU_TRI_HOOK(idx)
#endif

// Compare-exchange by key for the walker's child sort, branch-free.  Ties
// keep no particular order: the visit order of equal-distance children
// changes neither the closest point nor the walk.
__device__ __forceinline__ void cswap_key(float& ka, uint32_t& ra, float& kb, uint32_t& rb) {
  const bool sw = kb < ka;
  const float lo = sw ? kb : ka, hi = sw ? ka : kb;
  const uint32_t rlo = sw ? rb : ra, rhi = sw ? ra : rb;
  ka = lo;
  kb = hi;
  ra = rlo;
  rb = rhi;
}

// Squared distance from p to a box, with the sum of squares fused (two FMAs):
// a pruning key only, so a rounding different from box_dist2's by an ulp
// stays far inside the threshold's margin (see below).
__device__ __forceinline__ float box_dist2_fused(V3 p, float lox, float loy, float loz, float hix, float hiy,
                                                 float hiz) {
  const float dx = fmaxf(fmaxf(lox - p.x, p.x - hix), 0.0f);
  const float dy = fmaxf(fmaxf(loy - p.y, p.y - hiy), 0.0f);
  const float dz = fmaxf(fmaxf(loz - p.z, p.z - hiz), 0.0f);
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// The walker's pruning threshold: prune_threshold with the hardware square
// root and the relative margin raised from 2^-20 to 2^-19.  A square root at
// most 2 ulp low still gives sqrt(best)(1 + 2^-19)(1 - 2^-22) > sqrt(best)(1 +
// 2^-20), so this threshold is never below prune_threshold(best): the walker
// visits a superset of the nodes the exact search needs, and its closest
// points stay those of the brute-force scan.
__device__ __forceinline__ float walker_prune_threshold(float best, float delta) {
  const float t = hw_sqrt(best) * 1.0000019073486328125f + delta;
  return best >= FLT_MAX ? FLT_MAX : t * t;
}


// Chan et al. pairwise merge of (n, mean, M2); same expression order as the oracle.
__host__ __device__ __forceinline__ void chan_merge(double na, double ma, double sa, double nb, double mb, double sb,
                                                    double& n, double& m, double& s) {
  if (nb == 0.0) {
    n = na; m = ma; s = sa;
    return;
  }
  if (na == 0.0) {
    n = nb; m = mb; s = sb;
    return;
  }
  const double nn = na + nb;
  const double d = mb - ma;
  n = nn;
  m = ma + d * (nb / nn);
  s = sa + sb + d * d * (na * nb / nn);
}

#ifndef N2_WAVES_PER_BLOCK
#define N2_WAVES_PER_BLOCK 1
#endif
constexpr int kWaves = N2_WAVES_PER_BLOCK;  // waves per block
constexpr int kBlock = 64 * kWaves;
// Waves per SIMD the register allocation must allow: 5 (at most 96 VGPRs,
// no spills).  A one-wave block holds 7.2 KB of LDS with a ring of 16
// entries per lane (6 bytes each), so 20 blocks fit a CU's 160 KB.
#ifndef N2_WALKER_WAVES_PER_SIMD
#define N2_WALKER_WAVES_PER_SIMD 5
#endif


// Walk modes: plain walks; prefix walks that stop at depth m and leave the
// cache query point X_m (neural-cache WoS); coupled walks that record the
// depth-m state and then continue to the boundary (two-level MC residual).
constexpr int kModePlain = 0;
constexpr int kModePrefix = 1;
constexpr int kModeCoupled = 2;

// Lane phases of the persistent walker.
constexpr int kIdle = 0;    // no walk: waiting for refill
constexpr int kNode = 1;    // query in progress, at an inner node
constexpr int kTri = 2;     // query in progress, testing the triangles of a leaf
constexpr int kUpdate = 3;  // query finished: take the walk step (or end the walk)

// Records of its leaf a lane offers to one cooperative triangle iteration
// (a leaf of more is finished in later iterations).
constexpr int kTriTake = 4;

// Ring overflow goes first to the lane's own line of kLineSpill entries in the
// workspace (one 128-byte line per lane, so a lane's spills and the loads
// that take them back touch one line, which stays in L2 between them), then
// to the per-lane scratch array of kSpill entries.  Per-lane scratch is
// interleaved across the wave's lanes dword by dword: an 8-byte entry there
// dirties two lines, each holding 4 bytes of it.
constexpr int kLineSpill = 16;

// Walk flags beside the unit slot and walk-in-unit index (bits 0-8).
constexpr int kFlagBounded = 1 << 9;  // the query started from the step bound, not unbounded
constexpr int kFlagOvf = 1 << 10;     // the query's traversal stack overflowed

// Every iteration the wave serves ONE phase, chosen by ballot (the phase with
// the most lanes per instruction of its path), so divergent lanes never pay
// for paths they do not take, and a lane whose query ends starts its next
// query without waiting for the rest of the wave.  Each lane's own sequence of
// node visits, triangle tests and walk steps is exactly the sequential one.
//
// EXT selects the extended walker (plain walks only): the along-walk control
// variate over a field (cvfield.hpp) and the optional path record.  The plain
// instantiation carries none of it.
template <int CAP, int S, bool POISSON, bool COUNT, bool EXT = false>
__global__ void __launch_bounds__(kBlock, N2_WALKER_WAVES_PER_SIMD) k_wos_persistent(WosKernelArgs a) {
  static_assert(S <= 8, "the unit slot is kept in 3 bits");
  // CAP ring slots per lane plus one spare slot that takes the writes of
  // absent entries, so a node visit pushes without branches.  An entry is a
  // node ref and the high 16 bits of its squared box distance: the key
  // truncated toward zero, a lower bound, so an entry is popped whenever its
  // exact key would be (a few more, never fewer), and the search stays exact.
  __shared__ uint32_t s_ref[(CAP + 1) * kBlock];
  __shared__ uint16_t s_key[(CAP + 1) * kBlock];
  __shared__ int s_unit[kWaves][S];
  __shared__ int s_rem[kWaves][S];
  __shared__ uint32_t s_steps[kWaves][S];
  __shared__ uint32_t s_forced[kWaves][S];
  __shared__ uint32_t s_ovf[kWaves][S];
  __shared__ float s_charge[4 * kMaxCharges];
  __shared__ uint32_t s_task[kWaves][64];  // triangle tasks: record << 6 | owner lane

  // arguments used only on rare paths (refill, walk step, walk end) are read
  // where they are used instead of being held in scalar registers across the
  // loop, which the node and triangle paths need
#ifndef N2_EMULATED_WAVE
#define KA (::n2::kernel_args<WosKernelArgs>())  // `a` is this kernel's only argument
#else
#define KA (&a)
#endif
  const int lane = lane_id();
  const int wv = static_cast<int>(threadIdx.x) >> 6;
  uint32_t* sref = s_ref + threadIdx.x;
  uint16_t* skey = s_key + threadIdx.x;
  if (lane < S) s_unit[wv][lane] = -1;
  for (int i = static_cast<int>(threadIdx.x); i < 4 * kMaxCharges; i += kBlock) s_charge[i] = a.bc.charge[i >> 2][i & 3];
  __syncthreads();

  // wave-uniform fill cursor
  int fill_slot = -1, fill_next = 0, fill_count = 0, fill_point = 0, fill_w0 = 0;
  bool no_more = false;
  int ranges_done = 0;  // wave-uniform: unit ranges found exhausted

  // walk state
  int phase = kIdle;
  // the walk's index in its unit (bits 0-5), its unit slot (6-8) and the
  // flags kFlagBounded and kFlagOvf, in one register: all are used only on
  // the walk-step and walk-end paths
  int wflags = 0;
  int point = 0, walk = 0, step = 0;
  V3 p = v3(0.0f, 0.0f, 0.0f);
  float src = 0.0f;
  float cvs = 0.0f;  // extended walker: sum of the control-variate terms
  // query state
  // one position register: the node index in the node phase, the next
  // triangle record in the triangle phase (a lane is in one at a time)
  int cur = 0;
  int te = 0, top = 0, cnt = 0, scnt = 0, bid = -1;
  float thr = FLT_MAX, bd2 = FLT_MAX;
  V3 bcp = v3(0.0f, 0.0f, 0.0f);
  uint64_t spill_e[kSpill];
  uint32_t n_nodes = 0, n_tris = 0;

  // this lane's spill line, computed where it is used (spills are rare): the
  // opaque lane index keeps its address out of the registers the loop holds
#define N2_SPILL_LINE()                                                                         \
  (KA->spill_lines + (static_cast<size_t>(blockIdx.x) * kBlock + static_cast<size_t>(wv) * 64 + \
                      static_cast<uint32_t>(opaque_lane_int(lane_id()))) * kLineSpill)
  // go to the entry `r` of the traversal
#define N2_ENTER(r)                                       \
  do {                                                    \
    const uint32_t enter_ref_ = (r);                      \
    if (enter_ref_ & kLeafFlag) {                         \
      cur = static_cast<int>((enter_ref_ >> 5) & 0x3ffffffu); \
      te = cur + static_cast<int>(enter_ref_ & 31u);      \
      phase = kTri;                                       \
    } else {                                              \
      cur = static_cast<int>(enter_ref_);                 \
      phase = kNode;                                      \
    }                                                     \
  } while (0)
  // pop the next entry that can still hold a closer triangle, or finish
#define N2_POP_NEXT()                                \
  do {                                               \
    bool pop_found_ = false;                         \
    uint32_t pop_ref_ = 0;                           \
    /* the loop carries only the stack state; the entry found is entered after it */ \
    while (!pop_found_ && (cnt > 0 || scnt > 0)) {   \
      float pop_key_;                                \
      if (cnt > 0) {                                 \
        top = (top - 1) & (CAP - 1);                 \
        --cnt;                                       \
        pop_key_ = __uint_as_float(static_cast<uint32_t>(skey[top * kBlock]) << 16); \
        pop_ref_ = sref[top * kBlock];               \
      } else {                                       \
        --scnt;                                      \
        const uint64_t se_ = scnt < kLineSpill ? N2_SPILL_LINE()[scnt] : spill_e[scnt - kLineSpill]; \
        pop_key_ = __uint_as_float(static_cast<uint32_t>(se_ >> 32)); \
        pop_ref_ = static_cast<uint32_t>(se_);       \
      }                                              \
      pop_found_ = pop_key_ <= thr;                  \
    }                                                \
    if (pop_found_) {                                \
      N2_ENTER(pop_ref_);                            \
    } else {                                         \
      phase = kUpdate;                               \
    }                                                \
  } while (0)
  // push the last M (0-3) of (R3, K3), (R2, K2), (R1, K1) in that order, as
  // M single pushes would: R1 ends on top.  When the ring overflows, its
  // oldest entries move to the spill stack first (oldest first), or are
  // dropped with the overflow flag set when that is full too.
#define N2_PUSH3(M, R1, K1, R2, K2, R3, K3)                                       \
  do {                                                                            \
    const int m_ = (M);                                                           \
    if (cnt + m_ > CAP) {                                                         \
      for (int j_ = CAP - cnt; j_ < m_; ++j_) {                                   \
        if (scnt < kLineSpill + kSpill) {                                         \
          const int vs_ = ((top + j_) & (CAP - 1)) * kBlock;                       \
          const uint64_t e_ = (static_cast<uint64_t>(skey[vs_]) << 48) | sref[vs_]; \
          if (scnt < kLineSpill)                                                  \
            N2_SPILL_LINE()[scnt] = e_;                                           \
          else                                                                    \
            spill_e[scnt - kLineSpill] = e_;                                      \
          ++scnt;                                                                 \
        } else {                                                                  \
          wflags |= kFlagOvf;                                                     \
        }                                                                         \
      }                                                                           \
    }                                                                             \
    const int s1_ = m_ >= 1 ? ((top + m_ - 1) & (CAP - 1)) : CAP;                 \
    const int s2_ = m_ >= 2 ? ((top + m_ - 2) & (CAP - 1)) : CAP;                 \
    const int s3_ = m_ >= 3 ? top : CAP;                                          \
    sref[s3_ * kBlock] = (R3);                                                    \
    skey[s3_ * kBlock] = static_cast<uint16_t>(__float_as_uint(K3) >> 16);        \
    sref[s2_ * kBlock] = (R2);                                                    \
    skey[s2_ * kBlock] = static_cast<uint16_t>(__float_as_uint(K2) >> 16);        \
    sref[s1_ * kBlock] = (R1);                                                    \
    skey[s1_ * kBlock] = static_cast<uint16_t>(__float_as_uint(K1) >> 16);        \
    top = (top + m_) & (CAP - 1);                                                 \
    cnt = cnt + m_ < CAP ? cnt + m_ : CAP;                                        \
  } while (0)

  // one node visit: the node's line is (lox, loy, loz, hix, hiy, hiz, child refs)
#define N2_NODE_VISIT(lox, loy, loz, hix, hiy, hiz, chf)                          \
  do {                                                                             \
    if (COUNT) ++n_nodes;                                                          \
    float k0 = box_dist2_fused(p, lox.x, loy.x, loz.x, hix.x, hiy.x, hiz.x);       \
    float k1 = box_dist2_fused(p, lox.y, loy.y, loz.y, hix.y, hiy.y, hiz.y);       \
    float k2 = box_dist2_fused(p, lox.z, loy.z, loz.z, hix.z, hiy.z, hiz.z);       \
    float k3 = box_dist2_fused(p, lox.w, loy.w, loz.w, hix.w, hiy.w, hiz.w);       \
    const int nvalid = (k0 <= thr) + (k1 <= thr) + (k2 <= thr) + (k3 <= thr);      \
    if (nvalid == 0) {                                                             \
      N2_POP_NEXT();                                                               \
    } else {                                                                       \
      uint32_t r0 = static_cast<uint32_t>(__float_as_int(chf.x));                  \
      uint32_t r1 = static_cast<uint32_t>(__float_as_int(chf.y));                  \
      uint32_t r2 = static_cast<uint32_t>(__float_as_int(chf.z));                  \
      uint32_t r3 = static_cast<uint32_t>(__float_as_int(chf.w));                  \
      k0 = k0 <= thr ? k0 : INFINITY;                                              \
      k1 = k1 <= thr ? k1 : INFINITY;                                              \
      k2 = k2 <= thr ? k2 : INFINITY;                                              \
      k3 = k3 <= thr ? k3 : INFINITY;                                              \
      cswap_key(k0, r0, k1, r1);                                                   \
      cswap_key(k2, r2, k3, r3);                                                   \
      cswap_key(k0, r0, k2, r2);                                                   \
      cswap_key(k1, r1, k3, r3);                                                   \
      cswap_key(k1, r1, k2, r2);                                                   \
      N2_PUSH3(nvalid - 1, r1, k1, r2, k2, r3, k3);                                \
      N2_ENTER(r0);                                                                \
    }                                                                              \
  } while (0)

  while (true) {
    N2_EMU_LOOP_HOOK();
    // ---------------------------------------------------------- refill
    uint64_t idle = __ballot(phase == kIdle);
    if (__popcll(idle) >= a.refill_min || __popcll(idle) == 64) {
      while (idle != 0) {
        if (fill_next == fill_count) {
          if (no_more) break;
          int fs = -1;
          for (int s = 0; s < S; ++s) {
            if (s_unit[wv][s] < 0) {
              fs = s;
              break;
            }
          }
          if (fs < 0) break;
          // claim from this XCD's range first, then from the others in turn
          int u = -1;
          const int nr = KA->unit_ranges;
          const int home = nr > 1 ? xcd_id() % nr : 0;
          for (int k = 0; k < nr && u < 0; ++k) {
            const int r = home + k < nr ? home + k : home + k - nr;
            if ((ranges_done >> r) & 1) continue;
            const int b0 = static_cast<int>(static_cast<int64_t>(r) * KA->total_units / nr);
            const int b1 = static_cast<int>(static_cast<int64_t>(r + 1) * KA->total_units / nr);
            int t = 0;
            if (lane == 0) t = atomicAdd(KA->next_unit + r * kUnitCounterStride, 1);
            t = __shfl(t, 0);
            if (t < b1 - b0) {
              u = b0 + t;
            } else {
              ranges_done |= 1 << r;
            }
          }
          if (u < 0) {
            no_more = true;
            break;
          }
          // pt: the unit's first position in the claim order (KA->order
          // maps positions to points; identity when null)
          int pt, chunk, ucnt;
          if (KA->pack_shift >= 0) {
            // packed unit: 64 / wpp whole points
            pt = u << (6 - KA->pack_shift);
            chunk = 0;
            ucnt = min(64, (KA->n_points - pt) << KA->pack_shift);
          } else {
            pt = u / KA->units_per_point;
            chunk = u - pt * KA->units_per_point;
            ucnt = min(64, KA->walks_per_point - 64 * chunk);
          }
          if (lane == 0) {
            s_unit[wv][fs] = u;
            s_rem[wv][fs] = ucnt;
            s_steps[wv][fs] = 0u;
            s_forced[wv][fs] = 0u;
            s_ovf[wv][fs] = 0u;
          }
          wave_sync();
          fill_slot = fs;
          fill_next = 0;
          fill_count = ucnt;
          fill_point = pt;
          fill_w0 = 64 * chunk;
        }
        const int take = min(__popcll(idle), fill_count - fill_next);
        const bool is_idle = (idle >> lane) & 1ull;
        const int r = rank_in(idle);
        if (is_idle && r < take) {
          // every walk of a point shares its first query: computed once per point
          const int wiu = fill_next + r;
          if (KA->pack_shift >= 0) {
            point = fill_point + (wiu >> KA->pack_shift);
            walk = wiu & ((1 << KA->pack_shift) - 1);
          } else {
            point = fill_point;
            walk = fill_w0 + wiu;
          }
          if (KA->order) point = KA->order[point];
          step = 0;
          src = 0.0f;
          if (EXT) cvs = 0.0f;
          p = v3(KA->points[3 * point], KA->points[3 * point + 1], KA->points[3 * point + 2]);
          bd2 = KA->first_d2[point];
          bcp = v3(KA->first_cp[3 * point], KA->first_cp[3 * point + 1], KA->first_cp[3 * point + 2]);
          bid = 0;
          // the first query is unbounded (kFlagBounded clear)
          wflags = (fill_slot << 6) | wiu | (KA->first_ovf[point] ? kFlagOvf : 0);
          phase = kUpdate;
        }
        fill_next += take;
        idle = __ballot(phase == kIdle);
      }
    }

    const uint64_t mN = __ballot(phase == kNode);
    const uint64_t mT = __ballot(phase == kTri);
    const uint64_t mU = __ballot(phase == kUpdate);
    if ((mN | mT | mU) == 0) {
      if (no_more || __popcll(idle) < 64) break;
      continue;
    }
    const int sN = __popcll(mN) * a.weight_node;
    const int sT = __popcll(mT) * a.weight_tri;
    const int sU = __popcll(mU) * a.weight_update;
    N2_EMU_TRACE(sN, sT, sU);

    if (sU >= sN && sU >= sT) {
      // ------------------------------------------------------ walk step
      bool unit_done = false;
      if (phase == kUpdate) {
        if (bid < 0 && (wflags & kFlagBounded) && finite3(p)) {
          // the step bound was too tight (cannot happen in exact arithmetic): redo unbounded
          wflags &= ~kFlagBounded;
          thr = FLT_MAX;
          top = cnt = scnt = 0;
          cur = 0;
          phase = kNode;
        } else {
          const float d = sqrtf(fmaxf(bd2, 0.0f));
          const bool forced = step >= KA->max_steps;
          const bool at_boundary = d <= KA->epsilon || forced || !isfinite(d);
          // prefix mode ends a walk at depth m (after the boundary test, as the
          // reference's hybrid prefix does, src/cuda/cubql_bvh.cu:1486-1510)
          const bool stop_prefix = !at_boundary && KA->mode == kModePrefix && step >= KA->depth;
          if (EXT) {
            // path record: X_step and q_step, the terminal position included
            if (step < KA->path_len) {
              const int64_t pidx =
                  (static_cast<int64_t>(point) * KA->walks_per_point + walk) * KA->path_len + step;
              KA->path_pos[3 * pidx] = p.x;
              KA->path_pos[3 * pidx + 1] = p.y;
              KA->path_pos[3 * pidx + 2] = p.z;
              KA->path_q[pidx] = src - KA->cv_beta * cvs;
            }
          }
          if (at_boundary || stop_prefix) {
            float gval = 0.0f;
            if (at_boundary) {
              if (KA->bc.mode == kBcCharges) {
                // the charges are read from LDS at an offset the compiler
                // cannot see through, so their 64 values are not held in
                // registers across the loop
                gval = charges_value(s_charge + opaque_lane_int(0), KA->bc.n_charges, bcp);
              } else {
                Boundary b0;
                b0.mode = KA->bc.mode;
                b0.n_charges = 0;
                gval = boundary_value(b0, bcp);
              }
            }
            float value = at_boundary ? src + gval : src;
            if (EXT) value = value - KA->cv_beta * cvs;
            if (KA->mode != kModePlain) {
              const int64_t sidx = static_cast<int64_t>(point) * KA->walks_per_point + walk;
              if (KA->mode == kModeCoupled) KA->nc_full[sidx] = value;
              // coupled walks that passed depth m recorded their prefix there
              if (KA->mode == kModePrefix || step <= KA->depth) {
                KA->nc_prefix[sidx] = value;
                KA->nc_need[sidx] = stop_prefix ? 1 : 0;
                KA->nc_xm[3 * sidx] = p.x;
                KA->nc_xm[3 * sidx + 1] = p.y;
                KA->nc_xm[3 * sidx + 2] = p.z;
              }
            }
            // staged per unit slot in global memory, indexed by walk-in-unit
            const int slot = (wflags >> 6) & 7;
            const bool ovf = (wflags & kFlagOvf) != 0;
            KA->unit_vals[((static_cast<int64_t>(blockIdx.x) * kWaves + wv) * S + slot) * 64 + (wflags & 63)] =
                value;
            if (KA->pack_shift >= 0) {
              // packed unit: the point's own counters (integer sums, any order)
              atomicAdd(&KA->unit_steps[point], static_cast<uint32_t>(step));
              if (forced || ovf)
                atomicAdd(&KA->unit_flags[point], (forced ? 1u : 0u) + (ovf ? 0x10000u : 0u));
            } else {
              atomicAdd(&s_steps[wv][slot], static_cast<uint32_t>(step));
              if (forced) atomicAdd(&s_forced[wv][slot], 1u);
              if (ovf) atomicAdd(&s_ovf[wv][slot], 1u);
            }
            if (KA->walk_values) {
              const int64_t sidx = static_cast<int64_t>(point) * KA->walks_per_point + walk;
              KA->walk_values[sidx] = value;
              KA->walk_steps[sidx] = step;
            }
            const int old = atomicSub(&s_rem[wv][slot], 1);
            unit_done = (old == 1);
            phase = kIdle;
          } else {
            if (KA->mode == kModeCoupled && step == KA->depth) {
              // the coupled walk's NC state at depth m: prefix source and X_m
              const int64_t sidx = static_cast<int64_t>(point) * KA->walks_per_point + walk;
              KA->nc_prefix[sidx] = src;
              KA->nc_need[sidx] = 1;
              KA->nc_xm[3 * sidx] = p.x;
              KA->nc_xm[3 * sidx + 1] = p.y;
              KA->nc_xm[3 * sidx + 2] = p.z;
            }
            // antithetic pairs: the odd walk of a pair takes its first step
            // along the negated direction drawn for its even partner
            const bool anti = KA->antithetic && step == 0 && (walk & 1);
            // the walk's Philox counter words, derived here rather than held
            // in registers across the loop
            uint32_t ctr_point, ctr_walk;
            if (KA->point_rng) {
              ctr_point = KA->point_rng[2 * point];
              ctr_walk = KA->point_rng[2 * point + 1] + static_cast<uint32_t>(walk);
            } else {
              ctr_point = static_cast<uint32_t>(point) + KA->point_offset;
              ctr_walk = static_cast<uint32_t>(walk);
            }
            const uint32_t cw = anti ? ctr_walk - 1u : ctr_walk;
            const float sgn = anti ? -1.0f : 1.0f;
            // the key words are read here, so the compiler derives the ten
            // round keys here (scalar adds) instead of keeping 20 of them in
            // registers across the whole loop
            const uint32_t key_lo = KA->seed_lo, key_hi = KA->seed_hi;
            const U4 rr = philox4x32_10(U4{static_cast<uint32_t>(step), cw, ctr_point, 2u * KA->stream},
                                        key_lo, key_hi);
            V3 dir = sphere_dir(rr.x, rr.y);
            dir = v3(sgn * dir.x, sgn * dir.y, sgn * dir.z);
            const float R = KA->step_scale * d;
            // the control variate's frame angle: word 2 of the direction draw
            // (Laplace) or word 3 of the source draw (Poisson), both unused
            uint32_t psi_word = rr.z;
            if (POISSON) {
              const U4 t = philox4x32_10(U4{static_cast<uint32_t>(step), cw, ctr_point, 2u * KA->stream + 1u},
                                         key_lo, key_hi);
              psi_word = t.w;
              const float rho = median3(u01(rr.z), u01(rr.w), u01(t.x));
              V3 e = sphere_dir(t.y, t.z);
              e = v3(sgn * e.x, sgn * e.y, sgn * e.z);
              const float Rr = R * rho;
              const V3 y = v3(p.x + Rr * e.x, p.y + Rr * e.y, p.z + Rr * e.z);
              src = src + (R * R) * poisson_source(y) / 6.0f;
            }
            if (!EXT) (void)psi_word;
            if (EXT) {
              if (KA->cv_points > 0)
                cvs = cvs + cv_term(KA->cv_values, KA->cv_res, KA->cv_origin[0], KA->cv_origin[1],
                                    KA->cv_origin[2], KA->cv_inv_h[0], KA->cv_inv_h[1], KA->cv_inv_h[2],
                                    KA->cv_points, p, R, dir, psi_word);
            }
            p = v3(p.x + R * dir.x, p.y + R * dir.y, p.z + R * dir.z);
            ++step;
            // the previous closest point lies on the mesh, so the next
            // distance is at most |p_new - closest| (about 1.3 d on average,
            // against d + R): bound the next query by it.  The margin covers
            // the FP32 rounding of the closest point and of the distance.
            const V3 eo = p - bcp;
            // clamped to FLT_MAX: a bound that overflows is the unbounded
            // query, and an empty child slot (box key +inf) stays unvisited
            const float tb = hw_sqrt(dot(eo, eo)) * 1.0000152587890625f + 4.0f * a.delta;
            const float thr0 = fminf(tb * tb, FLT_MAX);
            bd2 = FLT_MAX;
            bcp = v3(0.0f, 0.0f, 0.0f);
            bid = -1;
            // a new query: the overflow flag cleared, bounded unless thr0 is FLT_MAX
            wflags = (wflags & 511) | (thr0 < FLT_MAX ? kFlagBounded : 0);
            if (finite3(p)) {
              thr = thr0;
              top = cnt = scnt = 0;
              cur = 0;
              phase = kNode;
            } else {
              bd2 = INFINITY;  // non-finite position: the walk ends at the next update
            }
          }
        }
      }
      wave_sync();
      // ---------------------------------------------------- finish units
      uint64_t done = __ballot(unit_done);
      // the unit values this wave stored are read back below
      if (done != 0) wave_global_sync();
      while (done != 0) {
        const int l = __ffsll(static_cast<unsigned long long>(done)) - 1;
        done &= done - 1;
        const int s = (__shfl(wflags, l) >> 6) & 7;
        const int u = s_unit[wv][s];
        const bool packed = KA->pack_shift >= 0;
        int pt, ucnt, span;
        if (packed) {
          pt = u << (6 - KA->pack_shift);
          ucnt = min(64, (KA->n_points - pt) << KA->pack_shift);
          span = 1 << KA->pack_shift;  // one point's walks: the tree stops at its segment
        } else {
          pt = u / KA->units_per_point;
          ucnt = min(64, KA->walks_per_point - 64 * (u - pt * KA->units_per_point));
          span = 64;
        }
        double n = lane < ucnt ? 1.0 : 0.0;
        const float* uv = KA->unit_vals + ((static_cast<int64_t>(blockIdx.x) * kWaves + wv) * S + s) * 64;
        double m = lane < ucnt ? static_cast<double>(load_l2(uv + lane)) : 0.0;
        double q = 0.0;
        // Chan tree over lane pairs at strides 1, 2, 4, ...: within a packed
        // unit's segments it gives each point exactly what the full tree
        // gives a one-point unit (merges with empty lanes change nothing)
#pragma unroll
        for (int stride = 1; stride < 64; stride *= 2) {
          const double nb = __shfl_down(n, stride);
          const double mb = __shfl_down(m, stride);
          const double qb = __shfl_down(q, stride);
          if (stride < span && (lane & (2 * stride - 1)) == 0) chan_merge(n, m, q, nb, mb, qb, n, m, q);
        }
        if (packed) {
          if ((lane & (span - 1)) == 0 && lane < ucnt) {
            int pp = pt + (lane >> KA->pack_shift);
            if (KA->order) pp = KA->order[pp];
            KA->unit_mean[pp] = m;
            KA->unit_m2[pp] = q;
          }
          if (lane == 0) s_unit[wv][s] = -1;
        } else if (lane == 0) {
          // the point's units are stored in walk order at point * units_per_point
          const int uo = KA->order ? KA->order[pt] * KA->units_per_point + (u - pt * KA->units_per_point) : u;
          KA->unit_mean[uo] = m;
          KA->unit_m2[uo] = q;
          KA->unit_steps[uo] = s_steps[wv][s];
          KA->unit_flags[uo] = (s_forced[wv][s] & 0xffffu) | (s_ovf[wv][s] << 16);
          s_unit[wv][s] = -1;
        }
        wave_sync();
      }
    } else if (sN >= sT) {
      // ------------------------------------------------------ one node visit
      if (phase == kNode) {
        const float4* np = reinterpret_cast<const float4*>(a.nodes + cur);
        const float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3], n4 = np[4], n5 = np[5], n6 = np[6];
        N2_NODE_VISIT(n0, n1, n2, n3, n4, n5, n6);
      }
    } else {
      // ------------------------------------- triangle tests, wave-cooperative
      // Each lane in the triangle phase offers the next (up to kTriTake)
      // records of its leaf as tasks; the first 64 tasks in lane order are
      // spread one per lane and tested at once, so a wave with 24 lanes in
      // the phase tests about 60 records per iteration instead of 24.  The
      // owner of a task then folds the results into its best, which becomes
      // what testing them one by one gives (every record of a leaf is
      // tested: no pre-test).
      const int offer = phase == kTri ? max(0, min(te - cur, kTriTake)) : 0;
      const uint64_t tb0 = __ballot(offer & 1), tb1 = __ballot(offer & 2), tb2 = __ballot(offer & 4);
      const int pre = rank_in(tb0) + 2 * rank_in(tb1) + 4 * rank_in(tb2);  // first task of this lane
      const int ntask_all = __popcll(tb0) + 2 * __popcll(tb1) + 4 * __popcll(tb2);
      const int ntask = ntask_all < 64 ? ntask_all : 64;
      const int take = max(0, min(offer, 64 - pre));  // this lane's tasks served now
      for (int j = 0; j < take; ++j) s_task[wv][pre + j] = (static_cast<uint32_t>(cur + j) << 6) | lane;
      wave_sync();
      const bool tasked = lane < ntask;
      const uint32_t task = tasked ? s_task[wv][lane] : static_cast<uint32_t>(lane);
      const int owner = static_cast<int>(task & 63u);
      const V3 tp = v3(lane_read_f(p.x, owner), lane_read_f(p.y, owner), lane_read_f(p.z, owner));
      float td2 = 0.0f;
      int tid = 0;
      V3 tq = v3(0.0f, 0.0f, 0.0f);
      if (tasked) {
        int rec = static_cast<int>(task >> 6);
        N2_EMU_TRI_HOOK(rec);
        const float4* trp = reinterpret_cast<const float4*>(a.tris + rec);
        const float4 t0 = trp[0], t1 = trp[1], t2 = trp[2];
        const V3 ta = v3(t0.x, t0.y, t0.z), tb = v3(t0.w, t1.x, t1.y), tc = v3(t1.z, t1.w, t2.x);
        tid = __float_as_int(t2.y);
        tq = closest_on_triangle_flat(tp, ta, tb, tc);
        const V3 e = tq - tp;
        td2 = dot(e, e);
      }
      // fold in record order: the lexicographic (d2, id) minimum of the best
      // so far and the served tasks, as testing them one by one gives.  The
      // lane reads wrap modulo 64, so lanes without task j read any lane.
      int bj = -1;
      for (int j = 0; j < kTriTake; ++j) {
        const bool has = j < take;
        if (__ballot(has) == 0) break;
        const float d2 = lane_read_f(td2, pre + j);
        const int id = static_cast<int>(lane_read(static_cast<uint32_t>(tid), pre + j));
        const bool better = has & ((d2 < bd2) | ((d2 == bd2) & (id < bid)));
        bd2 = better ? d2 : bd2;
        bid = better ? id : bid;
        bj = better ? j : bj;
      }
      // the closest point of the last (so the best) improvement, and the
      // threshold from the new best alone: one by one, each improvement
      // would lower it in turn to the threshold of a distance >= the new
      // best, and every such threshold is >= prune_threshold(new best) (see
      // walker_prune_threshold), so the search stays exact either way.  A
      // threshold can be NaN only for a NaN point (vertices are finite),
      // whose d2 is NaN: the select is fminf's result wherever it is used
      const V3 q = v3(lane_read_f(tq.x, pre + bj), lane_read_f(tq.y, pre + bj), lane_read_f(tq.z, pre + bj));
      const float ptn = walker_prune_threshold(bd2, a.delta);
      if (bj >= 0) {
        bcp = q;
        thr = ptn < thr ? ptn : thr;
      }
      if (phase == kTri) {
        if (COUNT) n_tris += static_cast<uint32_t>(take);
        cur += take;
        if (cur >= te) N2_POP_NEXT();
      }
    }
  }
#undef KA
#undef N2_ENTER
#undef N2_POP_NEXT
#undef N2_PUSH3
#undef N2_SPILL_LINE
#undef N2_NODE_VISIT
  if (COUNT) {
    atomicAdd(reinterpret_cast<unsigned long long*>(a.visits), static_cast<unsigned long long>(n_nodes));
    atomicAdd(reinterpret_cast<unsigned long long*>(a.visits) + 1, static_cast<unsigned long long>(n_tris));
  }
}


}  // namespace n2
