// wos.hip -- persistent Walk-on-Spheres kernel and per-point FP64 statistics.
//
// Work decomposition.  A job is P points x W walks.  Walks are grouped into
// work units of up to 64 consecutive walk indices of one point (unit u covers
// point u / U, walks [64 (u % U), ...), U = ceil(W / 64)).  The grid is
// persistent: every wave loops until the global unit counter is exhausted.
//
// Per wave (no block-level synchronisation anywhere):
//   * S unit slots in LDS, each with the unit id, the number of walks still
//     running, and a 64-entry value array indexed by walk-in-unit;
//   * a "fill" cursor into the newest unit.  Idle lanes are refilled every
//     iteration: a ballot of idle lanes, each lane's rank among them (mbcnt),
//     and consecutive walk indices handed out by rank; when the fill unit is
//     exhausted lane 0 claims the next unit with one atomic on the global
//     counter.  This is the compaction/refill step: finished walks never leave
//     lanes idle while work remains.
//   * every active lane advances its walk by one step per iteration (one
//     exact closest-point query, see traverse.hpp), so short and long walks
//     share the wave without waiting for each other;
//   * when the last walk of a unit finishes, the whole wave reduces the
//     unit's 64 values with a fixed Chan-merge tree in FP64 (strides 1..32,
//     lower lane keeps the result) and writes (mean, M2) for the unit.
// A second kernel merges each point's units left to right (Chan), optionally
// into existing per-point (n, mean, M2) state.  Every reduction order is fixed,
// so results are bit-reproducible regardless of scheduling.
//
// Walk semantics follow the reference's walk loop (src/cuda/cubql_bvh.cu:1323-1357):
//   step = 0..max_steps: query; d = sqrt(max(d2, 0));
//   stop with src + g(closest point) when d <= eps, step == max_steps or d
//   is not finite; otherwise R = step_scale d, add the ball's source term, and
//   move to p + R dir.
// Randomness is Philox4x32-10 keyed by the 64-bit seed with counter
// (step, walk, point, 2 stream) for the direction and (.., 2 stream + 1) for
// the Poisson source sample (Green's-function sampling: radius fraction
// ~ Beta(2,2) as the median of three uniforms, uniform direction,
// contribution R^2/6 f(y)).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "api_internal.hpp"
#include "traverse.hpp"
#include "wos.hpp"

namespace n2 {

namespace {

__device__ __forceinline__ int lane_id() {
  return static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)));
}

// number of set bits of `mask` below this lane
__device__ __forceinline__ int rank_in(uint64_t mask) {
  return static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u)));
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
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

constexpr int kWaves = 4;        // waves per block
constexpr int kBlock = 64 * kWaves;

template <int CAP, int S, bool POISSON>
__global__ void __launch_bounds__(kBlock) k_wos_persistent(WosKernelArgs a) {
  __shared__ uint32_t s_ref[CAP * kBlock];
  __shared__ float s_d2[CAP * kBlock];
  __shared__ float s_val[kWaves][S][64];
  __shared__ int s_unit[kWaves][S];
  __shared__ int s_rem[kWaves][S];
  __shared__ uint32_t s_steps[kWaves][S];
  __shared__ uint32_t s_forced[kWaves][S];
  __shared__ uint32_t s_ovf[kWaves][S];

  const int lane = lane_id();
  const int wv = static_cast<int>(threadIdx.x) >> 6;
  uint32_t* sref = s_ref + threadIdx.x;
  float* sd2 = s_d2 + threadIdx.x;
  if (lane < S) s_unit[wv][lane] = -1;
  wave_sync();

  // wave-uniform fill cursor
  int fill_slot = -1, fill_next = 0, fill_count = 0, fill_point = 0, fill_w0 = 0, fill_unit = 0;
  bool no_more = false;

  // lane state
  bool active = false;
  int slot = 0, wiu = 0, point = 0, walk = 0, step = 0;
  V3 p = v3(0.0f, 0.0f, 0.0f);
  float src = 0.0f;
  float thr0 = FLT_MAX;

  while (true) {
    // ---------------------------------------------------------- refill
    uint64_t idle = __ballot(!active);
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
        int u = 0;
        if (lane == 0) u = atomicAdd(a.next_unit, 1);
        u = __shfl(u, 0);
        if (u >= a.total_units) {
          no_more = true;
          break;
        }
        const int pt = u / a.units_per_point;
        const int chunk = u - pt * a.units_per_point;
        const int cnt = min(64, a.walks_per_point - 64 * chunk);
        if (lane == 0) {
          s_unit[wv][fs] = u;
          s_rem[wv][fs] = cnt;
          s_steps[wv][fs] = 0u;
          s_forced[wv][fs] = 0u;
          s_ovf[wv][fs] = 0u;
        }
        wave_sync();
        fill_slot = fs;
        fill_next = 0;
        fill_count = cnt;
        fill_point = pt;
        fill_w0 = 64 * chunk;
        fill_unit = u;
      }
      const int take = min(__popcll(idle), fill_count - fill_next);
      const bool is_idle = (idle >> lane) & 1ull;
      const int r = rank_in(idle);
      if (is_idle && r < take) {
        active = true;
        slot = fill_slot;
        wiu = fill_next + r;
        point = fill_point;
        walk = fill_w0 + wiu;
        step = 0;
        src = 0.0f;
        thr0 = FLT_MAX;
        p = v3(a.points[3 * point], a.points[3 * point + 1], a.points[3 * point + 2]);
      }
      fill_next += take;
      idle = __ballot(!active);
    }
    (void)fill_unit;
    if (__ballot(active) == 0) break;

    // ---------------------------------------------------------- one step
    bool unit_done = false;
    if (active) {
      Hit h;
      if (step == 0) {
        // every walk of a point shares its first query: computed once per point
        h.d2 = a.first_d2[point];
        h.cp = v3(a.first_cp[3 * point], a.first_cp[3 * point + 1], a.first_cp[3 * point + 2]);
        h.id = 0;
        h.overflow = a.first_ovf[point];
      } else if (finite3(p)) {
        h = bvh4_closest<CAP, false>(a.nodes, a.tris, p, a.delta, thr0, sref, sd2, kBlock);
        if (h.id < 0 && thr0 < FLT_MAX)  // bound too tight (cannot happen in exact arithmetic)
          h = bvh4_closest<CAP, false>(a.nodes, a.tris, p, a.delta, FLT_MAX, sref, sd2, kBlock);
      } else {
        h.d2 = FLT_MAX; h.cp = v3(0.0f, 0.0f, 0.0f); h.id = -1; h.overflow = 0;
      }
      const float d = sqrtf(fmaxf(h.d2, 0.0f));
      const bool forced = step >= a.max_steps;
      if (d <= a.epsilon || forced || !isfinite(d)) {
        const float value = src + boundary_value(a.bc, h.cp);
        s_val[wv][slot][wiu] = value;
        atomicAdd(&s_steps[wv][slot], static_cast<uint32_t>(step));
        if (forced) atomicAdd(&s_forced[wv][slot], 1u);
        if (h.overflow) atomicAdd(&s_ovf[wv][slot], 1u);
        if (a.walk_values) {
          const int64_t sidx = static_cast<int64_t>(point) * a.walks_per_point + walk;
          a.walk_values[sidx] = value;
          a.walk_steps[sidx] = step;
        }
        const int old = atomicSub(&s_rem[wv][slot], 1);
        unit_done = (old == 1);
        active = false;
      } else {
        const uint32_t gp = static_cast<uint32_t>(point) + a.point_offset;
        const U4 rr = philox4x32_10(U4{static_cast<uint32_t>(step), static_cast<uint32_t>(walk), gp, 2u * a.stream},
                                    a.seed_lo, a.seed_hi);
        const V3 dir = sphere_dir(rr.x, rr.y);
        const float R = a.step_scale * d;
        if (POISSON) {
          const U4 t = philox4x32_10(
              U4{static_cast<uint32_t>(step), static_cast<uint32_t>(walk), gp, 2u * a.stream + 1u}, a.seed_lo,
              a.seed_hi);
          const float rho = median3(u01(rr.z), u01(rr.w), u01(t.x));
          const V3 e = sphere_dir(t.y, t.z);
          const float Rr = R * rho;
          const V3 y = v3(p.x + Rr * e.x, p.y + Rr * e.y, p.z + Rr * e.z);
          src = src + (R * R) * poisson_source(y) / 6.0f;
        }
        p = v3(p.x + R * dir.x, p.y + R * dir.y, p.z + R * dir.z);
        ++step;
        // |p_new - closest| <= d + |R dir|: bound the next query (loose margin
        // covers FP32 rounding of d, R, dir and p)
        const float t = (d + R) * 1.0000152587890625f + 4.0f * a.delta;
        thr0 = t * t;
      }
    }
    wave_sync();

    // ---------------------------------------------------------- finish units
    uint64_t done = __ballot(unit_done);
    while (done != 0) {
      const int l = __ffsll(static_cast<unsigned long long>(done)) - 1;
      done &= done - 1;
      const int s = __shfl(slot, l);
      const int u = s_unit[wv][s];
      const int pt = u / a.units_per_point;
      const int cnt = min(64, a.walks_per_point - 64 * (u - pt * a.units_per_point));
      double n = lane < cnt ? 1.0 : 0.0;
      double m = lane < cnt ? static_cast<double>(s_val[wv][s][lane]) : 0.0;
      double q = 0.0;
#pragma unroll
      for (int stride = 1; stride < 64; stride *= 2) {
        const double nb = __shfl_down(n, stride);
        const double mb = __shfl_down(m, stride);
        const double qb = __shfl_down(q, stride);
        if ((lane & (2 * stride - 1)) == 0) chan_merge(n, m, q, nb, mb, qb, n, m, q);
      }
      if (lane == 0) {
        a.unit_mean[u] = m;
        a.unit_m2[u] = q;
        a.unit_steps[u] = s_steps[wv][s];
        a.unit_flags[u] = (s_forced[wv][s] & 0xffffu) | (s_ovf[wv][s] << 16);
        s_unit[wv][s] = -1;
      }
      wave_sync();
    }
  }
}

// Merge each point's units left to right into its (n, mean, M2) state.
__global__ void k_point_merge(int n_points, int units_per_point, int walks_per_point, int accumulate,
                              const double* __restrict__ unit_mean, const double* __restrict__ unit_m2,
                              const uint32_t* __restrict__ unit_steps, const uint32_t* __restrict__ unit_flags,
                              double* __restrict__ out_n, double* __restrict__ out_mean, double* __restrict__ out_m2,
                              uint64_t* __restrict__ out_steps, uint32_t* __restrict__ out_forced) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_points) return;
  double N = 0.0, M = 0.0, Q = 0.0;
  uint64_t steps = 0;
  uint32_t forced = 0;
  for (int c = 0; c < units_per_point; ++c) {
    const int64_t u = static_cast<int64_t>(i) * units_per_point + c;
    const int cnt = min(64, walks_per_point - 64 * c);
    chan_merge(N, M, Q, static_cast<double>(cnt), unit_mean[u], unit_m2[u], N, M, Q);
    steps += unit_steps[u];
    forced += unit_flags[u] & 0xffffu;
  }
  if (accumulate) {
    double n0 = out_n[i], m0 = out_mean[i], q0 = out_m2[i];
    chan_merge(n0, m0, q0, N, M, Q, N, M, Q);
    steps += out_steps[i];
    forced += out_forced[i];
  }
  out_n[i] = N;
  out_mean[i] = M;
  out_m2[i] = Q;
  out_steps[i] = steps;
  out_forced[i] = forced;
}

__global__ void k_zero_counter(int* c) { *c = 0; }

template <int CAP, int S, bool POISSON>
void launch_wos(const WosKernelArgs& a, int grid, hipStream_t stream) {
  hipLaunchKernelGGL((k_wos_persistent<CAP, S, POISSON>), dim3(grid), dim3(kBlock), 0, stream, a);
}

template <int CAP, int S, bool POISSON>
int max_blocks_per_cu() {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_wos_persistent<CAP, S, POISSON>, kBlock, 0) != hipSuccess)
    return 1;
  return std::max(1, b);
}

}  // namespace

int64_t wos_workspace_bytes(int n_points, int walks_per_point) {
  const int64_t units = static_cast<int64_t>(n_points) * ((walks_per_point + 63) / 64);
  // counter (256 B pad) + mean + m2 (8 B each) + steps + flags (4 B each)
  // + first query per point: d2, closest xyz, tri, overflow (24 B), 256-aligned
  const int64_t first = (static_cast<int64_t>(n_points) * 24 + 255) / 256 * 256;
  return 256 + (units * 24 + 255) / 256 * 256 + first;
}

void run_wos(const n2wos_bvh* bvh, const Boundary& bc, const WosRun& r, hipStream_t stream) {
  WosKernelArgs a{};
  a.nodes = bvh->d_nodes;
  a.tris = bvh->d_tris;
  a.delta = bvh->delta;
  a.bc = bc;
  a.points = r.d_points;
  a.walks_per_point = r.walks_per_point;
  a.units_per_point = (r.walks_per_point + 63) / 64;
  a.total_units = r.n_points * a.units_per_point;
  a.max_steps = r.max_steps;
  a.epsilon = r.epsilon;
  a.step_scale = r.step_scale;
  a.seed_lo = static_cast<uint32_t>(r.seed);
  a.seed_hi = static_cast<uint32_t>(r.seed >> 32);
  a.stream = r.rng_stream;
  a.point_offset = r.point_offset;
  a.walk_values = r.d_walk_values;
  a.walk_steps = r.d_walk_steps;
  char* ws = static_cast<char*>(r.d_workspace);
  const int64_t units = a.total_units;
  a.next_unit = reinterpret_cast<int*>(ws);
  a.unit_mean = reinterpret_cast<double*>(ws + 256);
  a.unit_m2 = a.unit_mean + units;
  a.unit_steps = reinterpret_cast<uint32_t*>(a.unit_m2 + units);
  a.unit_flags = a.unit_steps + units;
  char* first = ws + 256 + (units * 24 + 255) / 256 * 256;
  float* first_d2 = reinterpret_cast<float*>(first);
  float* first_cp = first_d2 + r.n_points;
  int* first_tri = reinterpret_cast<int*>(first_cp + 3 * static_cast<int64_t>(r.n_points));
  int* first_ovf = first_tri + r.n_points;
  a.first_d2 = first_d2;
  a.first_cp = first_cp;
  a.first_ovf = first_ovf;

  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, stream, a.next_unit);
  launch_closest_point(bvh, r.d_points, r.n_points, first_d2, first_cp, first_tri, first_ovf, stream);
  int dev = 0, cus = 0;
  N2_HIP(hipGetDevice(&dev));
  N2_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  constexpr int kSlots = 8;
  const bool poisson = r.source == N2WOS_SOURCE_POISSON_SINES;
  const int cap = tuning_value("ring", 8);
  int per_cu;
  if (cap == 16)
    per_cu = poisson ? max_blocks_per_cu<16, kSlots, true>() : max_blocks_per_cu<16, kSlots, false>();
  else
    per_cu = poisson ? max_blocks_per_cu<8, kSlots, true>() : max_blocks_per_cu<8, kSlots, false>();
  int grid = cus * per_cu;
  if (r.max_blocks > 0) grid = std::min(grid, r.max_blocks);
  const int64_t needed = (units + kWaves - 1) / kWaves;
  grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(grid, needed)));
  if (cap == 16) {
    if (poisson) launch_wos<16, kSlots, true>(a, grid, stream); else launch_wos<16, kSlots, false>(a, grid, stream);
  } else {
    if (poisson) launch_wos<8, kSlots, true>(a, grid, stream); else launch_wos<8, kSlots, false>(a, grid, stream);
  }
  N2_HIP(hipGetLastError());
  const int mb = 256;
  hipLaunchKernelGGL(k_point_merge, dim3((r.n_points + mb - 1) / mb), dim3(mb), 0, stream, r.n_points,
                     a.units_per_point, r.walks_per_point, r.accumulate, a.unit_mean, a.unit_m2, a.unit_steps,
                     a.unit_flags, r.d_n, r.d_mean, r.d_m2, r.d_steps, r.d_forced);
  N2_HIP(hipGetLastError());
}

}  // namespace n2

using namespace n2;

extern "C" int64_t n2wos_wos_workspace_size(int n_points, int walks_per_point) {
  if (n_points < 0 || walks_per_point < 1) return N2WOS_ERR_INVALID;
  return wos_workspace_bytes(n_points, walks_per_point);
}

extern "C" int n2wos_wos_solve(const n2wos_bvh* bvh, const n2wos_boundary* boundary, const n2wos_walk_params* wp,
                               const float* d_points, int n_points, int walks_per_point, uint32_t point_offset,
                               double* d_n, double* d_mean, double* d_m2, uint64_t* d_steps, uint32_t* d_forced,
                               int accumulate, void* d_workspace, int64_t workspace_bytes, float* d_walk_values,
                               int* d_walk_steps, void* stream) {
  return guarded([&]() -> int {
    N2_REQUIRE(bvh && boundary && wp, N2WOS_ERR_NULL, "NULL handle, boundary or parameters");
    N2_REQUIRE(bvh->d_nodes && bvh->d_tris, N2WOS_ERR_EMPTY, "solve called on an empty BVH");
    N2_REQUIRE(n_points >= 0, N2WOS_ERR_INVALID, "n_points must be non-negative");
    N2_REQUIRE(walks_per_point >= 1, N2WOS_ERR_INVALID, "walks_per_point must be >= 1");
    if (n_points == 0) return N2WOS_OK;
    N2_REQUIRE(static_cast<int64_t>(n_points) * ((walks_per_point + 63) / 64) < (int64_t{1} << 31),
               N2WOS_ERR_INVALID, "too many work units for one launch");
    N2_REQUIRE(static_cast<int64_t>(n_points) + point_offset <= (int64_t{1} << 32), N2WOS_ERR_INVALID,
               "point ids exceed 32 bits");
    N2_REQUIRE(d_points && d_n && d_mean && d_m2 && d_steps && d_forced && d_workspace, N2WOS_ERR_NULL,
               "NULL device pointer");
    N2_REQUIRE((d_walk_values == nullptr) == (d_walk_steps == nullptr), N2WOS_ERR_NULL,
               "per-walk outputs must be both set or both NULL");
    N2_REQUIRE(workspace_bytes >= wos_workspace_bytes(n_points, walks_per_point), N2WOS_ERR_INVALID,
               "workspace too small (n2wos_wos_workspace_size)");
    N2_REQUIRE(wp->max_steps >= 0, N2WOS_ERR_INVALID, "max_steps must be >= 0");
    N2_REQUIRE(wp->epsilon >= 0.0f, N2WOS_ERR_INVALID, "epsilon must be >= 0");
    N2_REQUIRE(wp->step_scale > 0.0f && wp->step_scale <= 1.0f, N2WOS_ERR_INVALID, "step_scale must be in (0, 1]");
    N2_REQUIRE(wp->source == N2WOS_SOURCE_NONE || wp->source == N2WOS_SOURCE_POISSON_SINES, N2WOS_ERR_INVALID,
               "unknown source term");
    N2_REQUIRE(boundary->n_charges >= 0 && boundary->n_charges <= kMaxCharges, N2WOS_ERR_INVALID,
               "n_charges must be in [0, 16]");
    Boundary bc{};
    bc.mode = boundary->mode;
    bc.n_charges = boundary->n_charges;
    for (int i = 0; i < bc.n_charges; ++i)
      for (int k = 0; k < 4; ++k) bc.charge[i][k] = boundary->charges[i][k];
    switch (bc.mode) {
      case kBcX2MinusY2: case kBcZebraK4: case kBcZebraK8: case kBcZebraK12: case kBcConstantOne:
      case kBcCharges: case kBcPoissonSines: break;
      default: throw Error(N2WOS_ERR_INVALID, "unknown boundary mode");
    }
    WosRun r{};
    r.d_points = d_points;
    r.n_points = n_points;
    r.walks_per_point = walks_per_point;
    r.point_offset = point_offset;
    r.max_steps = wp->max_steps;
    r.epsilon = wp->epsilon;
    r.step_scale = wp->step_scale;
    r.seed = wp->seed;
    r.rng_stream = wp->rng_stream;
    r.source = wp->source;
    r.max_blocks = wp->max_blocks;
    r.d_n = d_n;
    r.d_mean = d_mean;
    r.d_m2 = d_m2;
    r.d_steps = d_steps;
    r.d_forced = d_forced;
    r.accumulate = accumulate;
    r.d_workspace = d_workspace;
    r.d_walk_values = d_walk_values;
    r.d_walk_steps = d_walk_steps;
    run_wos(bvh, bc, r, static_cast<hipStream_t>(stream));
    return N2WOS_OK;
  });
}
