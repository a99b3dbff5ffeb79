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
#include <cmath>
#include <mutex>
#include <string>

#include "api_internal.hpp"
#include "walker_kernel.hpp"

namespace n2 {

namespace {

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

__global__ void k_zero_counters(int* c) {
  if (threadIdx.x < kMaxUnitRanges) c[threadIdx.x * kUnitCounterStride] = 0;
}

__global__ void k_zero_u32(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------- claim order
// The walker claims units in the order of the points' cells on a 32^3 grid
// over the mesh bounds, the cells taken in Morton (Z) order: walks that run at
// the same time then start close together and touch the same mesh lines (on
// the 1 M-triangle mesh of problem 3, 14-17% faster than random point order).
// A counting sort: cells and counts, an exclusive scan of the counts, then a
// scatter.  The order inside a cell follows the atomics and may differ from run
// to run; results do not depend on it (each point's walks, statistics and
// merge order are fixed by the point alone).
__device__ __forceinline__ uint32_t spread3_5(uint32_t v) {
  // 5 bits b4..b0 -> b4 at bit 12, ..., b0 at bit 0 (every third bit)
  v &= 31u;
  v = (v | (v << 8)) & 0x100Fu;
  v = (v | (v << 4)) & 0x10C3u;
  v = (v | (v << 2)) & 0x1249u;
  return v;
}

__global__ void k_order_cells(const float* __restrict__ pts, int n, float lox, float loy, float loz, float sx,
                              float sy, float sz, uint32_t* __restrict__ cell, uint32_t* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // fmaxf(NaN, 0) = 0: a non-finite coordinate lands in cell 0
  const float fx = fminf(fmaxf((pts[3 * i] - lox) * sx, 0.0f), static_cast<float>(kOrderAxis - 1));
  const float fy = fminf(fmaxf((pts[3 * i + 1] - loy) * sy, 0.0f), static_cast<float>(kOrderAxis - 1));
  const float fz = fminf(fmaxf((pts[3 * i + 2] - loz) * sz, 0.0f), static_cast<float>(kOrderAxis - 1));
  const uint32_t c = spread3_5(static_cast<uint32_t>(fx)) | (spread3_5(static_cast<uint32_t>(fy)) << 1) |
                     (spread3_5(static_cast<uint32_t>(fz)) << 2);
  cell[i] = c;
  atomicAdd(count + c, 1u);
}

// One block of 1024 threads: exclusive prefix sum of the kOrderCells counts
// into `cursor` (each thread scans 32 consecutive cells).
__global__ void __launch_bounds__(1024) k_order_scan(const uint32_t* __restrict__ count,
                                                     uint32_t* __restrict__ cursor) {
  constexpr int kPer = kOrderCells / 1024;
  __shared__ uint32_t part[1024];
  const int t = static_cast<int>(threadIdx.x);
  uint32_t sum = 0;
  for (int k = 0; k < kPer; ++k) sum += count[t * kPer + k];
  part[t] = sum;
  __syncthreads();
  // Hillis-Steele inclusive scan of the 1024 partial sums
  for (int off = 1; off < 1024; off *= 2) {
    const uint32_t add = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
  for (int k = 0; k < kPer; ++k) {
    const uint32_t c = count[t * kPer + k];
    cursor[t * kPer + k] = run;
    run += c;
  }
}

__global__ void k_order_scatter(const uint32_t* __restrict__ cell, int n, uint32_t* __restrict__ cursor,
                                int* __restrict__ order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t pos = atomicAdd(cursor + cell[i], 1u);
  order[pos] = i;
}

// Packed units: walks_per_point a power of two <= 32 (log2 of it, else -1).
int pack_shift_for(int walks_per_point) {
  for (int k = 0; k <= 5; ++k)
    if (walks_per_point == (1 << k)) return k;
  return -1;
}

template <int CAP, int S, bool POISSON>
void launch_wos(const WosKernelArgs& a, int grid, hipStream_t stream, bool ext) {
  if (ext)
    hipLaunchKernelGGL((k_wos_persistent<CAP, S, POISSON, false, true>), dim3(grid), dim3(kBlock), 0, stream, a);
  else if (a.visits)
    hipLaunchKernelGGL((k_wos_persistent<CAP, S, POISSON, true>), dim3(grid), dim3(kBlock), 0, stream, a);
  else
    hipLaunchKernelGGL((k_wos_persistent<CAP, S, POISSON, false>), dim3(grid), dim3(kBlock), 0, stream, a);
}

template <int CAP, int S, bool POISSON, bool EXT>
int max_blocks_per_cu() {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_wos_persistent<CAP, S, POISSON, false, EXT>, kBlock,
                                                   0) != hipSuccess)
    return 1;
  return std::max(1, b);
}

// Kernel variant: LDS ring capacity ("ring": 16 or 8 entries per lane).
template <int S, bool POISSON>
int variant_blocks_per_cu(int cap, bool ext) {
  if (ext) return cap == 16 ? max_blocks_per_cu<16, S, POISSON, true>() : max_blocks_per_cu<8, S, POISSON, true>();
  return cap == 16 ? max_blocks_per_cu<16, S, POISSON, false>() : max_blocks_per_cu<8, S, POISSON, false>();
}

template <int S, bool POISSON>
void launch_variant(const WosKernelArgs& a, int grid, hipStream_t stream, int cap, bool ext) {
  if (cap == 16) launch_wos<16, S, POISSON>(a, grid, stream, ext);
  else launch_wos<8, S, POISSON>(a, grid, stream, ext);
}

}  // namespace

Boundary to_boundary(const n2wos_boundary* boundary) {
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
  return bc;
}

void check_walk_params(const n2wos_walk_params* wp) {
  N2_REQUIRE(wp->max_steps >= 0, N2WOS_ERR_INVALID, "max_steps must be >= 0");
  N2_REQUIRE(wp->epsilon >= 0.0f, N2WOS_ERR_INVALID, "epsilon must be >= 0");
  N2_REQUIRE(wp->step_scale > 0.0f && wp->step_scale <= 1.0f, N2WOS_ERR_INVALID, "step_scale must be in (0, 1]");
  N2_REQUIRE(wp->source == N2WOS_SOURCE_NONE || wp->source == N2WOS_SOURCE_POISSON_SINES, N2WOS_ERR_INVALID,
             "unknown source term");
}

constexpr int kSlots = 8;  // unit slots per wave

// LDS ring entries per lane: 16 (default) or 8 ("ring" tuning key).
int ring_capacity() { return tuning_value("ring", 16) == 8 ? 8 : 16; }

int device_cus() {
  int dev = 0, cus = 0;
  N2_HIP(hipGetDevice(&dev));
  N2_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return cus;
}

int wos_full_grid(int source, bool ext) {
  const bool poisson = source == N2WOS_SOURCE_POISSON_SINES;
  const int cap = ring_capacity();
  const int per_cu =
      poisson ? variant_blocks_per_cu<kSlots, true>(cap, ext) : variant_blocks_per_cu<kSlots, false>(cap, ext);
  return device_cus() * per_cu;
}

// Units of a job: 64-walk units, or packed units of 64 / wpp whole points.
int64_t job_units(int n_points, int walks_per_point) {
  const int ps = pack_shift_for(walks_per_point);
  if (ps >= 0) return (static_cast<int64_t>(n_points) + (64 >> ps) - 1) / (64 >> ps);
  return static_cast<int64_t>(n_points) * ((walks_per_point + 63) / 64);
}

// Walker blocks a launch of this job can have on the current device: at most
// one per unit and at most the largest full-device grid of any variant.
int64_t staging_blocks(int n_points, int walks_per_point) {
  static int per_cu_cache[64] = {};  // per device: the same for the process's lifetime
  int dev = 0;
  N2_HIP(hipGetDevice(&dev));
  int per_cu = dev < 64 ? per_cu_cache[dev] : 0;
  if (per_cu == 0) {
    per_cu = 1;
    for (int cap : {8, 16})
      for (bool ext : {false, true})
        per_cu = std::max({per_cu, variant_blocks_per_cu<kSlots, false>(cap, ext),
                           variant_blocks_per_cu<kSlots, true>(cap, ext)});
    if (dev < 64) per_cu_cache[dev] = per_cu;
  }
  const int64_t needed = (job_units(n_points, walks_per_point) + kWaves - 1) / kWaves;
  return std::max<int64_t>(1, std::min<int64_t>(needed, static_cast<int64_t>(device_cus()) * per_cu));
}

int64_t wos_workspace_bytes(int n_points, int walks_per_point) {
  const int64_t units = static_cast<int64_t>(n_points) * ((walks_per_point + 63) / 64);
  // counters (256 B pad) + mean + m2 (8 B each) + steps + flags (4 B each)
  // + first query per point: d2, closest xyz, tri, overflow (24 B), 256-aligned
  // + claim order: cell and order per point (8 B), cell counters (2 x 128 KiB)
  // + unit-value staging: 256 B per unit slot of every block the launch may have
  // + spill lines: 128 B per lane of those blocks
  const int64_t first = (static_cast<int64_t>(n_points) * 24 + 255) / 256 * 256;
  const int64_t order = (static_cast<int64_t>(n_points) * 8 + 255) / 256 * 256 + 2 * 4 * kOrderCells;
  const int64_t blocks = staging_blocks(n_points, walks_per_point);
  const int64_t staging = blocks * kWaves * kSlots * 64 * 4;
  const int64_t lines = blocks * kBlock * kLineSpill * 8;
  return 256 + (units * 24 + 255) / 256 * 256 + first + order + staging + lines;
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
  a.n_points = r.n_points;
  // packed units: a unit of fewer than 64 walks would hold one of the wave's
  // 8 unit slots for a few lanes (85,120 one-walk units of plain walks:
  // 10.9 ms with most lanes idle)
  a.pack_shift = pack_shift_for(r.walks_per_point);
  if (a.pack_shift >= 0) {
    const int ppu = 64 >> a.pack_shift;
    a.total_units = (r.n_points + ppu - 1) / ppu;
  }
  a.max_steps = r.max_steps;
  a.epsilon = r.epsilon;
  a.step_scale = r.step_scale;
  a.seed_lo = static_cast<uint32_t>(r.seed);
  a.seed_hi = static_cast<uint32_t>(r.seed >> 32);
  a.stream = r.rng_stream;
  a.point_offset = r.point_offset;
  a.point_rng = r.d_point_rng;
  a.walk_values = r.d_walk_values;
  a.walk_steps = r.d_walk_steps;
  char* ws = static_cast<char*>(r.d_workspace);
  // unit arrays: one entry per unit, or per point with packed units
  const int64_t units = static_cast<int64_t>(r.n_points) * a.units_per_point;
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
  char* ord = first + (static_cast<int64_t>(r.n_points) * 24 + 255) / 256 * 256;
  uint32_t* order_cell = reinterpret_cast<uint32_t*>(ord);
  int* order = reinterpret_cast<int*>(order_cell + r.n_points);
  uint32_t* order_count = reinterpret_cast<uint32_t*>(ord + (static_cast<int64_t>(r.n_points) * 8 + 255) / 256 * 256);
  uint32_t* order_cursor = order_count + kOrderCells;
  a.unit_vals = reinterpret_cast<float*>(reinterpret_cast<char*>(order_count) + 2 * 4 * kOrderCells);
  a.spill_lines = reinterpret_cast<uint64_t*>(
      a.unit_vals + staging_blocks(r.n_points, r.walks_per_point) * kWaves * kSlots * 64);
  a.order = nullptr;
  if (r.n_points >= kOrderMinPoints && tuning_value("spatial_order", 1) != 0) {
    // claim order by cell of the mesh bounds (the points themselves may lie outside)
    float lo[3], sc[3];
    for (int k = 0; k < 3; ++k) {
      const float ext = bvh->bounds_hi[k] - bvh->bounds_lo[k];
      lo[k] = bvh->bounds_lo[k];
      sc[k] = ext > 0.0f ? static_cast<float>(kOrderAxis) / ext : 0.0f;
    }
    const int ob = 256;
    const int og = (r.n_points + ob - 1) / ob;
    hipLaunchKernelGGL(k_zero_u32, dim3(kOrderCells / ob), dim3(ob), 0, stream, order_count, kOrderCells);
    hipLaunchKernelGGL(k_order_cells, dim3(og), dim3(ob), 0, stream, r.d_points, r.n_points, lo[0], lo[1], lo[2],
                       sc[0], sc[1], sc[2], order_cell, order_count);
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, stream, order_count, order_cursor);
    hipLaunchKernelGGL(k_order_scatter, dim3(og), dim3(ob), 0, stream, order_cell, r.n_points, order_cursor, order);
    a.order = order;
  }

  a.unit_ranges = std::max(1, std::min(kMaxUnitRanges, tuning_value("unit_ranges", kMaxUnitRanges)));
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, stream, a.next_unit);
  if (a.pack_shift >= 0) {
    // packed units accumulate step and flag counts per point
    const int zb = 256;
    hipLaunchKernelGGL(k_zero_u32, dim3((r.n_points + zb - 1) / zb), dim3(zb), 0, stream, a.unit_steps,
                       r.n_points);
    hipLaunchKernelGGL(k_zero_u32, dim3((r.n_points + zb - 1) / zb), dim3(zb), 0, stream, a.unit_flags,
                       r.n_points);
  }
  launch_closest_point(bvh, r.d_points, r.n_points, first_d2, first_cp, first_tri, first_ovf, stream);
  const bool poisson = r.source == N2WOS_SOURCE_POISSON_SINES;
  const int cap = ring_capacity();
  const bool ext = r.ext != 0;
  int grid = wos_full_grid(r.source, ext);
  if (r.max_blocks > 0) grid = std::min(grid, r.max_blocks);
  const int64_t needed = (static_cast<int64_t>(a.total_units) + kWaves - 1) / kWaves;
  grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(grid, needed)));
  // the staging area holds staging_blocks() blocks, which bounds the grid
  grid = static_cast<int>(std::min<int64_t>(grid, staging_blocks(r.n_points, r.walks_per_point)));
  a.visits = r.d_visits;
  a.refill_min = tuning_value("refill_min", 2);
  a.weight_node = tuning_value("weight_node", 10);
  // a triangle iteration tests up to 64 records of all the lanes in that
  // phase at once, so it pays to serve it with fewer lanes waiting there,
  // the more so the more records a leaf holds (measured on the benchmark
  // problem: leaves of 3, 14 is 2.6% faster than 10 and 2% faster than 24;
  // leaves of 4, 24 is 5% faster than 14 and 1.5% faster than 34)
  a.weight_tri = tuning_value("weight_tri", bvh->leaf_size <= 3 ? 14 : 24);
  // serving walk steps sooner keeps more lanes in the traversal phases
  // (measured: weight 30 is 1.5% faster than 10 on the benchmark problem)
  a.weight_update = tuning_value("weight_update", 30);
  a.mode = r.mode;
  a.depth = r.depth;
  a.antithetic = r.antithetic;
  a.nc_prefix = r.nc_prefix;
  a.nc_xm = r.nc_xm;
  a.nc_need = r.nc_need;
  a.nc_full = r.nc_full;
  a.cv_values = r.cv_values;
  a.cv_res = r.cv_res;
  a.cv_points = r.cv_points;
  for (int k = 0; k < 3; ++k) {
    a.cv_origin[k] = r.cv_origin[k];
    a.cv_inv_h[k] = r.cv_inv_h[k];
  }
  a.cv_beta = r.cv_beta;
  a.path_pos = r.path_pos;
  a.path_q = r.path_q;
  a.path_len = r.path_len;
  if (r.ev_begin) N2_HIP(hipEventRecord(static_cast<hipEvent_t>(r.ev_begin), stream));
  if (poisson) launch_variant<kSlots, true>(a, grid, stream, cap, ext);
  else launch_variant<kSlots, false>(a, grid, stream, cap, ext);
  if (r.ev_end) N2_HIP(hipEventRecord(static_cast<hipEvent_t>(r.ev_end), stream));
  N2_HIP(hipGetLastError());
  const int mb = 256;
  hipLaunchKernelGGL(k_point_merge, dim3((r.n_points + mb - 1) / mb), dim3(mb), 0, stream, r.n_points,
                     a.units_per_point, r.walks_per_point, r.accumulate, a.unit_mean, a.unit_m2, a.unit_steps,
                     a.unit_flags, r.d_n, r.d_mean, r.d_m2, r.d_steps, r.d_forced);
  N2_HIP(hipGetLastError());
}

// Per point: the estimator samples of prefix (coarse) or coupled (residual)
// walks combined with the cache values, reduced in walk order (Welford) and
// merged (Chan) into the point's (n, mean, M2):
//   prefix:  v = a + (need ? C(X_m) : 0)            (neural-cache estimate)
//   coupled: v = need ? W - (a + C(X_m)) : 0         (two-level residual)
// as the reference's combine kernel (src/cuda/cubql_bvh.cu:1608-1627).
__global__ void k_nc_combine(int n_points, int wpp, int mode, const float* __restrict__ prefix,
                             const uint8_t* __restrict__ need, const float* __restrict__ cache,
                             const float* __restrict__ full, int accumulate, double* __restrict__ out_n,
                             double* __restrict__ out_mean, double* __restrict__ out_m2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_points) return;
  double m = 0.0, q = 0.0;
  for (int w = 0; w < wpp; ++w) {
    const int64_t s = static_cast<int64_t>(i) * wpp + w;
    const float c = need[s] ? cache[s] : 0.0f;
    float v;
    if (mode == kModePrefix)
      v = prefix[s] + c;
    else
      v = need[s] ? full[s] - (prefix[s] + c) : 0.0f;
    const double x = static_cast<double>(v);
    const double dlt = x - m;
    m += dlt / static_cast<double>(w + 1);
    q += dlt * (x - m);
  }
  double N = static_cast<double>(wpp), M = m, Q = q;
  if (accumulate) chan_merge(out_n[i], out_mean[i], out_m2[i], N, M, Q, N, M, Q);
  out_n[i] = N;
  out_mean[i] = M;
  out_m2[i] = Q;
}

}  // namespace n2

using namespace n2;

extern "C" int n2wos_wos_full_grid_blocks(int source, int* blocks) {
  return guarded([&]() -> int {
    N2_REQUIRE(blocks, N2WOS_ERR_NULL, "NULL output");
    N2_REQUIRE(source == N2WOS_SOURCE_NONE || source == N2WOS_SOURCE_POISSON_SINES, N2WOS_ERR_INVALID,
               "unknown source term");
    *blocks = wos_full_grid(source);
    return N2WOS_OK;
  });
}

extern "C" int64_t n2wos_wos_workspace_size(int n_points, int walks_per_point) {
  if (n_points < 0 || walks_per_point < 1) return N2WOS_ERR_INVALID;
  int64_t bytes = 0;
  const int rc = guarded([&]() -> int {
    bytes = wos_workspace_bytes(n_points, walks_per_point);
    return N2WOS_OK;
  });
  return rc == N2WOS_OK ? bytes : rc;
}

namespace {

// The body of both plain-solve entry points; d_workspace holds at least
// wos_workspace_bytes(n_points, walks_per_point) bytes.
int wos_solve_impl(const n2wos_bvh* bvh, const n2wos_boundary* boundary, const n2wos_walk_params* wp,
                   const float* d_points, int n_points, int walks_per_point, uint32_t point_offset, double* d_n,
                   double* d_mean, double* d_m2, uint64_t* d_steps, uint32_t* d_forced, int accumulate,
                   void* d_workspace, float* d_walk_values, int* d_walk_steps, void* stream) {
  check_walk_params(wp);
  const Boundary bc = to_boundary(boundary);
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
  r.d_visits = static_cast<uint64_t*>(wp->d_visit_counts);
  r.mode = kModePlain;
  r.depth = 0;
  r.ev_begin = wp->event_begin;
  r.ev_end = wp->event_end;
  run_wos(bvh, bc, r, static_cast<hipStream_t>(stream));
  return N2WOS_OK;
}

// Argument checks shared by both plain-solve entry points (the workspace is
// checked by each).
void check_solve_args(const n2wos_bvh* bvh, const n2wos_boundary* boundary, const n2wos_walk_params* wp,
                      const float* d_points, int n_points, int walks_per_point, uint32_t point_offset, double* d_n,
                      double* d_mean, double* d_m2, uint64_t* d_steps, uint32_t* d_forced, float* d_walk_values,
                      int* d_walk_steps) {
  N2_REQUIRE(bvh && boundary && wp, N2WOS_ERR_NULL, "NULL handle, boundary or parameters");
  N2_REQUIRE(bvh->d_nodes && bvh->d_tris, N2WOS_ERR_EMPTY, "solve called on an empty BVH");
  N2_REQUIRE(n_points >= 0, N2WOS_ERR_INVALID, "n_points must be non-negative");
  N2_REQUIRE(walks_per_point >= 1, N2WOS_ERR_INVALID, "walks_per_point must be >= 1");
  if (n_points == 0) return;
  N2_REQUIRE(static_cast<int64_t>(n_points) * ((walks_per_point + 63) / 64) < (int64_t{1} << 31),
             N2WOS_ERR_INVALID, "too many work units for one launch");
  N2_REQUIRE(static_cast<int64_t>(n_points) + point_offset <= (int64_t{1} << 32), N2WOS_ERR_INVALID,
             "point ids exceed 32 bits");
  N2_REQUIRE(d_points && d_n && d_mean && d_m2 && d_steps && d_forced, N2WOS_ERR_NULL, "NULL device pointer");
  N2_REQUIRE((d_walk_values == nullptr) == (d_walk_steps == nullptr), N2WOS_ERR_NULL,
             "per-walk outputs must be both set or both NULL");
}

}  // namespace

extern "C" int n2wos_wos_solve(const n2wos_bvh* bvh, const n2wos_boundary* boundary, const n2wos_walk_params* wp,
                               const float* d_points, int n_points, int walks_per_point, uint32_t point_offset,
                               double* d_n, double* d_mean, double* d_m2, uint64_t* d_steps, uint32_t* d_forced,
                               int accumulate, void* d_workspace, int64_t workspace_bytes, float* d_walk_values,
                               int* d_walk_steps, void* stream) {
  return guarded([&]() -> int {
    check_solve_args(bvh, boundary, wp, d_points, n_points, walks_per_point, point_offset, d_n, d_mean, d_m2,
                     d_steps, d_forced, d_walk_values, d_walk_steps);
    if (n_points == 0) return N2WOS_OK;
    N2_REQUIRE(d_workspace, N2WOS_ERR_NULL, "NULL device pointer");
    N2_REQUIRE(workspace_bytes >= wos_workspace_bytes(n_points, walks_per_point), N2WOS_ERR_INVALID,
               "workspace too small (n2wos_wos_workspace_size)");
    return wos_solve_impl(bvh, boundary, wp, d_points, n_points, walks_per_point, point_offset, d_n, d_mean, d_m2,
                          d_steps, d_forced, accumulate, d_workspace, d_walk_values, d_walk_steps, stream);
  });
}

// ------------------------------------------------------------ no allocation
// The walker workspace of n2wos_wos_solve_noalloc: one file-scope device
// array, which every GPU the library runs on holds until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address on the current GPU comes from hipGetSymbolAddress, taken
// once per device on that device's first call and kept in scratch_ptr.
static __device__ __align__(256) unsigned char scratch[N2WOS_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of this GPU's scratch array (the current device's).
void* scratch_ptr() {
  int dev = 0;
  N2_HIP(hipGetDevice(&dev));
  N2_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, N2WOS_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    N2_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    N2_REQUIRE(p != nullptr, N2WOS_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" int n2wos_wos_solve_noalloc(const n2wos_bvh* bvh, const n2wos_boundary* boundary,
                                       const n2wos_walk_params* wp, const float* d_points, int n_points,
                                       int walks_per_point, uint32_t point_offset, double* d_n, double* d_mean,
                                       double* d_m2, uint64_t* d_steps, uint32_t* d_forced, int accumulate,
                                       float* d_walk_values, int* d_walk_steps, void* stream) {
  return guarded([&]() -> int {
    check_solve_args(bvh, boundary, wp, d_points, n_points, walks_per_point, point_offset, d_n, d_mean, d_m2,
                     d_steps, d_forced, d_walk_values, d_walk_steps);
    if (n_points == 0) return N2WOS_OK;
    const int64_t need = wos_workspace_bytes(n_points, walks_per_point);
    N2_REQUIRE(need <= N2WOS_NOALLOC_SCRATCH_BYTES, N2WOS_ERR_INVALID,
               "job needs " + std::to_string(need) + " bytes of walker workspace; n2wos_wos_solve_noalloc holds " +
                   std::to_string(static_cast<long long>(N2WOS_NOALLOC_SCRATCH_BYTES)) +
                   " (split the points into smaller calls, or use n2wos_wos_solve with a workspace)");
    // the unit counter is cleared by a kernel in run_wos; every other part
    // of the workspace is written before it is read
    return wos_solve_impl(bvh, boundary, wp, d_points, n_points, walks_per_point, point_offset, d_n, d_mean, d_m2,
                          d_steps, d_forced, accumulate, scratch_ptr(), d_walk_values, d_walk_steps, stream);
  });
}

extern "C" int n2wos_cv_solve(const n2wos_bvh* bvh, const n2wos_boundary* boundary, const n2wos_walk_params* wp,
                              const float* d_points, int n_points, int walks_per_point, uint32_t point_offset,
                              const n2wos_cv_field* field, double* d_n, double* d_mean, double* d_m2,
                              uint64_t* d_steps, uint32_t* d_forced, int accumulate, void* d_workspace,
                              int64_t workspace_bytes, float* d_walk_values, int* d_walk_steps, float* d_path_pos,
                              float* d_path_q, int path_len, void* stream) {
  return guarded([&]() -> int {
    check_solve_args(bvh, boundary, wp, d_points, n_points, walks_per_point, point_offset, d_n, d_mean, d_m2,
                     d_steps, d_forced, d_walk_values, d_walk_steps);
    N2_REQUIRE(path_len >= 0 && path_len <= 4096, N2WOS_ERR_INVALID, "path_len must be in [0, 4096]");
    const int points = field ? field->points : 0;
    N2_REQUIRE(points == 0 || points == 2 || points == 6, N2WOS_ERR_INVALID, "field points must be 0, 2 or 6");
    if (points > 0) {
      N2_REQUIRE(field->d_values, N2WOS_ERR_NULL, "NULL field values");
      N2_REQUIRE(field->res >= 1 && field->res <= 1024, N2WOS_ERR_INVALID, "field res must be in [1, 1024]");
      for (int k = 0; k < 3; ++k)
        N2_REQUIRE(std::isfinite(field->origin[k]) && std::isfinite(field->cell_inverse[k]) &&
                       field->cell_inverse[k] > 0.0f,
                   N2WOS_ERR_INVALID, "field box must be finite with positive cell_inverse");
      N2_REQUIRE(std::isfinite(field->beta), N2WOS_ERR_INVALID, "field beta must be finite");
    }
    if (path_len > 0)
      N2_REQUIRE(d_path_pos && d_path_q && d_walk_values, N2WOS_ERR_NULL,
                 "a path record needs d_path_pos, d_path_q and the per-walk outputs");
    if (n_points == 0) return N2WOS_OK;
    N2_REQUIRE(d_workspace, N2WOS_ERR_NULL, "NULL device pointer");
    N2_REQUIRE(workspace_bytes >= wos_workspace_bytes(n_points, walks_per_point), N2WOS_ERR_INVALID,
               "workspace too small (n2wos_wos_workspace_size)");
    check_walk_params(wp);
    const Boundary bc = to_boundary(boundary);
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
    r.d_visits = nullptr;
    r.mode = kModePlain;
    r.ev_begin = wp->event_begin;
    r.ev_end = wp->event_end;
    r.ext = 1;
    r.cv_points = points;
    r.cv_beta = points > 0 ? field->beta : 0.0f;
    if (points > 0) {
      r.cv_values = field->d_values;
      r.cv_res = field->res;
      for (int k = 0; k < 3; ++k) {
        r.cv_origin[k] = field->origin[k];
        r.cv_inv_h[k] = field->cell_inverse[k];
      }
    }
    r.path_pos = d_path_pos;
    r.path_q = d_path_q;
    r.path_len = path_len;
    run_wos(bvh, bc, r, static_cast<hipStream_t>(stream));
    return N2WOS_OK;
  });
}

extern "C" int n2wos_nc_walks(const n2wos_bvh* bvh, const n2wos_boundary* boundary, const n2wos_walk_params* wp,
                              const float* d_points, int n_points, int walks_per_point, uint32_t point_offset,
                              int mode, int depth_m, float* d_prefix, float* d_xm, uint8_t* d_need, float* d_full,
                              uint64_t* d_steps, void* d_workspace, int64_t workspace_bytes, void* stream) {
  return guarded([&]() -> int {
    const int antithetic = (mode & N2WOS_NC_ANTITHETIC) ? 1 : 0;
    mode &= ~N2WOS_NC_ANTITHETIC;
    N2_REQUIRE(mode == N2WOS_NC_PREFIX || mode == N2WOS_NC_COUPLED, N2WOS_ERR_INVALID, "mode must be prefix or coupled");
    N2_REQUIRE(depth_m >= 0, N2WOS_ERR_INVALID, "depth_m must be >= 0");
    N2_REQUIRE(bvh && boundary && wp, N2WOS_ERR_NULL, "NULL handle, boundary or parameters");
    N2_REQUIRE(n_points >= 0 && walks_per_point >= 1, N2WOS_ERR_INVALID, "bad sizes");
    if (n_points == 0) return N2WOS_OK;
    N2_REQUIRE(d_points && d_prefix && d_xm && d_need && d_workspace && d_steps, N2WOS_ERR_NULL,
               "NULL device pointer");
    N2_REQUIRE(mode == N2WOS_NC_PREFIX || d_full, N2WOS_ERR_NULL, "coupled mode needs d_full");
    N2_REQUIRE(workspace_bytes >= wos_workspace_bytes(n_points, walks_per_point) +
                                      static_cast<int64_t>(n_points) * 28 + 256,
               N2WOS_ERR_INVALID, "workspace too small (n2wos_nc_workspace_size)");
    check_walk_params(wp);
    const Boundary bc = to_boundary(boundary);
    // per-point statistics of the raw walk values go to scratch space after
    // the walker workspace (the estimator statistics come from n2wos_nc_combine)
    char* extra = static_cast<char*>(d_workspace) + wos_workspace_bytes(n_points, walks_per_point);
    extra = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(extra) + 255) & ~uintptr_t{255});
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
    r.d_n = reinterpret_cast<double*>(extra);
    r.d_mean = r.d_n + n_points;
    r.d_m2 = r.d_mean + n_points;
    r.d_steps = d_steps;
    r.d_forced = reinterpret_cast<uint32_t*>(r.d_m2 + n_points);
    r.accumulate = 0;
    r.d_workspace = d_workspace;
    r.d_walk_values = nullptr;
    r.d_walk_steps = nullptr;
    r.d_visits = static_cast<uint64_t*>(wp->d_visit_counts);
    r.ev_begin = wp->event_begin;
    r.ev_end = wp->event_end;
    r.mode = mode == N2WOS_NC_PREFIX ? kModePrefix : kModeCoupled;
    r.depth = depth_m;
    r.antithetic = antithetic;
    r.nc_prefix = d_prefix;
    r.nc_xm = d_xm;
    r.nc_need = d_need;
    r.nc_full = d_full;
    run_wos(bvh, bc, r, static_cast<hipStream_t>(stream));
    return N2WOS_OK;
  });
}

extern "C" int64_t n2wos_nc_workspace_size(int n_points, int walks_per_point) {
  const int64_t w = n2wos_wos_workspace_size(n_points, walks_per_point);
  return w < 0 ? w : w + static_cast<int64_t>(n_points) * 28 + 256;
}

extern "C" int n2wos_nc_combine(int n_points, int walks_per_point, int mode, const float* d_prefix,
                                const uint8_t* d_need, const float* d_cache, const float* d_full, int accumulate,
                                double* d_n, double* d_mean, double* d_m2, void* stream) {
  return guarded([&]() -> int {
    N2_REQUIRE(mode == N2WOS_NC_PREFIX || mode == N2WOS_NC_COUPLED, N2WOS_ERR_INVALID, "mode must be prefix or coupled");
    N2_REQUIRE(n_points >= 0 && walks_per_point >= 1, N2WOS_ERR_INVALID, "bad sizes");
    if (n_points == 0) return N2WOS_OK;
    N2_REQUIRE(d_prefix && d_need && d_cache && d_n && d_mean && d_m2, N2WOS_ERR_NULL, "NULL device pointer");
    N2_REQUIRE(mode == N2WOS_NC_PREFIX || d_full, N2WOS_ERR_NULL, "coupled mode needs d_full");
    const int b = 256;
    hipLaunchKernelGGL(k_nc_combine, dim3((n_points + b - 1) / b), dim3(b), 0, static_cast<hipStream_t>(stream),
                       n_points, walks_per_point, mode == N2WOS_NC_PREFIX ? kModePrefix : kModeCoupled, d_prefix,
                       d_need, d_cache, d_full, accumulate, d_n, d_mean, d_m2);
    N2_HIP(hipGetLastError());
    return N2WOS_OK;
  });
}
