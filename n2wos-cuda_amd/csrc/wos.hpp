// wos.hpp -- argument blocks of the persistent walker.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "api_internal.hpp"
#include "bvh.hpp"
#include "geom.hpp"

namespace n2 {

// Units are split into at most kMaxUnitRanges contiguous ranges, one per XCD:
// a wave claims from its own XCD's range first (the points of a range share
// the mesh lines their walks touch, which then stay in that XCD's L2), and
// from the others once it is exhausted.  The counters are 32 bytes apart.
constexpr int kMaxUnitRanges = 8;
constexpr int kUnitCounterStride = 8;

// Claim order (spatial): 32 cells per axis over the mesh bounds, Morton order;
// solves of fewer points claim in the caller's order.
constexpr int kOrderAxis = 32;
constexpr int kOrderCells = kOrderAxis * kOrderAxis * kOrderAxis;
constexpr int kOrderMinPoints = 8192;

struct WosKernelArgs {
  const Bvh4Node* nodes;
  const TriRec* tris;
  float delta;
  Boundary bc;
  const float* points;
  int walks_per_point;
  int units_per_point;
  int total_units;
  int max_steps;
  float epsilon;
  float step_scale;
  uint32_t seed_lo, seed_hi;
  uint32_t stream;
  uint32_t point_offset;
  // optional, 2 per point: Philox counter point id and walk-index base (used
  // instead of point + point_offset and walk, so a point's walks are the
  // same whichever batch it runs in)
  const uint32_t* point_rng;
  // unit claim counters: range r of unit_ranges contiguous unit ranges is
  // claimed through next_unit[r * kUnitCounterStride]
  int* next_unit;
  int unit_ranges;
  // claim order: units are claimed in the order of positions, position i
  // holding point order[i] (null: point i), so that the walks in flight at
  // once start close together and share the mesh lines they touch
  const int* order;
  // walk values staged per unit slot until the unit's last walk ends:
  // [block][wave][slot][64 walks in the unit]
  float* unit_vals;
  // traversal-stack entries spilled from the LDS ring: kLineSpill per lane,
  // lane-major ([block][wave][lane][kLineSpill], 8 B each), so one lane's
  // spills share one 128-byte line
  uint64_t* spill_lines;
  double* unit_mean;
  double* unit_m2;
  uint32_t* unit_steps;
  uint32_t* unit_flags;  // forced count (low 16 bits) | overflow count << 16
  float* walk_values;    // optional per-walk outputs
  int* walk_steps;
  const float* first_d2;  // first query of each point (shared by its walks)
  const float* first_cp;
  const int* first_ovf;
  uint64_t* visits;       // optional: [node lines loaded, triangle records tested]
  int refill_min;         // refill when at least this many lanes are idle
  int weight_node;        // phase choice: lanes x weight, largest wins
  int weight_tri;
  int weight_update;
  int mode;               // kModePlain / kModePrefix / kModeCoupled
  int depth;              // m of the prefix / coupled modes
  int antithetic;         // 1: walks 2j and 2j+1 take mirrored first steps
  float* nc_prefix;       // per walk: prefix value (or the whole value if the walk ended first)
  float* nc_xm;           // per walk: X_m (3 floats)
  uint8_t* nc_need;       // per walk: 1 if the cache must be added at X_m
  float* nc_full;         // coupled mode: the full walk's value
  // extended walker only (plain mode): along-walk control variate over the
  // field cv_values ((cv_res + 1)^3 vertex values, cv_points 0 = off, 2 or 6)
  const float* cv_values;
  int cv_res;
  int cv_points;
  float cv_origin[3];
  float cv_inv_h[3];
  float cv_beta;
  // extended walker only: per walk, the first path_len visited positions X_k
  // and q_k = S_k - beta (control-variate terms of steps < k), so that the
  // walk's value minus q_k estimates the solution at X_k
  float* path_pos;
  float* path_q;
  int path_len;
  // packed units (pack_shift >= 0, walks_per_point = 2^pack_shift <= 32):
  // unit u holds the walks of points
  // [u 64 / wpp, (u + 1) 64 / wpp); unit arrays are indexed by point and the
  // step and flag counts are accumulated there (zeroed before the launch)
  int pack_shift;
  int n_points;
};

struct WosRun {
  const float* d_points;
  int n_points;
  int walks_per_point;
  uint32_t point_offset;
  const uint32_t* d_point_rng;
  int max_steps;
  float epsilon;
  float step_scale;
  uint64_t seed;
  uint32_t rng_stream;
  int source;
  int max_blocks;
  double* d_n;
  double* d_mean;
  double* d_m2;
  uint64_t* d_steps;
  uint32_t* d_forced;
  int accumulate;
  void* d_workspace;
  float* d_walk_values;
  int* d_walk_steps;
  uint64_t* d_visits;
  void* ev_begin;
  void* ev_end;
  int mode;
  int depth;
  int antithetic;
  float* nc_prefix;
  float* nc_xm;
  uint8_t* nc_need;
  float* nc_full;
  int ext;  // extended walker (control variate / path record)
  const float* cv_values;
  int cv_res;
  int cv_points;
  float cv_origin[3];
  float cv_inv_h[3];
  float cv_beta;
  float* path_pos;
  float* path_q;
  int path_len;
};

// Workspace of a walker job on the current device (its staging area for unit
// values is sized by the device's resident walker blocks).
int64_t wos_workspace_bytes(int n_points, int walks_per_point);
// Validated conversions of the ABI structs (throw Error on bad values).
Boundary to_boundary(const n2wos_boundary* boundary);
void check_walk_params(const n2wos_walk_params* wp);
// Blocks of a full-device walker launch (CUs x resident blocks per CU).
int wos_full_grid(int source, bool ext = false);
// Enqueues the stand-alone closest-point kernel (bvh.hip).
void launch_closest_point(const n2wos_bvh* bvh, const float* d_points, int n, float* d_d2, float* d_cp, int* d_tri,
                          int* d_ovf, hipStream_t stream);
// Tuning knobs set through n2wos_set_tuning (default when unset).
int tuning_value(const char* key, int fallback);
void run_wos(const n2wos_bvh* bvh, const Boundary& bc, const WosRun& r, hipStream_t stream);

}  // namespace n2
