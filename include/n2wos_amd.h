/*
 * n2wos_amd.h -- C ABI of the MI355X (gfx950) Walk-on-Spheres engine.
 *
 * Conventions
 *   - Every entry point returns 0 (N2WOS_OK) or a negative N2WOS_ERR_* code;
 *     n2wos_last_error() then describes the failure (thread-local string).
 *     No C++ exception crosses this boundary.
 *   - Pointers named d_* are HIP device pointers owned by the caller.  Calls
 *     that take a stream only enqueue work on it: they neither allocate,
 *     transfer nor synchronise unless their comment says otherwise.
 *   - Points are arrays of float triples (x, y, z), 12 bytes per point, the
 *     element layout of the reference's DeviceVec3 buffers.
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   n2wos_bvh_create / _destroy      CudaBvh::CudaBvh / ~CudaBvh
 *                                    (src/cuda/bvh_query.cu:171-226,
 *                                    include/n2wos/cuda_bvh.hpp:39-42) and
 *                                    CuBqlBvh::CuBqlBvh (include/n2wos/cubql_bvh.hpp:33-35)
 *   n2wos_bvh_query_device           CudaBvh::query_device (include/n2wos/cuda_bvh.hpp:50-57,
 *                                    src/cuda/bvh_query.cu:228-261)
 *   n2wos_bvh_query_host             CudaBvh::query (include/n2wos/cuda_bvh.hpp:44,
 *                                    src/cuda/bvh_query.cu:263-335)
 *   n2wos_bvh_query_device_masked    CuBqlBvh::query_device_masked
 *                                    (include/n2wos/cubql_bvh.hpp:57-65,
 *                                    src/cuda/cubql_bvh.cu:680-714)
 *   n2wos_bvh_get_stats              triangle_count/node_count/index_count/leaf_size/
 *                                    max_depth accessors (include/n2wos/cuda_bvh.hpp:59-63)
 *   n2wos_uv_sphere                  make_procedural_bumpy_sphere
 *                                    (src/mesh/procedural_bumpy_sphere.cpp:32-90)
 *   n2wos_normalize_to_unit_radius   normalize_to_unit_radius (src/mesh/normalize.cpp:18-37)
 *   n2wos_count_degenerate_triangles count_degenerate_triangles
 *                                    (src/geometry/closest_point.cpp:70-83)
 *   n2wos_bvh_inside_device          point_inside_closed_mesh_ray_x
 *                                    (src/cuda/cubql_bvh.cu:407-455, eval_tcnn_nc_wos.cu:399-400)
 *   n2wos_wos_solve                  run_persistent_harmonic (include/n2wos/wos_wavefront.hpp:120-122,
 *                                    src/cuda/cubql_bvh.cu:819-960) and launch_nc_pure_wos_samples +
 *                                    host summarize (include/n2wos/tcnn_nc_wos.hpp:79-134,
 *                                    src/tools/eval_tcnn_nc_wos.cu:525-597): walks plus per-point
 *                                    FP64 mean/variance on the device
 */
#ifndef N2WOS_AMD_H
#define N2WOS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define N2WOS_ABI_VERSION 1

enum {
  N2WOS_OK = 0,
  N2WOS_ERR_INVALID = -1,     /* bad argument value (negative count, bad option) */
  N2WOS_ERR_NULL = -2,        /* required pointer is NULL */
  N2WOS_ERR_BLOCK = -3,       /* block size outside [1, 1024] */
  N2WOS_ERR_HIP = -4,         /* HIP runtime or launch failure */
  N2WOS_ERR_EMPTY = -5,       /* handle holds no geometry */
  N2WOS_ERR_NOMEM = -6,       /* host or device allocation failed */
  N2WOS_ERR_UNSUPPORTED = -7, /* option not implemented (e.g. build method) */
  N2WOS_ERR_INTERNAL = -8
};

typedef struct n2wos_bvh n2wos_bvh;

/* Description of the last failure on the calling thread ("" if none). */
const char* n2wos_last_error(void);
int n2wos_abi_version(void);
/* Launch tuning for benchmarks (process-wide): "ring" = LDS stack entries
 * per lane of the walker (8 or 16, default 16); "refill_min" = idle lanes
 * that trigger a refill (1..64, default 2); "weight_node" / "weight_tri" /
 * "weight_update" = per-lane weights of the walker's phase choice (10, 14 for
 * leaves of up to 3 triangles else 24, and 30); "query_quad" = 1 (default):
 * closest-point queries run on four lanes each, 0: one lane per query with
 * the caller's block size (same results); "query_cap" = node visits and
 * leaves after which a four-lane query stops and a second pass finishes it
 * with a whole wave (default 16; 0: no second pass);
 * "query_finish_slots" = consecutive query slots each wave of that pass
 * scans (1..64, default 8);
 * "spatial_order" = 1 (default): solves of at least 8,192 points claim their
 * work in Morton order of 32^3 cells over the mesh bounds, 0: in the
 * caller's point order; "unit_ranges" = 1..8 (default 8): the claim order is
 * split into that many ranges, a wave claiming from its XCD's range first.
 * Neither changes any result. */
int n2wos_set_tuning(const char* key, int value);
/* Device name and runtime version of the current HIP device, for reports. */
int n2wos_device_summary(char* buf, int buf_len);

/* ----------------------------------------------------------------- meshes */
/* UV-sphere radius kinds. */
enum {
  N2WOS_SPHERE_REFERENCE_BUMP = 0, /* reference bump_radius, params[0] = amplitude */
  N2WOS_SPHERE_SIN6 = 1,           /* 1 + params[0] sin(6 theta) sin(6 phi) */
  N2WOS_SPHERE_MULTI = 2,          /* 1 + .06 sin5t sin7p + .03 sin17t cos13p */
  N2WOS_SPHERE_SH_BLOB = 3         /* 1 + sum_{l<=8} params[l*l+l+m] Y_lm */
};
/* Vertex and triangle counts: nv = 2 + (stacks-1) slices, nt = 2 slices (stacks-1). */
int n2wos_uv_sphere_counts(int stacks, int slices, int* nv, int* nt);
/* Fills verts[3 nv] and tris[3 nt] (vertex and face order of the reference). */
int n2wos_uv_sphere(int stacks, int slices, int kind, const float* params, int n_params, float* verts,
                    uint32_t* tris);
/* In place: p <- (p - c) / R, c = AABB midpoint, R = max |p - c|.  Writes c and 1/R. */
int n2wos_normalize_to_unit_radius(float* verts, int nv, float* center, float* scale);
/* Number of triangles with |(b-a) x (c-a)|^2 <= area2_eps. */
int64_t n2wos_count_degenerate_triangles(const float* verts, int nv, const uint32_t* tris, int nt,
                                         float area2_eps);

/* Mesh files, replacing load_obj_mesh (src/mesh/obj_loader.cpp:63-107) and
 * load_ply_mesh (src/mesh/ply_loader.cpp): OBJ, or PLY ascii /
 * binary_little_endian; polygons fan-triangulated.  kind = "obj", "ply" or
 * NULL (by file extension).  *verts (3 nv floats) and *tris (3 nt) are
 * allocated by the library: release them with n2wos_mesh_free. */
int n2wos_mesh_load(const char* path, const char* kind, float** verts, int* n_verts, uint32_t** tris, int* n_tris);
int n2wos_mesh_free(float* verts, uint32_t* tris);
/* Host brute-force closest point (closest_point_bruteforce,
 * src/geometry/closest_point.cpp:50-68): every triangle, first strict
 * minimum of d2, the same FP32 arithmetic as the GPU query (bit-identical
 * results); nthreads <= 0 uses every hardware thread. */
int n2wos_closest_point_bruteforce(const float* verts, int nv, const uint32_t* tris, int nt, const float* points,
                                   int n, float* d2, float* closest, int* tri, int nthreads);

/* -------------------------------------------------------------------- BVH */
/* Builds the BVH on the host and uploads it to the current device.
 * build_method: NULL, "median" or "custom" (split at the centroid median of
 * the largest extent), or "sah" / "surface_area_heuristic" (binned surface
 * area heuristic); leaves of at most leaf_size triangles, leaf_size in
 * [1, 31]; others -> N2WOS_ERR_UNSUPPORTED. */
int n2wos_bvh_create(const float* verts, int nv, const uint32_t* tris, int nt, int leaf_size,
                     const char* build_method, n2wos_bvh** out);
int n2wos_bvh_destroy(n2wos_bvh* bvh);

typedef struct {
  int64_t triangle_count;
  int64_t node_count;      /* 4-wide nodes */
  int64_t index_count;     /* triangle records (== triangle_count) */
  int64_t binary_nodes;    /* nodes of the median-split tree before collapsing */
  int leaf_size;
  int max_depth;           /* depth of the median-split tree */
  int depth4;              /* depth of the 4-wide tree */
  float prune_margin;      /* absolute distance margin used for exact pruning */
  double build_milliseconds;
  int64_t device_bytes;
  float bounds_lo[3];
  float bounds_hi[3];
  char build_method[32];
} n2wos_bvh_stats;
int n2wos_bvh_get_stats(const n2wos_bvh* bvh, n2wos_bvh_stats* out);

/* Closest point on the mesh for each of query_count points.
 * Outputs: squared distance, closest point (3 floats), original triangle
 * index (lowest index among exact ties), overflow flag (0 unless the
 * traversal stack overflowed and the result may be inexact).  A point with a
 * non-finite coordinate yields d2 = FLT_MAX, closest = (0,0,0), tri = -1.
 * query_count == 0 is a no-op; < 0 -> N2WOS_ERR_INVALID.
 * A BVH handle is read-only after n2wos_bvh_create: queries and solves that
 * use it may run concurrently on different streams and host threads (each
 * with its own output buffers and workspace); n2wos_bvh_destroy must not
 * overlap any of them. */
int n2wos_bvh_query_device(const n2wos_bvh* bvh, const float* d_points, int query_count, float* d_distance2,
                           float* d_closest, int* d_triangle_id, int* d_overflow, int block_size,
                           void* stream);
/* Host-memory convenience form of n2wos_bvh_query_device (synchronous):
 * copies the points to the device, runs the query on the null stream and
 * copies the results back; *overflow_count receives the number of overflowed
 * queries and *kernel_milliseconds (optional) the query launch's duration
 * measured with HIP events.  Allocates and frees its device buffers: for
 * probes and tests, not for a loop (use n2wos_bvh_query_device there). */
int n2wos_bvh_query_host(const n2wos_bvh* bvh, const float* points, int query_count, float* distance2,
                         float* closest, int* triangle_id, int* overflow_count, float* kernel_milliseconds,
                         int block_size);
/* As above; where d_active[i] == 0 the outputs are d2 = 0, closest = the
 * point itself, tri = -1, overflow = 0 and no traversal runs. */
int n2wos_bvh_query_device_masked(const n2wos_bvh* bvh, const float* d_points, const uint8_t* d_active,
                                  int query_count, float* d_distance2, float* d_closest, int* d_triangle_id,
                                  int* d_overflow, int block_size, void* stream);
/* Query with per-query visit counters: d_node_visits[i] / d_tri_visits[i]
 * receive the number of node lines loaded and triangle records tested (the
 * inputs of the traffic model).  Same results as n2wos_bvh_query_device. */
int n2wos_bvh_query_device_counted(const n2wos_bvh* bvh, const float* d_points, int query_count,
                                   float* d_distance2, float* d_closest, int* d_triangle_id, int* d_overflow,
                                   uint32_t* d_node_visits, uint32_t* d_tri_visits, int block_size,
                                   void* stream);
/* Host-only build (no device needed): the node and triangle-record arrays
 * exactly as n2wos_bvh_create would upload them.  Call with nodes_out ==
 * NULL to get *node_count; tris_out holds nt records of 48 bytes.
 * *prune_margin receives the absolute pruning margin. */
int n2wos_bvh_build_host(const float* verts, int nv, const uint32_t* tris, int nt, int leaf_size,
                         const char* build_method, void* nodes_out, int64_t* node_count, void* tris_out,
                         float* prune_margin);
/* Host copies of the device arrays: nodes (128 B each, node_count of them)
 * and triangle records (48 B each).  Either pointer may be NULL. */
int n2wos_bvh_export(const n2wos_bvh* bvh, void* nodes_out, void* tris_out);

/* Inside/outside classification of points against the (closed) mesh: ray
 * parity along three fixed generic directions, majority vote.  d_inside[i]
 * = 1 inside, 0 outside. */
int n2wos_bvh_inside_device(const n2wos_bvh* bvh, const float* d_points, int n, uint8_t* d_inside, void* stream);

/* -------------------------------------------------------- Walk-on-Spheres */
/* Dirichlet boundary data g.  Modes 0, 3, 4, 5, 12 are the reference's
 * NcBoundaryMode ids (include/n2wos/tcnn_nc_wos.hpp:15-36). */
enum {
  N2WOS_BC_X2_MINUS_Y2 = 0,
  N2WOS_BC_ZEBRA_K4 = 3,
  N2WOS_BC_ZEBRA_K8 = 4,
  N2WOS_BC_ZEBRA_K12 = 5,
  N2WOS_BC_CONSTANT_ONE = 12,
  N2WOS_BC_CHARGES = 100,       /* g = sum_i q_i / |x - c_i| */
  N2WOS_BC_POISSON_SINES = 101  /* g = u* = sin(2pi x) sin(2pi y) sin(2pi z) + |x|^2/2 */
};
typedef struct {
  int mode;
  int n_charges;          /* <= 16, for N2WOS_BC_CHARGES */
  float charges[16][4];   /* cx, cy, cz, q */
} n2wos_boundary;

enum {
  N2WOS_SOURCE_NONE = 0,         /* Laplace */
  N2WOS_SOURCE_POISSON_SINES = 1 /* -Laplace u = f = 12 pi^2 sss - 3, Green's-function sampling */
};
typedef struct {
  int max_steps;      /* reference default 256 */
  float epsilon;      /* stop when d <= epsilon; reference default 1e-4 */
  float step_scale;   /* R = step_scale * d; reference default 0.999 */
  int source;         /* N2WOS_SOURCE_* */
  uint64_t seed;      /* Philox key */
  uint32_t rng_stream;/* Philox counter word 3 = 2 rng_stream (+1) */
  int max_blocks;     /* 0 = fill the device */
  /* Optional instrumentation (NULL = off): two uint64 device counters that
   * the walker kernel adds its node-line loads and triangle tests to (a
   * counting build of the same kernel, same results), and two hipEvent_t
   * recorded on the stream right before and after the walker kernel. */
  void* d_visit_counts;
  void* event_begin;
  void* event_end;
} n2wos_walk_params;

/* Bytes of device workspace n2wos_wos_solve needs for this job on the current
 * device (a negative error code if the device cannot be queried): per-unit
 * and per-point state plus a staging area for walk values and a 128-byte
 * line per walker lane for traversal-stack spills, whose size follows the
 * device's resident walker blocks (at most 52.4 MB on MI355X). */
int64_t n2wos_wos_workspace_size(int n_points, int walks_per_point);
/* Blocks of the persistent walker's full-device grid on the current device
 * for this source term (the launch that max_blocks = 0 asks for), so that a
 * caller sharing the device can set max_blocks to a fraction of it.  No
 * counterpart in the reference (its walkers launch one thread per walk). */
int n2wos_wos_full_grid_blocks(int source, int* blocks);
/* walks_per_point walks from each point; per-point FP64 (n, mean, M2),
 * total steps and forced-at-max-steps counts.  With accumulate != 0 the new
 * statistics are merged (Chan) into the existing contents of d_n, d_mean,
 * d_m2, d_steps, d_forced; otherwise they are overwritten.  Point i uses RNG
 * point id point_offset + i, so a job split into point ranges reproduces the
 * unsplit job bit for bit.  d_walk_values / d_walk_steps (both NULL or both
 * set, n_points * walks_per_point entries) receive each walk's value and step
 * count.  No allocation: the caller passes a workspace of at least
 * n2wos_wos_workspace_size bytes. */
int n2wos_wos_solve(const n2wos_bvh* bvh, const n2wos_boundary* boundary, const n2wos_walk_params* params,
                    const float* d_points, int n_points, int walks_per_point, uint32_t point_offset, double* d_n,
                    double* d_mean, double* d_m2, uint64_t* d_steps, uint32_t* d_forced, int accumulate,
                    void* d_workspace, int64_t workspace_bytes, float* d_walk_values, int* d_walk_steps,
                    void* stream);

/* Largest walker workspace (n2wos_wos_workspace_size, bytes) that
 * n2wos_wos_solve_noalloc serves: 96 MiB.  For example 170,132 points x 256
 * walks need 74.5 MB on MI355X; 400,000 points x 256 walks do not fit. */
#define N2WOS_NOALLOC_SCRATCH_BYTES (96LL << 20)
/* n2wos_wos_solve without the workspace arguments, for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  No counterpart in the reference.  The
 * walker's workspace is a file-scope device array of
 * N2WOS_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
 * the process ends; a job whose workspace would exceed it returns
 * N2WOS_ERR_INVALID before anything is enqueued.  The call allocates nothing
 * and gives n2wos_wos_solve's results bit for bit.  Every call on one GPU
 * shares that GPU's array (it keeps the last call's contents): calls that
 * overlap on one GPU, from two streams or two threads, race, so the caller
 * serialises them. */
int n2wos_wos_solve_noalloc(const n2wos_bvh* bvh, const n2wos_boundary* boundary, const n2wos_walk_params* params,
                            const float* d_points, int n_points, int walks_per_point, uint32_t point_offset,
                            double* d_n, double* d_mean, double* d_m2, uint64_t* d_steps, uint32_t* d_forced,
                            int accumulate, float* d_walk_values, int* d_walk_steps, void* stream);

/* ------------------------------------------------ along-walk control variate */
/* A field C for the control variate: values at the (res + 1)^3 vertices of
 * a box grid, vertex (x, y, z) at d_values[x + (res+1) (y + (res+1) z)],
 * vertex position origin + (x, y, z) / cell_inverse; read by trilinear
 * interpolation, coordinates clamped to the box.  points: 2 (antipodal pair)
 * or 6 (octahedron turned by a uniform angle about the step direction);
 * beta scales the correction (1 = full). */
typedef struct {
  const float* d_values;
  int res;
  int points;
  float origin[3];
  float cell_inverse[3];
  float beta;
} n2wos_cv_field;

/* Plain walks with the along-walk control variate: each walk's value is
 *   W - beta sum_k [C(X_k + R_k d_k) - (1/m) sum_j C(X_k + R_k s_j)],
 * where X_k + R_k d_k is the walk's next position and s_j the m design
 * directions around d_k, each uniform on the sphere: every term has zero
 * mean, so the estimate is unbiased for any field, and a field close to the
 * solution (the learned cache evaluated on the grid) cancels most of the
 * walk's variance.  field NULL or points 0: plain walks.  Statistics,
 * workspace, point_offset and per-walk outputs as n2wos_wos_solve.  With
 * path_len > 0 the walker also records, per walk (walk index i = point *
 * walks_per_point + walk), its first path_len positions X_k (d_path_pos[3
 * (i path_len + k)]) and q_k = S_k - beta (terms of steps < k)
 * (d_path_q[i path_len + k]), S_k the source sum before step k: entries k <=
 * steps of walk i are written, and d_walk_values[i] - q_k is an unbiased
 * estimate of the solution at X_k (training labels for the cache).  No
 * counterpart in the reference; its two-level estimator
 * (src/tools/eval_tcnn_nc_wos.cu:767-783) uses the cache at one depth only. */
int n2wos_cv_solve(const n2wos_bvh* bvh, const n2wos_boundary* boundary, const n2wos_walk_params* params,
                   const float* d_points, int n_points, int walks_per_point, uint32_t point_offset,
                   const n2wos_cv_field* field, double* d_n, double* d_mean, double* d_m2, uint64_t* d_steps,
                   uint32_t* d_forced, int accumulate, void* d_workspace, int64_t workspace_bytes,
                   float* d_walk_values, int* d_walk_steps, float* d_path_pos, float* d_path_q, int path_len,
                   void* stream);

/* ------------------------------------------------ neural-cache WoS and 2LMC */
/* Walks of the neural-cache estimators (replacing launch_nc_hybrid_prefix and
 * launch_nc_2lmc_prefix_continue, include/n2wos/tcnn_nc_wos.hpp:79-134):
 *   N2WOS_NC_PREFIX:  each walk stops at depth m (or earlier at the boundary);
 *     d_prefix = accumulated source (+ boundary value if it ended there),
 *     d_xm = position X_m, d_need = 1 if the cache must be added at X_m;
 *   N2WOS_NC_COUPLED: each walk records the same depth-m state and continues
 *     to the boundary; d_full = the full walk's value (shared prefix).
 * Outputs have n_points * walks_per_point entries (d_xm three floats each);
 * d_steps receives per-point step totals.  Workspace: n2wos_nc_workspace_size. */
/* N2WOS_NC_ANTITHETIC may be OR-ed into the mode: walks 2j and 2j+1 of a
 * point then take mirrored first steps (the odd walk uses its even partner's
 * step-0 draw with the direction, and the source-sample offset, negated).
 * Each walk keeps its marginal law, so every estimator stays as it was in
 * expectation; the pair mean cancels the odd part of the step-0 variation.
 * (No counterpart in the reference: a variance-reduction option.) */
enum { N2WOS_NC_PREFIX = 1, N2WOS_NC_COUPLED = 2, N2WOS_NC_ANTITHETIC = 0x100 };
int64_t n2wos_nc_workspace_size(int n_points, int walks_per_point);
int n2wos_nc_walks(const n2wos_bvh* bvh, const n2wos_boundary* boundary, const n2wos_walk_params* params,
                   const float* d_points, int n_points, int walks_per_point, uint32_t point_offset, int mode,
                   int depth_m, float* d_prefix, float* d_xm, uint8_t* d_need, float* d_full, uint64_t* d_steps,
                   void* d_workspace, int64_t workspace_bytes, void* stream);
/* Per-point FP64 statistics of the estimator samples (the reference's combine
 * step, src/cuda/cubql_bvh.cu:1608-1627): prefix v = a + need C, coupled
 * v = need (W - a - C).  d_cache holds C(X_m) per walk (entries with need = 0
 * are ignored).  accumulate != 0 merges into the existing statistics. */
int n2wos_nc_combine(int n_points, int walks_per_point, int mode, const float* d_prefix, const uint8_t* d_need,
                     const float* d_cache, const float* d_full, int accumulate, double* d_n, double* d_mean,
                     double* d_m2, void* stream);

/* ------------------------------------------------------------ learned cache */
/* Multiresolution hash grid + ReLU MLP (2 n_levels inputs, n_hidden_layers
 * hidden layers of n_neurons, 1 output), trained with Adam on the MSE loss.
 * Replaces the tiny-cuda-nn NetworkWithInputEncoding / Trainer / inference calls of
 * src/tools/eval_tcnn_nc_wos.cu:747-769 (HashGrid + FullyFusedMLP, L2 loss,
 * Adam).  MLP operands are FP16 on MFMA with FP32 accumulation and a static
 * loss scale; tables and master weights are FP32.  World points map to the
 * unit cube by (p - origin) / extent, clamped. */
typedef struct n2wos_cache n2wos_cache;
typedef struct {
  int n_levels;          /* <= 16 */
  int log2_table_size;   /* T = 2^log2_table_size entries per hashed level */
  int base_resolution;   /* N_min */
  float growth;          /* b: N_l = floor(N_min b^l) */
  float origin[3];
  float extent[3];
  float learning_rate;   /* Adam */
  float beta1, beta2, epsilon;
  float loss_scale;      /* FP16 gradient scale */
  uint64_t seed;         /* parameter initialisation */
  int max_batch;         /* largest training batch (0: 262144); its activation
                            store (2 (33 + 2 L H) B per sample) is allocated here */
  int n_neurons;         /* hidden width H: 16, 32 or 64 (0: 64) */
  int n_hidden_layers;   /* hidden layers L: 1 .. 4 (0: 4) */
} n2wos_cache_config;
/* MLP weights, FP32 row-major, concatenated: A_1 [H x 32] (input features
 * 2l, 2l+1 of level l; levels >= n_levels read zero), A_2 .. A_L [H x H],
 * A_out [16 x H] (row 0 is the output, rows 1..15 are zero):
 * 32 H + (L - 1) H^2 + 16 H floats (15,360 for 64 x 4). */
int n2wos_cache_create(const n2wos_cache_config* config, n2wos_cache** out);
int n2wos_cache_destroy(n2wos_cache* cache);
/* A constant added to the network's output, in inference and in the
 * training loss (default 0).  Setting it to the mean training target lets
 * the bias-free MLP fit the variation only, which trains in far fewer steps.
 * No counterpart in the reference (tiny-cuda-nn's MLP is also bias-free;
 * the reference trains on raw targets). */
int n2wos_cache_set_output_offset(n2wos_cache* cache, float offset);
/* Queue a device-to-device copy of src's parameters (grid, FP32 and FP16
 * weights, output offset) into dst, which must have the same grid layout, on
 * a stream.  A frozen inference snapshot while src keeps training: the
 * two-cache snapshot discipline of docs/progressive_snapshot_accounting_policy.md
 * (the reference describes it; it has no such entry point). */
int n2wos_cache_copy_params(n2wos_cache* dst, const n2wos_cache* src, void* stream);
/* d_out[i] = C(d_points[i]) for n points (12 B each). */
int n2wos_cache_infer(const n2wos_cache* cache, const float* d_points, int n, float* d_out, void* stream);
/* One Adam step on the batch (n points, n targets).  *d_loss_out (device,
 * optional) receives the summed squared error.  No allocation: n above the
 * config's max_batch returns N2WOS_ERR_INVALID. */
int n2wos_cache_train_step(n2wos_cache* cache, const float* d_points, const float* d_targets, int n,
                           float* d_loss_out, void* stream);
/* Grid layout: per-level offsets, sizes (entries), resolutions and dense flags. */
int n2wos_cache_layout(const n2wos_cache* cache, int* n_levels, uint32_t* table_size, uint32_t* offsets,
                       uint32_t* sizes, int* res, int* dense, uint32_t* total_entries);
/* Host copies / uploads of the FP32 parameters (synchronous). */
int n2wos_cache_get_params(const n2wos_cache* cache, float* grid, float* weights);
int n2wos_cache_set_params(n2wos_cache* cache, const float* grid, const float* weights);
/* Gradients of the mean squared error on one batch, without an update
 * (synchronous; for gradient checks).  *loss = mean squared error. */
int n2wos_cache_gradients(n2wos_cache* cache, const float* d_points, const float* d_targets, int n, float* grid_grad,
                          float* weight_grad, float* loss);

/* ---------------------------------------------------------------- serving
 * A request server over resident meshes.  The reference has no serving loop;
 * these entry points have no reference counterpart and serve callers that
 * evaluate point sets of uneven size against meshes kept on the device.
 *
 * A request is a set of interior points of one mesh plus a target RMSE.  It
 * is admitted when its device state (48 B per point) fits the server's HBM
 * budget (first come, first served), then runs in rounds: round 0 gives every
 * point `first_round_walks` walks; each later round gives every point the
 * walks that its estimated RMSE, sqrt(mean_i var_i / n_i), says are still
 * missing (x 1.1, in units of 64, at most `max_round_walks`), until the
 * estimate is <= the target or `max_walks_per_point` is reached.  The
 * estimate and the next round's size are computed on the device and read
 * back asynchronously, so a round ends without a host synchronisation.
 *
 * Continuous batching: each mesh has two batch slots, each with its own
 * stream.  A slot's launch holds pieces of the current rounds of all the
 * mesh's active requests (max-min fair share of `tick_walks` walks; whole
 * points only): ONE persistent walker launch over a full-device grid, after
 * which each piece is merged into its request's per-point FP64
 * (n, mean, M2).  n2wos_server_step() admits queued requests, fills every
 * free slot, waits for the oldest launch, retires every completed one, ends
 * the rounds whose pieces are all merged (retiring finished requests) and
 * refills the freed slots before it returns, so launches of all meshes
 * overlap on the device and a new launch takes over the blocks that the
 * tail of the previous one frees.
 * Walk j of request point i uses the Philox counter (step, j, rng_base + i)
 * under the mesh's seed, and a point's 64-walk units are merged in walk
 * order, so a request's results are bit-identical whatever else is in the
 * batch, and equal to n2wos_wos_solve of the same points with point_offset =
 * rng_base and walks_per_point = the request's final walk count.
 *
 * A server and its meshes' BVH handles must be used from one thread at a
 * time; the BVH handles must outlive the server. */
typedef struct n2wos_server n2wos_server;
typedef struct {
  int64_t hbm_budget_bytes;   /* device bytes for request state (one arena) */
  int64_t tick_walks;         /* walks per batch slot launch (>= 64) */
  int first_round_walks;      /* multiple of 64 */
  int max_round_walks;        /* multiple of 64 */
  int max_walks_per_point;    /* multiple of 64 */
} n2wos_server_config;
enum { N2WOS_REQ_QUEUED = 0, N2WOS_REQ_ACTIVE = 1, N2WOS_REQ_DONE = 2 };
typedef struct {
  int state;                  /* N2WOS_REQ_* */
  int mesh;
  int n_points;
  int rounds;                 /* rounds completed */
  int walks_per_point;        /* walks merged per point so far */
  double estimated_rmse;      /* after the last completed round (-1 before) */
  uint64_t walk_steps;        /* total steps of the finished request (0 before) */
  double t_submit, t_admit, t_done;  /* seconds since server creation (-1 if not yet) */
  int control_variate;        /* 1: the mesh's control-variate field was used */
} n2wos_request_info;
typedef struct {
  int64_t ticks;
  int64_t requests_submitted, requests_done;
  int64_t walks;              /* walks run so far */
  uint64_t walk_steps;        /* steps of finished requests */
  int64_t arena_bytes, arena_used_bytes, fixed_bytes;
  double busy_seconds;        /* wall time inside n2wos_server_step */
  int64_t launches;           /* batch launches so far */
  double gpu_busy_seconds;    /* device time covered by at least one retired launch */
  double gpu_span_seconds;    /* device time from the first launch's start to the last's end */
} n2wos_server_stats;
int n2wos_server_create(const n2wos_server_config* config, n2wos_server** out);
int n2wos_server_destroy(n2wos_server* server);
/* Registers a resident mesh with its boundary data and walk parameters
 * (seed = Philox key of all its requests).  Allocates the mesh's batch
 * buffers (counted in fixed_bytes) and its stream. */
int n2wos_server_add_mesh(n2wos_server* server, const n2wos_bvh* bvh, const n2wos_boundary* boundary,
                          const n2wos_walk_params* params, int* mesh_id);
/* Attaches a control-variate field to a mesh (before any request to it):
 * its batches then run the along-walk control variate (n2wos_cv_solve) over
 * `field`, whose device values must outlive the server, in entries of
 * unit_walks walks (a power of two <= 64) with first rounds of
 * first_round_walks (a multiple of unit_walks).  field NULL or points 0
 * restores plain walks.  The caller's policy decides per mesh: attach the
 * field when it gains over plain walks (variance reduction / cost per walk
 * > 1), keep plain walks otherwise.  No counterpart in the reference. */
int n2wos_server_set_mesh_field(n2wos_server* server, int mesh_id, const n2wos_cv_field* field, int unit_walks,
                                int first_round_walks);
/* Queues a request (host points are copied; synchronous).  Fails with
 * N2WOS_ERR_NOMEM if the request can never fit the budget. */
int n2wos_server_submit(n2wos_server* server, int mesh_id, const float* host_points, int n_points,
                        double target_rmse, uint32_t rng_base, int64_t* request_id);
/* One scheduling step (synchronous).  *n_done = requests finished in it. */
int n2wos_server_step(n2wos_server* server, int* n_done);
int n2wos_server_request_info(const n2wos_server* server, int64_t request_id, n2wos_request_info* info);
/* Copies a finished request's per-point results to host arrays (any may be
 * NULL), frees its arena range and forgets the request.  Until then a
 * finished request keeps its range (results wait on the device). */
int n2wos_server_take_result(n2wos_server* server, int64_t request_id, double* n, double* mean, double* m2,
                             uint64_t* steps, uint32_t* forced);
int n2wos_server_get_stats(const n2wos_server* server, n2wos_server_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* N2WOS_AMD_H */
