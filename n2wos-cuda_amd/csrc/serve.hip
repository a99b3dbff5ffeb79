// serve.hip -- request server over resident meshes: HBM-budget admission,
// continuous batching of walker pools across requests, per-request FP64
// statistics that do not depend on what else shares the batch.
//
// Device state.  One arena (a single hipMalloc of the HBM budget) holds every
// admitted request's points and per-point (n, mean, M2, steps, forced); the
// host keeps a first-fit free list over it, so admission and retirement never
// call the HIP allocator.  Each mesh owns two batch slots, each with its own
// stream, event and buffers: capacity tick_walks / 64 entries of (point,
// Philox counter words, one 64-walk unit's statistics) plus the walker
// workspace, allocated once when the mesh is added.
//
// A slot launch: share the slot's capacity max-min fairly among the current
// rounds of the mesh's active requests (whole points; a point of a round with
// k units takes k entries), upload the piece table, gather the batch (entry ->
// request point, counter point id rng_base + i and walk base round_base +
// 64 c), one persistent walker launch over a full-device grid (64 walks per
// entry), and merge each point's k entries, in walk order, into the
// request's running statistics.  Slots of all meshes run concurrently; the
// blocks a draining launch frees go to the next one.
//
// A step: admit queued requests in arrival order while the arena has room,
// fill free slots, wait for the oldest launch, retire every completed one;
// for requests whose round is fully merged a fixed-order block reduction on
// the device gives the estimated RMSE and sizes the next round (0 = done),
// and the host reads the decisions back asynchronously (an event it polls),
// so no round ends in a host synchronisation; a finished request's results
// are copied to its pinned host buffer the same way (its arena range is
// freed when that copy's event has completed); freed slots are refilled at
// once.
//
// A mesh may carry a control-variate field (n2wos_server_set_mesh_field):
// its batches then run the extended walker with the along-walk control
// variate, and its rounds count in units of fewer walks (packed walker
// units), since each walk carries far less variance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <unordered_map>
#include <vector>

#include "api_internal.hpp"
#include "walker_kernel.hpp"
#include "wos.hpp"

namespace n2 {
namespace {

constexpr int kUnit = 64;          // walks per batch entry (one walker work unit)
constexpr int kMaxPieces = 4096;   // pieces per mesh per step (bounds active requests)

// One piece of a request's round inside a mesh batch.
struct ServePiece {
  const float* points;  // request points (12 B each)
  double* n;
  double* mean;
  double* m2;
  uint64_t* steps;
  uint32_t* forced;
  int first;            // first request point of the piece
  int count;            // points in the piece
  int units;            // 64-walk units per point in this round
  int batch_start;      // first batch entry (count * units entries)
  int point_start;      // prefix sum of counts over the pieces before this one
  uint32_t id_base;     // the request's counter base
  uint32_t walk_base;   // first walk index of the round
  int unit;             // walks per batch entry
};

__device__ __forceinline__ int find_piece(const ServePiece* pc, int n_pieces, int i, bool by_entry) {
  int lo = 0, hi = n_pieces - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int key = by_entry ? pc[mid].batch_start : pc[mid].point_start;
    if (key <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void k_serve_gather(const ServePiece* __restrict__ pc, int n_pieces, int n_entries, float* __restrict__ bpts,
                               uint32_t* __restrict__ brng) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_entries) return;
  const ServePiece& p = pc[find_piece(pc, n_pieces, i, true)];
  const int local = i - p.batch_start;
  const int j = p.first + local / p.units;
  const int c = local - (local / p.units) * p.units;
  bpts[3 * i] = p.points[3 * j];
  bpts[3 * i + 1] = p.points[3 * j + 1];
  bpts[3 * i + 2] = p.points[3 * j + 2];
  brng[2 * i] = p.id_base + static_cast<uint32_t>(j);
  brng[2 * i + 1] = p.walk_base + static_cast<uint32_t>(p.unit * c);
}

// Merges each piece point's entries, in walk order, into the request state:
// the same sequence of Chan merges as one walker launch over all its walks.
__global__ void k_serve_merge(const ServePiece* __restrict__ pc, int n_pieces, int n_points,
                              const double* __restrict__ bn, const double* __restrict__ bmean,
                              const double* __restrict__ bm2, const uint64_t* __restrict__ bsteps,
                              const uint32_t* __restrict__ bforced) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_points) return;
  const ServePiece& p = pc[find_piece(pc, n_pieces, t, false)];
  const int local = t - p.point_start;
  const int j = p.first + local;
  double N = p.n[j], M = p.mean[j], Q = p.m2[j];
  uint64_t steps = p.steps[j];
  uint32_t forced = p.forced[j];
  for (int c = 0; c < p.units; ++c) {
    const int e = p.batch_start + local * p.units + c;
    chan_merge(N, M, Q, bn[e], bmean[e], bm2[e], N, M, Q);
    steps += bsteps[e];
    forced += bforced[e];
  }
  p.n[j] = N;
  p.mean[j] = M;
  p.m2[j] = Q;
  p.steps[j] = steps;
  p.forced[j] = forced;
}

struct EstItem {
  const double* n;
  const double* m2;
  const uint64_t* steps;
  double target;
  int count;
  int wpp;        // walks per point merged, this round included
  int unit;       // round granularity (walks per batch entry)
  int max_round;  // largest next round: max_round_walks capped by the walks left
};

struct EstOut {
  double est;      // estimated RMSE: sqrt(mean over points of var_i / n_i)
  uint64_t steps;  // walk steps of the request so far
  int next;        // walks per point of the next round, 0 = finished
  int pad;
};

constexpr int kEstBlock = 256;

// Per request (one block): the estimated RMSE and the request's step total
// by fixed-order block reductions, and the next round's size: the walks the
// estimate says are still missing (variance ~ 1 / n), 10 % margin, rounded up
// to whole units and clamped to [unit, max_round]; 0 when the estimate meets
// the target or no walks are left.
__global__ void __launch_bounds__(kEstBlock) k_serve_estimate(const EstItem* __restrict__ items,
                                                              EstOut* __restrict__ out) {
  __shared__ double s[kEstBlock];
  __shared__ uint64_t st[kEstBlock];
  const EstItem it = items[blockIdx.x];
  double acc = 0.0;
  uint64_t steps = 0;
  for (int j = threadIdx.x; j < it.count; j += kEstBlock) {
    const double n = it.n[j];
    if (n > 1.0) acc += it.m2[j] / (n - 1.0) / n;
    steps += it.steps[j];
  }
  s[threadIdx.x] = acc;
  st[threadIdx.x] = steps;
  __syncthreads();
  for (int w = kEstBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s[threadIdx.x] += s[threadIdx.x + w];
      st[threadIdx.x] += st[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double est = sqrt(fmax(0.0, s[0] / static_cast<double>(it.count)));
    int next = 0;
    if (!(est <= it.target) && it.max_round > 0) {
      const double ratio = (est / it.target) * (est / it.target);
      const double want = ceil(static_cast<double>(it.wpp) * (ratio - 1.0) * 1.1);
      int64_t w = want < 4.0e9 ? static_cast<int64_t>(want) : int64_t{4000000000};
      w = (w + it.unit - 1) / it.unit * it.unit;
      w = w < it.unit ? it.unit : w;
      w = w > it.max_round ? it.max_round : w;
      next = static_cast<int>(w);
    }
    out[blockIdx.x] = EstOut{est, st[0], next, 0};
  }
}

__global__ void k_fill_zero(uint32_t* p, int64_t words) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < words) p[i] = 0u;
}

int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// First-fit allocator over [0, size) in 256-byte granules; host bookkeeping only.
class Arena {
 public:
  explicit Arena(int64_t size) : size_(size) {
    if (size > 0) free_[0] = size;
  }
  int64_t alloc(int64_t bytes) {
    bytes = align256(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const int64_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest > 0) free_[off + bytes] = rest;
        used_ += bytes;
        return off;
      }
    }
    return -1;
  }
  void release(int64_t off, int64_t bytes) {
    bytes = align256(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t size() const { return size_; }
  int64_t used() const { return used_; }

 private:
  int64_t size_;
  int64_t used_ = 0;
  std::map<int64_t, int64_t> free_;
};

// Device bytes of one request's state (each array 256-byte aligned).
int64_t request_bytes(int n) {
  const int64_t p = n;
  return align256(12 * p) + 4 * align256(8 * p) + align256(4 * p);
}

// One in-flight batch of a mesh: its own stream, buffers and completion event.
struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t begin = nullptr;  // timing: the launch's first and last commands
  hipEvent_t done = nullptr;
  bool busy = false;
  int64_t seq = 0;        // launch order (the oldest busy slot is waited on first)
  int entries = 0;
  std::vector<int64_t> reqs;  // requests with a piece in this batch
  void* buf = nullptr;    // one allocation holding everything below
  float* bpts = nullptr;
  uint32_t* brng = nullptr;
  double* bn = nullptr;
  double* bmean = nullptr;
  double* bm2 = nullptr;
  uint64_t* bsteps = nullptr;
  uint32_t* bforced = nullptr;
  void* ws = nullptr;
  ServePiece* d_pieces = nullptr;
  ServePiece* h_pieces = nullptr;  // pinned staging
};

constexpr int kSlotsPerMesh = 2;

struct Mesh {
  const n2wos_bvh* bvh = nullptr;
  Boundary bc{};
  n2wos_walk_params wp{};
  int cap = 0;            // batch entries per slot
  int unit = kUnit;       // walks per batch entry (a round is a multiple)
  int first_round = 0;    // walks per point of a request's first round
  bool cv = false;        // along-walk control variate over the field below
  n2wos_cv_field field{};
  Slot slot[kSlotsPerMesh];
  int64_t bytes = 0;
};

// Decisions of one batch of ended rounds, read back asynchronously.
struct EstBatch {
  EstItem* d_items = nullptr;
  EstItem* h_items = nullptr;  // pinned
  EstOut* d_out = nullptr;
  EstOut* h_out = nullptr;     // pinned
  hipEvent_t done = nullptr;
  std::vector<int64_t> ids;
  bool busy = false;
  int64_t seq = 0;
};

constexpr int kEstBatches = 4;

// Result copies of finished requests (device arena -> pinned host), one event.
struct ResBatch {
  hipEvent_t done = nullptr;
  std::vector<int64_t> ids;
  bool busy = false;
  int64_t seq = 0;
};

// Pinned host bytes of one request's results: n, mean, m2, steps (8 B each), forced.
int64_t result_bytes(int n) { return 36 * static_cast<int64_t>(n); }

struct Request {
  int64_t id = 0;
  int mesh = 0;
  int n = 0;
  double target = 0.0;
  uint32_t rng_base = 0;
  std::vector<float> host_points;
  int state = N2WOS_REQ_QUEUED;
  int64_t off = -1, bytes = 0;
  float* d_points = nullptr;
  double* d_n = nullptr;
  double* d_mean = nullptr;
  double* d_m2 = nullptr;
  uint64_t* d_steps = nullptr;
  uint32_t* d_forced = nullptr;
  int rounds = 0;
  int unit = kUnit;       // the mesh's walks per batch entry
  int wpp = 0;            // walks merged per point
  int round_walks = 0;    // walks per point of the current round
  int cursor = 0;         // points of the current round already batched
  int inflight = 0;       // busy slots holding a piece of this request
  double est = -1.0;
  uint64_t total_steps = 0;
  double t_submit = -1.0, t_admit = -1.0, t_done = -1.0;
  char* h_res = nullptr;  // pinned results (result_bytes), filled when the request finishes
};

}  // namespace
}  // namespace n2

struct n2wos_server {
  n2wos_server_config cfg{};
  int device = 0;
  char* arena_base = nullptr;
  std::unique_ptr<n2::Arena> arena;
  std::vector<std::unique_ptr<n2::Mesh>> meshes;
  std::map<int64_t, std::unique_ptr<n2::Request>> requests;  // by id (arrival order)
  std::deque<int64_t> queue;
  std::vector<int64_t> active;  // admission order
  int64_t next_id = 1;
  int64_t ticks = 0, submitted = 0, done = 0, walks = 0, launches = 0;
  uint64_t steps = 0;
  int64_t fixed_bytes = 0;
  double busy = 0.0;
  hipStream_t control = nullptr;  // admission uploads, estimates, result downloads
  n2::EstBatch est[n2::kEstBatches];
  std::vector<n2::ResBatch> res;
  int64_t est_seq = 0;
  hipEvent_t t_ref = nullptr;     // device-time origin of the launch intervals
  std::vector<std::pair<float, float>> intervals;  // device ms of each retired launch
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

using namespace n2;

namespace {

void server_admit(n2wos_server* s) {
  while (!s->queue.empty()) {
    Request& r = *s->requests.at(s->queue.front());
    if (static_cast<int>(s->active.size()) >= kMaxPieces) break;
    const int64_t off = s->arena->alloc(r.bytes);
    if (off < 0) break;  // first come, first served: later requests wait too
    char* base = s->arena_base + off;
    const int64_t p = r.n;
    r.off = off;
    r.d_points = reinterpret_cast<float*>(base);
    base += align256(12 * p);
    r.d_n = reinterpret_cast<double*>(base);
    base += align256(8 * p);
    r.d_mean = reinterpret_cast<double*>(base);
    base += align256(8 * p);
    r.d_m2 = reinterpret_cast<double*>(base);
    base += align256(8 * p);
    r.d_steps = reinterpret_cast<uint64_t*>(base);
    base += align256(8 * p);
    r.d_forced = reinterpret_cast<uint32_t*>(base);
    N2_HIP(hipMemcpyAsync(r.d_points, r.host_points.data(), 12 * p, hipMemcpyHostToDevice, s->control));
    const int64_t stat_words = (r.bytes - align256(12 * p)) / 4;
    hipLaunchKernelGGL(k_fill_zero, dim3(static_cast<unsigned>((stat_words + 255) / 256)), dim3(256), 0, s->control,
                       reinterpret_cast<uint32_t*>(r.d_n), stat_words);
    N2_HIP(hipGetLastError());
    N2_HIP(hipStreamSynchronize(s->control));  // the host copy of the points can go
    std::vector<float>().swap(r.host_points);
    r.state = N2WOS_REQ_ACTIVE;
    r.round_walks = s->meshes[static_cast<size_t>(r.mesh)]->first_round;
    r.cursor = 0;
    r.t_admit = s->now();
    s->active.push_back(r.id);
    s->queue.pop_front();
  }
}

// Max-min fair split of `cap` entries among needs given in whole points of
// `units` entries each; returns points granted per request.
std::vector<int> fair_share(const std::vector<int>& pts_left, const std::vector<int>& units, int cap) {
  const int k = static_cast<int>(pts_left.size());
  std::vector<int> grant(k, 0), order(k);
  for (int i = 0; i < k; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    const int64_t na = static_cast<int64_t>(pts_left[a]) * units[a], nb = static_cast<int64_t>(pts_left[b]) * units[b];
    return na != nb ? na < nb : a < b;
  });
  int64_t remaining = cap;
  for (int idx = 0; idx < k; ++idx) {
    const int i = order[idx];
    const int64_t share = remaining / (k - idx);
    int64_t g = std::min<int64_t>(pts_left[i], share / units[i]);
    if (g == 0 && pts_left[i] > 0 && remaining >= units[i]) g = 1;  // progress for wide rounds
    grant[i] = static_cast<int>(g);
    remaining -= g * units[i];
  }
  return grant;
}

// Plans and launches one batch of mesh `mi` into its free slot `sl`; false if
// no active request of the mesh has unscheduled points in its current round.
bool launch_slot(n2wos_server* s, int mi, Slot& sl) {
  Mesh& m = *s->meshes[mi];
  std::vector<int64_t> ids;
  for (int64_t id : s->active) {
    const Request& r = *s->requests.at(id);
    if (r.mesh == mi && r.cursor < r.n) ids.push_back(id);
  }
  if (ids.empty()) return false;
  std::vector<int> left(ids.size()), units(ids.size());
  for (size_t q = 0; q < ids.size(); ++q) {
    const Request& r = *s->requests.at(ids[q]);
    left[q] = r.n - r.cursor;
    units[q] = r.round_walks / r.unit;
  }
  const std::vector<int> grant = fair_share(left, units, m.cap);
  int entries = 0, points = 0, np = 0;
  sl.reqs.clear();
  for (size_t q = 0; q < ids.size(); ++q) {
    if (grant[q] == 0) continue;
    Request& r = *s->requests.at(ids[q]);
    ServePiece& pc = sl.h_pieces[np++];
    pc.points = r.d_points;
    pc.n = r.d_n;
    pc.mean = r.d_mean;
    pc.m2 = r.d_m2;
    pc.steps = r.d_steps;
    pc.forced = r.d_forced;
    pc.first = r.cursor;
    pc.count = grant[q];
    pc.units = units[q];
    pc.batch_start = entries;
    pc.point_start = points;
    pc.id_base = r.rng_base;
    pc.walk_base = static_cast<uint32_t>(r.wpp);
    pc.unit = r.unit;
    entries += grant[q] * units[q];
    points += grant[q];
    r.cursor += grant[q];
    r.inflight += 1;
    sl.reqs.push_back(r.id);
  }
  if (entries == 0) return false;
  hipStream_t st = sl.stream;
  N2_HIP(hipEventRecord(sl.begin, st));
  N2_HIP(hipMemcpyAsync(sl.d_pieces, sl.h_pieces, sizeof(ServePiece) * np, hipMemcpyHostToDevice, st));
  const int b = 256;
  hipLaunchKernelGGL(k_serve_gather, dim3((entries + b - 1) / b), dim3(b), 0, st, sl.d_pieces, np, entries, sl.bpts,
                     sl.brng);
  N2_HIP(hipGetLastError());
  WosRun w{};
  w.d_points = sl.bpts;
  w.n_points = entries;
  w.walks_per_point = m.unit;
  w.point_offset = 0;
  w.d_point_rng = sl.brng;
  w.max_steps = m.wp.max_steps;
  w.epsilon = m.wp.epsilon;
  w.step_scale = m.wp.step_scale;
  w.seed = m.wp.seed;
  w.rng_stream = m.wp.rng_stream;
  w.source = m.wp.source;
  w.max_blocks = 0;  // a full-device grid: a later launch takes over the blocks this one frees
  w.d_n = sl.bn;
  w.d_mean = sl.bmean;
  w.d_m2 = sl.bm2;
  w.d_steps = sl.bsteps;
  w.d_forced = sl.bforced;
  w.accumulate = 0;
  w.d_workspace = sl.ws;
  w.mode = kModePlain;
  if (m.cv) {
    // the extended walker: control variate over the mesh's field (packed
    // units when the entry holds fewer than 64 walks)
    w.ext = 1;
    w.cv_values = m.field.d_values;
    w.cv_res = m.field.res;
    w.cv_points = m.field.points;
    for (int k = 0; k < 3; ++k) {
      w.cv_origin[k] = m.field.origin[k];
      w.cv_inv_h[k] = m.field.cell_inverse[k];
    }
    w.cv_beta = m.field.beta;
  }
  run_wos(m.bvh, m.bc, w, st);
  hipLaunchKernelGGL(k_serve_merge, dim3((points + b - 1) / b), dim3(b), 0, st, sl.d_pieces, np, points, sl.bn,
                     sl.bmean, sl.bm2, sl.bsteps, sl.bforced);
  N2_HIP(hipGetLastError());
  N2_HIP(hipEventRecord(sl.done, st));
  sl.busy = true;
  sl.seq = ++s->launches;
  sl.entries = entries;
  s->walks += static_cast<int64_t>(entries) * m.unit;
  return true;
}

void launch_free_slots(n2wos_server* s) {
  for (int mi = 0; mi < static_cast<int>(s->meshes.size()); ++mi)
    for (Slot& sl : s->meshes[mi]->slot)
      if (!sl.busy && !launch_slot(s, mi, sl)) break;
}

// Retires a completed slot: requests whose round is fully merged go to `round_done`.
void retire(n2wos_server* s, Slot& sl, std::vector<int64_t>& round_done) {
  float a = 0.0f, b = 0.0f;
  if (hipEventElapsedTime(&a, s->t_ref, sl.begin) == hipSuccess &&
      hipEventElapsedTime(&b, s->t_ref, sl.done) == hipSuccess)
    s->intervals.emplace_back(a, b);
  else
    (void)hipGetLastError();
  for (int64_t id : sl.reqs) {
    Request& r = *s->requests.at(id);
    r.inflight -= 1;
    if (r.inflight == 0 && r.cursor == r.n) round_done.push_back(id);
  }
  sl.reqs.clear();
  sl.busy = false;
}

// Applies the decisions of a completed estimate batch: finished requests
// (their results stay in the arena) or the size of their next round.
void apply_estimates(n2wos_server* s, EstBatch& eb) {
  ResBatch* rb = nullptr;
  for (size_t q = 0; q < eb.ids.size(); ++q) {
    Request& r = *s->requests.at(eb.ids[q]);
    const EstOut& o = eb.h_out[q];
    r.est = o.est;
    if (o.next == 0) {
      // finished: copy the results to the request's pinned buffer
      r.total_steps = o.steps;
      if (!rb) {
        for (ResBatch& x : s->res)
          if (!x.busy) {
            rb = &x;
            break;
          }
        if (!rb) {
          s->res.emplace_back();
          rb = &s->res.back();
          N2_HIP(hipEventCreateWithFlags(&rb->done, hipEventDisableTiming));
        }
      }
      const int64_t p = r.n;
      char* h = r.h_res;
      hipStream_t st = s->control;
      N2_HIP(hipMemcpyAsync(h, r.d_n, 8 * p, hipMemcpyDeviceToHost, st));
      N2_HIP(hipMemcpyAsync(h + 8 * p, r.d_mean, 8 * p, hipMemcpyDeviceToHost, st));
      N2_HIP(hipMemcpyAsync(h + 16 * p, r.d_m2, 8 * p, hipMemcpyDeviceToHost, st));
      N2_HIP(hipMemcpyAsync(h + 24 * p, r.d_steps, 8 * p, hipMemcpyDeviceToHost, st));
      N2_HIP(hipMemcpyAsync(h + 32 * p, r.d_forced, 4 * p, hipMemcpyDeviceToHost, st));
      rb->ids.push_back(r.id);
    } else {
      r.round_walks = o.next;
      r.cursor = 0;
    }
  }
  if (rb) {
    N2_HIP(hipEventRecord(rb->done, s->control));
    rb->busy = true;
    rb->seq = ++s->est_seq;
  }
  eb.ids.clear();
  eb.busy = false;
}

// Requests whose result copy has completed: free their arena range, done.
int poll_results(n2wos_server* s, bool wait_oldest) {
  int finished = 0;
  ResBatch* oldest = nullptr;
  for (ResBatch& rb : s->res)
    if (rb.busy && (!oldest || rb.seq < oldest->seq)) oldest = &rb;
  if (wait_oldest && oldest) N2_HIP(hipEventSynchronize(oldest->done));
  for (ResBatch& rb : s->res) {
    if (!rb.busy) continue;
    const hipError_t q = hipEventQuery(rb.done);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) N2_HIP(q);
    for (int64_t id : rb.ids) {
      Request& r = *s->requests.at(id);
      s->arena->release(r.off, r.bytes);
      r.off = -1;
      r.d_points = nullptr;
      s->steps += r.total_steps;
      r.state = N2WOS_REQ_DONE;
      r.t_done = s->now();
      s->active.erase(std::find(s->active.begin(), s->active.end(), r.id));
      ++s->done;
      ++finished;
    }
    rb.ids.clear();
    rb.busy = false;
  }
  return finished;
}

EstBatch* oldest_est(n2wos_server* s) {
  EstBatch* o = nullptr;
  for (EstBatch& eb : s->est)
    if (eb.busy && (!o || eb.seq < o->seq)) o = &eb;
  return o;
}

// Queues the estimate of every request whose round has just been merged: a
// kernel on the control stream, then an asynchronous copy of its decisions
// to pinned memory and an event the host polls.  The requests schedule no
// work until their decision is applied.
int finish_rounds(n2wos_server* s, const std::vector<int64_t>& round_done) {
  const int finished = 0;
  size_t at = 0;
  while (at < round_done.size()) {
    EstBatch* eb = nullptr;
    for (EstBatch& e : s->est)
      if (!e.busy) {
        eb = &e;
        break;
      }
    if (!eb) {
      // every batch in flight: wait for the oldest and apply it
      EstBatch* o = oldest_est(s);
      N2_HIP(hipEventSynchronize(o->done));
      apply_estimates(s, *o);
      continue;
    }
    const int k = static_cast<int>(std::min<size_t>(round_done.size() - at, kMaxPieces));
    eb->ids.assign(round_done.begin() + static_cast<int64_t>(at), round_done.begin() + static_cast<int64_t>(at + k));
    at += static_cast<size_t>(k);
    for (int q = 0; q < k; ++q) {
      Request& r = *s->requests.at(eb->ids[static_cast<size_t>(q)]);
      r.rounds += 1;
      r.wpp += r.round_walks;
      const int room = s->cfg.max_walks_per_point - r.wpp;
      eb->h_items[q] = EstItem{r.d_n, r.d_m2, r.d_steps, r.target, r.n, r.wpp, r.unit,
                               room <= 0 ? 0 : std::min(s->cfg.max_round_walks, room / r.unit * r.unit)};
    }
    hipStream_t st = s->control;
    N2_HIP(hipMemcpyAsync(eb->d_items, eb->h_items, sizeof(EstItem) * k, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_serve_estimate, dim3(k), dim3(kEstBlock), 0, st, eb->d_items, eb->d_out);
    N2_HIP(hipGetLastError());
    N2_HIP(hipMemcpyAsync(eb->h_out, eb->d_out, sizeof(EstOut) * k, hipMemcpyDeviceToHost, st));
    N2_HIP(hipEventRecord(eb->done, st));
    eb->busy = true;
    eb->seq = ++s->est_seq;
  }
  return finished;
}

// Applies every estimate batch whose decisions have arrived.
void poll_estimates(n2wos_server* s) {
  for (EstBatch& eb : s->est) {
    if (!eb.busy) continue;
    const hipError_t q = hipEventQuery(eb.done);
    if (q == hipSuccess) {
      apply_estimates(s, eb);
    } else if (q != hipErrorNotReady) {
      N2_HIP(q);
    }
  }
}

// One step: admit, keep every free slot busy, wait for the oldest launch (or,
// with none in flight, the oldest estimate batch), retire every completed
// launch, queue the estimates of the rounds that ended, apply the decisions
// that have arrived, and refill the freed slots before returning.
void server_step(n2wos_server* s, int* n_done) {
  const double t_begin = s->now();
  server_admit(s);
  launch_free_slots(s);
  Slot* oldest = nullptr;
  for (auto& m : s->meshes)
    for (Slot& sl : m->slot)
      if (sl.busy && (!oldest || sl.seq < oldest->seq)) oldest = &sl;
  int finished = 0;
  EstBatch* eo = oldest ? nullptr : oldest_est(s);
  bool res_pending = false;
  for (const ResBatch& rb : s->res) res_pending = res_pending || rb.busy;
  if (oldest || eo || res_pending) {
    // wait for the oldest launch, else the oldest estimate, else a result copy
    if (oldest || eo) N2_HIP(hipEventSynchronize(oldest ? oldest->done : eo->done));
    std::vector<int64_t> round_done;
    for (auto& m : s->meshes)
      for (Slot& sl : m->slot) {
        if (!sl.busy) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipSuccess) {
          retire(s, sl, round_done);
        } else if (q != hipErrorNotReady) {
          N2_HIP(q);
        }
      }
    finished += finish_rounds(s, round_done);
    poll_estimates(s);
    finished += poll_results(s, !oldest && !eo);
    server_admit(s);
    launch_free_slots(s);
  }
  ++s->ticks;
  s->busy += s->now() - t_begin;
  if (n_done) *n_done = finished;
}

// Device time covered by at least one launch, and from the first launch's
// start to the last one's end, over the retired launches (seconds).
void device_busy(const n2wos_server* s, double* busy, double* span) {
  std::vector<std::pair<float, float>> iv = s->intervals;
  std::sort(iv.begin(), iv.end());
  double total = 0.0, cur_a = 0.0, cur_b = -1.0;
  for (const auto& x : iv) {
    if (x.first > cur_b) {
      if (cur_b > cur_a) total += cur_b - cur_a;
      cur_a = x.first;
      cur_b = x.second;
    } else if (x.second > cur_b) {
      cur_b = x.second;
    }
  }
  if (cur_b > cur_a) total += cur_b - cur_a;
  *busy = total * 1e-3;
  *span = iv.empty() ? 0.0 : (std::max_element(iv.begin(), iv.end(), [](const auto& a, const auto& b) {
                                return a.second < b.second;
                              })->second - iv.front().first) * 1e-3;
}

void free_slot(Slot& sl) {
  if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  (void)hipFree(sl.buf);
  (void)hipFree(sl.d_pieces);
  (void)hipHostFree(sl.h_pieces);
  if (sl.done) (void)hipEventDestroy(sl.done);
  if (sl.begin) (void)hipEventDestroy(sl.begin);
  if (sl.stream) (void)hipStreamDestroy(sl.stream);
  sl = Slot{};
}

// Allocates the mesh's batch slots for m.cap entries of m.unit walks each.
void alloc_slots(n2wos_server* s, Mesh& m) {
  const int64_t c = m.cap;
  const int64_t ws_bytes = wos_workspace_bytes(m.cap, m.unit);
  const int64_t sz[8] = {align256(12 * c), align256(8 * c), align256(8 * c), align256(8 * c),
                         align256(8 * c), align256(8 * c), align256(4 * c), align256(ws_bytes)};
  int64_t total = 0;
  for (int64_t v : sz) total += v;
  try {
    for (Slot& sl : m.slot) {
      if (hipMalloc(&sl.buf, total) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(N2WOS_ERR_NOMEM, "cannot allocate the mesh batch");
      }
      char* p = static_cast<char*>(sl.buf);
      sl.bpts = reinterpret_cast<float*>(p);
      p += sz[0];
      sl.brng = reinterpret_cast<uint32_t*>(p);
      p += sz[1];
      sl.bn = reinterpret_cast<double*>(p);
      p += sz[2];
      sl.bmean = reinterpret_cast<double*>(p);
      p += sz[3];
      sl.bm2 = reinterpret_cast<double*>(p);
      p += sz[4];
      sl.bsteps = reinterpret_cast<uint64_t*>(p);
      p += sz[5];
      sl.bforced = reinterpret_cast<uint32_t*>(p);
      p += sz[6];
      sl.ws = p;
      N2_HIP(hipMalloc(&sl.d_pieces, sizeof(ServePiece) * kMaxPieces));
      N2_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_pieces), sizeof(ServePiece) * kMaxPieces));
      N2_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
      N2_HIP(hipEventCreate(&sl.begin));
      N2_HIP(hipEventCreate(&sl.done));
    }
  } catch (...) {
    for (Slot& sl : m.slot) free_slot(sl);
    throw;
  }
  s->fixed_bytes -= m.bytes;
  m.bytes = kSlotsPerMesh * (total + static_cast<int64_t>(sizeof(ServePiece)) * kMaxPieces);
  s->fixed_bytes += m.bytes;
}

}  // namespace

extern "C" int n2wos_server_create(const n2wos_server_config* cfg, n2wos_server** out) {
  return guarded([&]() -> int {
    N2_REQUIRE(cfg && out, N2WOS_ERR_NULL, "NULL config or output");
    *out = nullptr;
    N2_REQUIRE(cfg->hbm_budget_bytes > 0, N2WOS_ERR_INVALID, "hbm_budget_bytes must be > 0");
    N2_REQUIRE(cfg->tick_walks >= kUnit && cfg->tick_walks / kUnit < (int64_t{1} << 30), N2WOS_ERR_INVALID,
               "tick_walks must be in [64, 2^36)");
    const auto mult = [](int v) { return v >= kUnit && v % kUnit == 0; };
    N2_REQUIRE(mult(cfg->first_round_walks) && mult(cfg->max_round_walks) && mult(cfg->max_walks_per_point),
               N2WOS_ERR_INVALID, "walk counts must be positive multiples of 64");
    N2_REQUIRE(cfg->max_round_walks <= cfg->tick_walks && cfg->first_round_walks <= cfg->tick_walks,
               N2WOS_ERR_INVALID, "a round of one point must fit one step (round walks <= tick_walks)");
    N2_REQUIRE(cfg->first_round_walks <= cfg->max_walks_per_point, N2WOS_ERR_INVALID,
               "first_round_walks exceeds max_walks_per_point");
    auto s = std::make_unique<n2wos_server>();
    s->cfg = *cfg;
    N2_HIP(hipGetDevice(&s->device));
    void* base = nullptr;
    if (hipMalloc(&base, cfg->hbm_budget_bytes) != hipSuccess) {
      (void)hipGetLastError();
      throw Error(N2WOS_ERR_NOMEM, "cannot allocate the request arena");
    }
    s->arena_base = static_cast<char*>(base);
    s->arena = std::make_unique<Arena>(cfg->hbm_budget_bytes);
    N2_HIP(hipStreamCreateWithFlags(&s->control, hipStreamNonBlocking));
    for (EstBatch& eb : s->est) {
      N2_HIP(hipMalloc(&eb.d_items, sizeof(EstItem) * kMaxPieces));
      N2_HIP(hipMalloc(&eb.d_out, sizeof(EstOut) * kMaxPieces));
      N2_HIP(hipHostMalloc(reinterpret_cast<void**>(&eb.h_items), sizeof(EstItem) * kMaxPieces));
      N2_HIP(hipHostMalloc(reinterpret_cast<void**>(&eb.h_out), sizeof(EstOut) * kMaxPieces));
      N2_HIP(hipEventCreateWithFlags(&eb.done, hipEventDisableTiming));
      s->fixed_bytes += (sizeof(EstItem) + sizeof(EstOut)) * kMaxPieces;
    }
    // the origin of the device-time intervals of the launches
    N2_HIP(hipEventCreate(&s->t_ref));
    N2_HIP(hipEventRecord(s->t_ref, s->control));
    N2_HIP(hipEventSynchronize(s->t_ref));
    *out = s.release();
    return N2WOS_OK;
  });
}

extern "C" int n2wos_server_destroy(n2wos_server* s) {
  return guarded([&]() -> int {
    if (!s) return N2WOS_OK;
    for (auto& m : s->meshes)
      for (Slot& sl : m->slot) free_slot(sl);
    if (s->control) (void)hipStreamSynchronize(s->control);
    (void)hipFree(s->arena_base);
    for (auto& kv : s->requests)
      if (kv.second->h_res) (void)hipHostFree(kv.second->h_res);
    for (ResBatch& rb : s->res)
      if (rb.done) (void)hipEventDestroy(rb.done);
    for (EstBatch& eb : s->est) {
      (void)hipFree(eb.d_items);
      (void)hipFree(eb.d_out);
      (void)hipHostFree(eb.h_items);
      (void)hipHostFree(eb.h_out);
      if (eb.done) (void)hipEventDestroy(eb.done);
    }
    if (s->t_ref) (void)hipEventDestroy(s->t_ref);
    if (s->control) (void)hipStreamDestroy(s->control);
    delete s;
    return N2WOS_OK;
  });
}

extern "C" int n2wos_server_add_mesh(n2wos_server* s, const n2wos_bvh* bvh, const n2wos_boundary* boundary,
                                     const n2wos_walk_params* params, int* mesh_id) {
  return guarded([&]() -> int {
    N2_REQUIRE(s && bvh && boundary && params && mesh_id, N2WOS_ERR_NULL, "NULL argument");
    N2_REQUIRE(bvh->d_nodes && bvh->d_tris, N2WOS_ERR_EMPTY, "mesh has no geometry");
    N2_REQUIRE(s->meshes.size() < 64, N2WOS_ERR_INVALID, "at most 64 meshes per server");
    check_walk_params(params);
    auto m = std::make_unique<Mesh>();
    m->bvh = bvh;
    m->bc = to_boundary(boundary);
    m->wp = *params;
    m->wp.d_visit_counts = nullptr;
    m->wp.event_begin = nullptr;
    m->wp.event_end = nullptr;
    m->cap = static_cast<int>(s->cfg.tick_walks / kUnit);
    m->unit = kUnit;
    m->first_round = s->cfg.first_round_walks;
    alloc_slots(s, *m);
    *mesh_id = static_cast<int>(s->meshes.size());
    s->meshes.push_back(std::move(m));
    return N2WOS_OK;
  });
}

extern "C" int n2wos_server_set_mesh_field(n2wos_server* s, int mesh_id, const n2wos_cv_field* field,
                                           int unit_walks, int first_round_walks) {
  return guarded([&]() -> int {
    N2_REQUIRE(s, N2WOS_ERR_NULL, "NULL server");
    N2_REQUIRE(mesh_id >= 0 && mesh_id < static_cast<int>(s->meshes.size()), N2WOS_ERR_INVALID, "unknown mesh");
    for (const auto& kv : s->requests)
      N2_REQUIRE(kv.second->mesh != mesh_id, N2WOS_ERR_INVALID,
                 "set a mesh's field before any request is submitted to it");
    Mesh& m = *s->meshes[static_cast<size_t>(mesh_id)];
    const bool cv = field && field->points != 0;
    int unit = kUnit, first = s->cfg.first_round_walks;
    if (cv) {
      N2_REQUIRE(field->points == 2 || field->points == 6, N2WOS_ERR_INVALID, "field points must be 2 or 6");
      N2_REQUIRE(field->d_values, N2WOS_ERR_NULL, "NULL field values");
      N2_REQUIRE(field->res >= 1 && field->res <= 1024, N2WOS_ERR_INVALID, "field res must be in [1, 1024]");
      for (int k = 0; k < 3; ++k)
        N2_REQUIRE(std::isfinite(field->origin[k]) && std::isfinite(field->cell_inverse[k]) &&
                       field->cell_inverse[k] > 0.0f,
                   N2WOS_ERR_INVALID, "field box must be finite with positive cell_inverse");
      N2_REQUIRE(std::isfinite(field->beta), N2WOS_ERR_INVALID, "field beta must be finite");
      N2_REQUIRE(unit_walks >= 1 && unit_walks <= kUnit && (unit_walks & (unit_walks - 1)) == 0, N2WOS_ERR_INVALID,
                 "unit_walks must be a power of two in [1, 64]");
      N2_REQUIRE(first_round_walks >= unit_walks && first_round_walks % unit_walks == 0 &&
                     first_round_walks <= s->cfg.max_round_walks && first_round_walks <= s->cfg.max_walks_per_point,
                 N2WOS_ERR_INVALID, "first_round_walks must be a multiple of unit_walks within the round limits");
      unit = unit_walks;
      first = first_round_walks;
    }
    const int64_t cap = s->cfg.tick_walks / unit;
    N2_REQUIRE(cap < (int64_t{1} << 30), N2WOS_ERR_INVALID, "tick_walks / unit_walks too large");
    N2_HIP(hipDeviceSynchronize());
    for (Slot& sl : m.slot) free_slot(sl);
    m.cv = cv;
    m.field = cv ? *field : n2wos_cv_field{};
    m.unit = unit;
    m.first_round = first;
    m.cap = static_cast<int>(cap);
    alloc_slots(s, m);
    return N2WOS_OK;
  });
}

extern "C" int n2wos_server_submit(n2wos_server* s, int mesh_id, const float* host_points, int n_points,
                                   double target_rmse, uint32_t rng_base, int64_t* request_id) {
  return guarded([&]() -> int {
    N2_REQUIRE(s && host_points && request_id, N2WOS_ERR_NULL, "NULL argument");
    N2_REQUIRE(mesh_id >= 0 && mesh_id < static_cast<int>(s->meshes.size()), N2WOS_ERR_INVALID, "unknown mesh");
    N2_REQUIRE(n_points > 0, N2WOS_ERR_INVALID, "a request needs at least one point");
    N2_REQUIRE(target_rmse > 0.0, N2WOS_ERR_INVALID, "target_rmse must be > 0");
    N2_REQUIRE(static_cast<int64_t>(rng_base) + n_points <= (int64_t{1} << 32), N2WOS_ERR_INVALID,
               "rng_base + n_points exceeds 32 bits");
    const int64_t bytes = request_bytes(n_points);
    N2_REQUIRE(bytes <= s->arena->size(), N2WOS_ERR_NOMEM, "request larger than the server's HBM budget");
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->mesh = mesh_id;
    r->n = n_points;
    r->target = target_rmse;
    r->rng_base = rng_base;
    r->unit = s->meshes[static_cast<size_t>(mesh_id)]->unit;
    r->bytes = bytes;
    r->host_points.assign(host_points, host_points + 3 * static_cast<int64_t>(n_points));
    N2_HIP(hipHostMalloc(reinterpret_cast<void**>(&r->h_res), result_bytes(n_points)));
    r->t_submit = s->now();
    *request_id = r->id;
    s->queue.push_back(r->id);
    s->requests.emplace(r->id, std::move(r));
    ++s->submitted;
    return N2WOS_OK;
  });
}

extern "C" int n2wos_server_step(n2wos_server* s, int* n_done) {
  return guarded([&]() -> int {
    N2_REQUIRE(s, N2WOS_ERR_NULL, "NULL server");
    if (n_done) *n_done = 0;
    N2_HIP(hipSetDevice(s->device));
    server_step(s, n_done);
    return N2WOS_OK;
  });
}

extern "C" int n2wos_server_request_info(const n2wos_server* s, int64_t request_id, n2wos_request_info* info) {
  return guarded([&]() -> int {
    N2_REQUIRE(s && info, N2WOS_ERR_NULL, "NULL argument");
    auto it = s->requests.find(request_id);
    N2_REQUIRE(it != s->requests.end(), N2WOS_ERR_INVALID, "unknown request id");
    const Request& r = *it->second;
    info->state = r.state;
    info->mesh = r.mesh;
    info->n_points = r.n;
    info->rounds = r.rounds;
    info->walks_per_point = r.wpp;
    info->estimated_rmse = r.est;
    info->walk_steps = r.total_steps;
    info->t_submit = r.t_submit;
    info->t_admit = r.t_admit;
    info->t_done = r.t_done;
    info->control_variate = s->meshes[static_cast<size_t>(r.mesh)]->cv ? 1 : 0;
    return N2WOS_OK;
  });
}

extern "C" int n2wos_server_take_result(n2wos_server* s, int64_t request_id, double* n, double* mean, double* m2,
                                        uint64_t* steps, uint32_t* forced) {
  return guarded([&]() -> int {
    N2_REQUIRE(s, N2WOS_ERR_NULL, "NULL server");
    auto it = s->requests.find(request_id);
    N2_REQUIRE(it != s->requests.end(), N2WOS_ERR_INVALID, "unknown request id");
    Request& r = *it->second;
    N2_REQUIRE(r.state == N2WOS_REQ_DONE, N2WOS_ERR_INVALID, "request has not finished");
    const size_t p = static_cast<size_t>(r.n);
    const char* h = r.h_res;
    if (n) std::memcpy(n, h, 8 * p);
    if (mean) std::memcpy(mean, h + 8 * p, 8 * p);
    if (m2) std::memcpy(m2, h + 16 * p, 8 * p);
    if (steps) std::memcpy(steps, h + 24 * p, 8 * p);
    if (forced) std::memcpy(forced, h + 32 * p, 4 * p);
    (void)hipHostFree(r.h_res);
    r.h_res = nullptr;
    s->requests.erase(it);
    return N2WOS_OK;
  });
}

extern "C" int n2wos_server_get_stats(const n2wos_server* s, n2wos_server_stats* st) {
  return guarded([&]() -> int {
    N2_REQUIRE(s && st, N2WOS_ERR_NULL, "NULL argument");
    st->ticks = s->ticks;
    st->requests_submitted = s->submitted;
    st->requests_done = s->done;
    st->walks = s->walks;
    st->walk_steps = s->steps;
    st->arena_bytes = s->arena->size();
    st->arena_used_bytes = s->arena->used();
    st->fixed_bytes = s->fixed_bytes;
    st->busy_seconds = s->busy;
    st->launches = s->launches;
    device_busy(s, &st->gpu_busy_seconds, &st->gpu_span_seconds);
    return N2WOS_OK;
  });
}
