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
// for requests whose round is fully merged a fixed-order block reduction
// gives the estimated RMSE; finished requests are copied out and their arena
// ranges freed; freed slots are refilled at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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
  int pad;
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
  brng[2 * i + 1] = p.walk_base + static_cast<uint32_t>(kUnit * c);
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
  int count;
  int pad;
};

constexpr int kEstBlock = 256;

// out[b] = mean over the request's points of var_i / n_i (fixed reduction order).
__global__ void __launch_bounds__(kEstBlock) k_serve_estimate(const EstItem* __restrict__ items,
                                                              double* __restrict__ out) {
  __shared__ double s[kEstBlock];
  const EstItem it = items[blockIdx.x];
  double acc = 0.0;
  for (int j = threadIdx.x; j < it.count; j += kEstBlock) {
    const double n = it.n[j];
    if (n > 1.0) acc += it.m2[j] / (n - 1.0) / n;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kEstBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0] / static_cast<double>(it.count);
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
  Slot slot[kSlotsPerMesh];
  int64_t bytes = 0;
};

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
  int wpp = 0;            // walks merged per point
  int round_walks = 0;    // walks per point of the current round
  int cursor = 0;         // points of the current round already batched
  int inflight = 0;       // busy slots holding a piece of this request
  double est = -1.0;
  uint64_t total_steps = 0;
  double t_submit = -1.0, t_admit = -1.0, t_done = -1.0;
  std::vector<double> r_n, r_mean, r_m2;
  std::vector<uint64_t> r_steps;
  std::vector<uint32_t> r_forced;
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
  n2::EstItem* d_est_items = nullptr;
  n2::EstItem* h_est_items = nullptr;
  double* d_est = nullptr;
  double* h_est = nullptr;
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
    r.round_walks = s->cfg.first_round_walks;
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
    units[q] = r.round_walks / kUnit;
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
    pc.pad = 0;
    entries += grant[q] * units[q];
    points += grant[q];
    r.cursor += grant[q];
    r.inflight += 1;
    sl.reqs.push_back(r.id);
  }
  if (entries == 0) return false;
  hipStream_t st = sl.stream;
  N2_HIP(hipMemcpyAsync(sl.d_pieces, sl.h_pieces, sizeof(ServePiece) * np, hipMemcpyHostToDevice, st));
  const int b = 256;
  hipLaunchKernelGGL(k_serve_gather, dim3((entries + b - 1) / b), dim3(b), 0, st, sl.d_pieces, np, entries, sl.bpts,
                     sl.brng);
  N2_HIP(hipGetLastError());
  WosRun w{};
  w.d_points = sl.bpts;
  w.n_points = entries;
  w.walks_per_point = kUnit;
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
  run_wos(m.bvh, m.bc, w, st);
  hipLaunchKernelGGL(k_serve_merge, dim3((points + b - 1) / b), dim3(b), 0, st, sl.d_pieces, np, points, sl.bn,
                     sl.bmean, sl.bm2, sl.bsteps, sl.bforced);
  N2_HIP(hipGetLastError());
  N2_HIP(hipEventRecord(sl.done, st));
  sl.busy = true;
  sl.seq = ++s->launches;
  sl.entries = entries;
  s->walks += static_cast<int64_t>(entries) * kUnit;
  return true;
}

void launch_free_slots(n2wos_server* s) {
  for (int mi = 0; mi < static_cast<int>(s->meshes.size()); ++mi)
    for (Slot& sl : s->meshes[mi]->slot)
      if (!sl.busy && !launch_slot(s, mi, sl)) break;
}

// Retires a completed slot: requests whose round is fully merged go to `round_done`.
void retire(n2wos_server* s, Slot& sl, std::vector<int64_t>& round_done) {
  for (int64_t id : sl.reqs) {
    Request& r = *s->requests.at(id);
    r.inflight -= 1;
    if (r.inflight == 0 && r.cursor == r.n) round_done.push_back(id);
  }
  sl.reqs.clear();
  sl.busy = false;
}

int finish_rounds(n2wos_server* s, const std::vector<int64_t>& round_done) {
  if (round_done.empty()) return 0;
  hipStream_t st = s->control;
  const int k = static_cast<int>(round_done.size());
  for (int q = 0; q < k; ++q) {
    const Request& r = *s->requests.at(round_done[q]);
    s->h_est_items[q] = EstItem{r.d_n, r.d_m2, r.n, 0};
  }
  N2_HIP(hipMemcpyAsync(s->d_est_items, s->h_est_items, sizeof(EstItem) * k, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_serve_estimate, dim3(k), dim3(kEstBlock), 0, st, s->d_est_items, s->d_est);
  N2_HIP(hipGetLastError());
  N2_HIP(hipMemcpyAsync(s->h_est, s->d_est, sizeof(double) * k, hipMemcpyDeviceToHost, st));
  N2_HIP(hipStreamSynchronize(st));
  int finished = 0;
  for (int q = 0; q < k; ++q) {
    Request& r = *s->requests.at(round_done[q]);
    r.rounds += 1;
    r.wpp += r.round_walks;
    r.est = std::sqrt(std::max(0.0, s->h_est[q]));
    const int room = s->cfg.max_walks_per_point - r.wpp;
    if (r.est <= r.target || room <= 0) {
      const size_t p = static_cast<size_t>(r.n);
      r.r_n.resize(p);
      r.r_mean.resize(p);
      r.r_m2.resize(p);
      r.r_steps.resize(p);
      r.r_forced.resize(p);
      N2_HIP(hipMemcpyAsync(r.r_n.data(), r.d_n, 8 * p, hipMemcpyDeviceToHost, st));
      N2_HIP(hipMemcpyAsync(r.r_mean.data(), r.d_mean, 8 * p, hipMemcpyDeviceToHost, st));
      N2_HIP(hipMemcpyAsync(r.r_m2.data(), r.d_m2, 8 * p, hipMemcpyDeviceToHost, st));
      N2_HIP(hipMemcpyAsync(r.r_steps.data(), r.d_steps, 8 * p, hipMemcpyDeviceToHost, st));
      N2_HIP(hipMemcpyAsync(r.r_forced.data(), r.d_forced, 4 * p, hipMemcpyDeviceToHost, st));
      N2_HIP(hipStreamSynchronize(st));
      uint64_t tot = 0;
      for (uint64_t v : r.r_steps) tot += v;
      r.total_steps = tot;
      s->steps += tot;
      s->arena->release(r.off, r.bytes);
      r.off = -1;
      r.d_points = nullptr;
      r.state = N2WOS_REQ_DONE;
      r.t_done = s->now();
      s->active.erase(std::find(s->active.begin(), s->active.end(), r.id));
      ++s->done;
      ++finished;
    } else {
      // walks still missing by the estimate (variance ~ 1 / n), 10 % margin
      const double ratio = (r.est / r.target) * (r.est / r.target);
      const double want = std::ceil(static_cast<double>(r.wpp) * (ratio - 1.0) * 1.1);
      int64_t w = (static_cast<int64_t>(want) + kUnit - 1) / kUnit * kUnit;
      w = std::max<int64_t>(kUnit, std::min<int64_t>({w, s->cfg.max_round_walks, room}));
      r.round_walks = static_cast<int>(w);
      r.cursor = 0;
    }
  }
  return finished;
}

// One step: admit, keep every free slot busy, wait for the oldest launch to
// complete, retire every completed launch, finish the rounds that ended, and
// refill the freed slots before returning.
void server_step(n2wos_server* s, int* n_done) {
  const double t_begin = s->now();
  server_admit(s);
  launch_free_slots(s);
  Slot* oldest = nullptr;
  for (auto& m : s->meshes)
    for (Slot& sl : m->slot)
      if (sl.busy && (!oldest || sl.seq < oldest->seq)) oldest = &sl;
  int finished = 0;
  if (oldest) {
    N2_HIP(hipEventSynchronize(oldest->done));
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
    finished = finish_rounds(s, round_done);
    server_admit(s);
    launch_free_slots(s);
  }
  ++s->ticks;
  s->busy += s->now() - t_begin;
  if (n_done) *n_done = finished;
}

void free_slot(Slot& sl) {
  if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  (void)hipFree(sl.buf);
  (void)hipFree(sl.d_pieces);
  (void)hipHostFree(sl.h_pieces);
  if (sl.done) (void)hipEventDestroy(sl.done);
  if (sl.stream) (void)hipStreamDestroy(sl.stream);
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
    N2_HIP(hipMalloc(&s->d_est_items, sizeof(EstItem) * kMaxPieces));
    N2_HIP(hipMalloc(&s->d_est, sizeof(double) * kMaxPieces));
    N2_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->h_est_items), sizeof(EstItem) * kMaxPieces));
    N2_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->h_est), sizeof(double) * kMaxPieces));
    s->fixed_bytes += sizeof(EstItem) * kMaxPieces + sizeof(double) * kMaxPieces;
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
    (void)hipFree(s->d_est_items);
    (void)hipFree(s->d_est);
    (void)hipHostFree(s->h_est_items);
    (void)hipHostFree(s->h_est);
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
    const int64_t c = m->cap;
    const int64_t ws_bytes = wos_workspace_bytes(m->cap, kUnit);
    const int64_t sz[8] = {align256(12 * c), align256(8 * c), align256(8 * c), align256(8 * c),
                           align256(8 * c), align256(8 * c), align256(4 * c), align256(ws_bytes)};
    int64_t total = 0;
    for (int64_t v : sz) total += v;
    try {
      for (Slot& sl : m->slot) {
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
        N2_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
      }
    } catch (...) {
      for (Slot& sl : m->slot) free_slot(sl);
      throw;
    }
    m->bytes = kSlotsPerMesh * (total + static_cast<int64_t>(sizeof(ServePiece)) * kMaxPieces);
    s->fixed_bytes += m->bytes;
    *mesh_id = static_cast<int>(s->meshes.size());
    s->meshes.push_back(std::move(m));
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
    r->bytes = bytes;
    r->host_points.assign(host_points, host_points + 3 * static_cast<int64_t>(n_points));
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
    if (n) std::copy(r.r_n.begin(), r.r_n.end(), n);
    if (mean) std::copy(r.r_mean.begin(), r.r_mean.end(), mean);
    if (m2) std::copy(r.r_m2.begin(), r.r_m2.end(), m2);
    if (steps) std::copy(r.r_steps.begin(), r.r_steps.end(), steps);
    if (forced) std::copy(r.r_forced.begin(), r.r_forced.end(), forced);
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
    return N2WOS_OK;
  });
}
