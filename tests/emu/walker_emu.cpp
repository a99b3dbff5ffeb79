// walker_emu.cpp -- CPU SIMT emulation of the persistent walker kernel.
//
// Test infrastructure.  One wave of 64 lanes runs over the kernel source itself
// (n2wos-cuda_amd/csrc/walker_kernel.hpp) as 64 coroutines (ucontext) on one
// host thread.  A lane runs until it reaches a wave-wide operation (__ballot,
// __shfl, __shfl_down, wave_sync); when all 64 lanes have arrived at it, the
// scheduler computes the result for every lane and resumes them in order.
// This is exact for code that reaches each wave-wide operation with the whole
// wave, as every call site in the kernel does; a lane arriving at a different
// operation is reported as a divergence error.  A watchdog bounds the main
// loop and prints each lane's state if the wave stops making progress.
#include <ucontext.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hip/hip_runtime.h"

EmuDim3 threadIdx;
EmuDim3 blockIdx;  // the emulated wave is block 0

namespace emu {
enum Op { kNone = 0, kBallot, kShfl, kShflDown, kSync, kDone };
struct Lane {
  ucontext_t ctx;
  int op = kNone;
  uint64_t u = 0;
  double d = 0.0;
  int arg = 0;
  uint64_t ru = 0;
  double rd = 0.0;
};
Lane lanes[64];
ucontext_t sched;
int cur = 0;
long iter_limit = 0;
long iters = 0;
long trace = 0;
int hung = 0;
int diverged = 0;
long path_iters[3];
long path_lanes[3];
long tri_count = -1;  // records in the triangle array (-1: not checked)
long tri_oob = 0;     // triangle-record loads outside [0, tri_count)

inline void arrive(int op) {
  lanes[cur].op = op;
  swapcontext(&lanes[cur].ctx, &sched);
}
}  // namespace emu

inline uint64_t __ballot(int pred) {
  emu::lanes[emu::cur].u = pred ? 1u : 0u;
  emu::arrive(emu::kBallot);
  return emu::lanes[emu::cur].ru;
}

inline int __shfl(int v, int src) {
  emu::lanes[emu::cur].u = static_cast<uint32_t>(v);
  emu::lanes[emu::cur].arg = src;
  emu::arrive(emu::kShfl);
  return static_cast<int>(static_cast<uint32_t>(emu::lanes[emu::cur].ru));
}

inline double __shfl_down(double v, int dlt) {
  emu::lanes[emu::cur].d = v;
  emu::lanes[emu::cur].arg = dlt;
  emu::arrive(emu::kShflDown);
  return emu::lanes[emu::cur].rd;
}

#define N2_EMULATED_WAVE 1
#define N2_WAVES_PER_BLOCK 1
// Watchdog: lane 0 counts iterations; past the limit every lane prints its
// state once and the wave leaves the loop together.
#define N2_EMU_LOOP_HOOK()                                                                         \
  if (lane == 0) ++emu::iters;                                                                     \
  wave_sync();                                                                                     \
  if (emu::iters > emu::iter_limit) {                                                              \
    std::fprintf(stderr, "lane %d phase %d step %d point %d walk %d cur %d te %d cnt %d "          \
                 "scnt %d bid %d bd2 %g thr %g flags %x\n", lane, phase, step, point, walk, cur, te,  \
                 cnt, scnt, bid, bd2, thr, wflags);                                                \
    if (lane == 0) emu::hung = 1;                                                                  \
    wave_sync();                                                                                   \
    break;                                                                                         \
  }
// Bounds check of every triangle-record load against the array's size.  A
// load outside it is counted, then served from record 0, so the failure is
// reported instead of reading past the array.
#define N2_EMU_TRI_HOOK(idx)                                                                       \
  if (emu::tri_count >= 0 && ((idx) < 0 || (idx) >= emu::tri_count)) {                             \
    ++emu::tri_oob;                                                                                \
    (idx) = 0;                                                                                     \
  }
#define N2_EMU_TRACE(sN, sT, sU)                                                                   \
  if (lane == 0) {                                                                                 \
    const int path = ((sU) >= (sN) && (sU) >= (sT)) ? 2 : ((sN) >= (sT) ? 0 : 1);                  \
    const int served = __builtin_popcountll(path == 2 ? mU : (path == 0 ? mN : mT));               \
    emu::path_iters[path] += 1;                                                                    \
    emu::path_lanes[path] += served;                                                               \
    if (emu::trace > 0 && emu::iters <= emu::trace)                                                \
      std::fprintf(stderr, "iter %ld sN %d sT %d sU %d\n", emu::iters, sN, sT, sU);               \
  }

namespace n2 {
inline float hw_sqrt(float x) { return std::sqrt(x); }
inline uint32_t opaque_u32(uint32_t v) { return v; }
inline int opaque_lane_int(int v) { return v; }
inline int lane_id() { return emu::cur; }
inline int rank_in(uint64_t mask) {
  const int l = emu::cur;
  return l == 0 ? 0 : __builtin_popcountll(mask & ((~0ull) >> (64 - l)));
}
inline void wave_sync() { emu::arrive(emu::kSync); }
inline uint32_t lane_read(uint32_t v, int src) { return static_cast<uint32_t>(__shfl(static_cast<int>(v), src)); }
inline void wave_global_sync() { wave_sync(); }
inline float load_l2(const float* p) { return *p; }
// the emulated wave's XCD: set through emu_set_unit_ranges
inline int emu_xcd = 3;
inline int emu_ranges = 8;
inline const int* emu_order = nullptr;
inline int xcd_id() { return emu_xcd; }
}  // namespace n2
// the emulated block is one wave
inline void __syncthreads() { n2::wave_sync(); }

#include "../../n2wos-cuda_amd/csrc/walker_kernel.hpp"

namespace {

n2::WosKernelArgs g_args;
bool g_poisson = false;
bool g_ext = false;          // run the extended walker (control variate / path record)
n2::WosKernelArgs g_ext_args{};  // its control-variate and path fields

int g_cap = 16;  // LDS ring entries per lane: 16 (the library's default) or 8

template <int CAP>
void run_lane() {
  if (g_ext) {
    if (g_poisson)
      n2::k_wos_persistent<CAP, 8, true, false, true>(g_args);
    else
      n2::k_wos_persistent<CAP, 8, false, false, true>(g_args);
  } else if (g_poisson) {
    n2::k_wos_persistent<CAP, 8, true, true>(g_args);
  } else {
    n2::k_wos_persistent<CAP, 8, false, true>(g_args);
  }
}

void lane_main() {
  if (g_cap == 8)
    run_lane<8>();
  else
    run_lane<16>();
  emu::lanes[emu::cur].op = emu::kDone;
  swapcontext(&emu::lanes[emu::cur].ctx, &emu::sched);
}

int run_wave() {
  const size_t stack_bytes = 512 * 1024;
  std::vector<std::vector<char>> stacks(64, std::vector<char>(stack_bytes));
  for (int l = 0; l < 64; ++l) {
    getcontext(&emu::lanes[l].ctx);
    emu::lanes[l].ctx.uc_stack.ss_sp = stacks[l].data();
    emu::lanes[l].ctx.uc_stack.ss_size = stack_bytes;
    emu::lanes[l].ctx.uc_link = nullptr;
    makecontext(&emu::lanes[l].ctx, lane_main, 0);
    emu::lanes[l].op = emu::kNone;
  }
  for (;;) {
    for (int l = 0; l < 64; ++l) {
      emu::cur = l;
      threadIdx.x = static_cast<unsigned>(l);
      swapcontext(&emu::sched, &emu::lanes[l].ctx);
    }
    const int op = emu::lanes[0].op;
    for (int l = 1; l < 64; ++l) {
      if (emu::lanes[l].op != op) {
        emu::diverged = 1;
        std::fprintf(stderr, "divergent wave operation: lane 0 at %d, lane %d at %d\n", op, l, emu::lanes[l].op);
        return -2;
      }
    }
    if (op == emu::kDone) return emu::hung ? -1 : 0;
    if (op == emu::kBallot) {
      uint64_t m = 0;
      for (int l = 0; l < 64; ++l) m |= (emu::lanes[l].u & 1u) << l;
      for (int l = 0; l < 64; ++l) emu::lanes[l].ru = m;
    } else if (op == emu::kShfl) {
      for (int l = 0; l < 64; ++l) emu::lanes[l].ru = emu::lanes[emu::lanes[l].arg & 63].u;
    } else if (op == emu::kShflDown) {
      for (int l = 0; l < 64; ++l) {
        const int s = l + emu::lanes[l].arg;
        emu::lanes[l].rd = s < 64 ? emu::lanes[s].d : emu::lanes[l].d;
      }
    }
  }
}

}  // namespace

extern "C" int emu_walker(const void* nodes, const void* tris, float delta, const float* points, int n_points,
                          int wpp, int max_steps, float eps, float step_scale, uint64_t seed, int poisson,
                          int bmode, const float* charges, int n_charges, const float* first_d2,
                          const float* first_cp, const int* first_ovf, float* walk_values, int* walk_steps,
                          double* unit_mean, double* unit_m2, uint32_t* unit_steps, uint32_t* unit_flags,
                          uint64_t* visits, int refill_min, int w_node, int w_tri, int w_update,
                          long iter_limit) {
  n2::WosKernelArgs a{};
  a.nodes = static_cast<const n2::Bvh4Node*>(nodes);
  a.tris = static_cast<const n2::TriRec*>(tris);
  a.delta = delta;
  a.bc.mode = bmode;
  a.bc.n_charges = n_charges;
  for (int i = 0; i < n_charges; ++i)
    for (int k = 0; k < 4; ++k) a.bc.charge[i][k] = charges[4 * i + k];
  a.points = points;
  a.walks_per_point = wpp;
  a.units_per_point = (wpp + 63) / 64;
  a.total_units = n_points * a.units_per_point;
  a.max_steps = max_steps;
  a.epsilon = eps;
  a.step_scale = step_scale;
  a.seed_lo = static_cast<uint32_t>(seed);
  a.seed_hi = static_cast<uint32_t>(seed >> 32);
  a.stream = 0;
  a.point_offset = 0;
  static int next_unit[n2::kMaxUnitRanges * n2::kUnitCounterStride];
  for (int& c : next_unit) c = 0;
  a.next_unit = next_unit;
  a.unit_ranges = n2::emu_ranges;
  a.order = n2::emu_order;
  static float unit_vals[8 * 64];  // staging of the one emulated wave's 8 unit slots
  a.unit_vals = unit_vals;
  static uint64_t spill_lines[64 * n2::kLineSpill];  // the one emulated wave's spill lines
  a.spill_lines = spill_lines;
  a.unit_mean = unit_mean;
  a.unit_m2 = unit_m2;
  a.unit_steps = unit_steps;
  a.unit_flags = unit_flags;
  a.walk_values = walk_values;
  a.walk_steps = walk_steps;
  a.first_d2 = first_d2;
  a.first_cp = first_cp;
  a.first_ovf = first_ovf;
  a.visits = visits;
  a.refill_min = refill_min;
  a.weight_node = w_node;
  a.weight_tri = w_tri;
  a.weight_update = w_update;
  a.n_points = n_points;
  a.pack_shift = -1;
  // packed units for wpp = 1, 2, 4, ..., 32, as the host launch sets them up
  for (int k = 0; k <= 5; ++k)
    if (wpp == (1 << k)) a.pack_shift = k;
  if (a.pack_shift >= 0) {
    const int ppu = 64 >> a.pack_shift;
    a.total_units = (n_points + ppu - 1) / ppu;
    for (int i = 0; i < n_points; ++i) unit_steps[i] = unit_flags[i] = 0u;
  }
  if (g_ext) {
    a.cv_values = g_ext_args.cv_values;
    a.cv_res = g_ext_args.cv_res;
    a.cv_points = g_ext_args.cv_points;
    for (int k = 0; k < 3; ++k) {
      a.cv_origin[k] = g_ext_args.cv_origin[k];
      a.cv_inv_h[k] = g_ext_args.cv_inv_h[k];
    }
    a.cv_beta = g_ext_args.cv_beta;
    a.path_pos = g_ext_args.path_pos;
    a.path_q = g_ext_args.path_q;
    a.path_len = g_ext_args.path_len;
  }
  g_args = a;
  g_poisson = poisson != 0;
  emu::iter_limit = iter_limit;
  emu::iters = 0;
  emu::hung = 0;
  emu::diverged = 0;
  for (int k = 0; k < 3; ++k) emu::path_iters[k] = emu::path_lanes[k] = 0;
  return run_wave();
}

// Selects the extended walker for the next emu_walker calls (points 0 and
// path_len 0: the plain walker again).
extern "C" void emu_set_ext(const float* values, int res, int points, const float* origin, const float* inv_h,
                            float beta, float* path_pos, float* path_q, int path_len) {
  g_ext = points > 0 || path_len > 0;
  g_ext_args = n2::WosKernelArgs{};
  g_ext_args.cv_values = values;
  g_ext_args.cv_res = res;
  g_ext_args.cv_points = points;
  for (int k = 0; k < 3; ++k) {
    g_ext_args.cv_origin[k] = origin ? origin[k] : 0.0f;
    g_ext_args.cv_inv_h[k] = inv_h ? inv_h[k] : 0.0f;
  }
  g_ext_args.cv_beta = beta;
  g_ext_args.path_pos = path_pos;
  g_ext_args.path_q = path_q;
  g_ext_args.path_len = path_len;
}

extern "C" long emu_iterations() { return emu::iters; }
// Sets the triangle array's size for the load bounds check (-1: off) and
// clears the count of loads outside it.
// Unit ranges of the claim (1-8) and the XCD the emulated wave reports.
extern "C" void emu_set_unit_ranges(int ranges, int xcd) {
  n2::emu_ranges = ranges;
  n2::emu_xcd = xcd;
}

// Claim order: positions -> points (NULL: identity).
extern "C" void emu_set_order(const int* order) { n2::emu_order = order; }

extern "C" void emu_set_tri_count(long n) {
  emu::tri_count = n;
  emu::tri_oob = 0;
}
extern "C" long emu_tri_oob() { return emu::tri_oob; }
extern "C" void emu_set_trace(long n) { emu::trace = n; }
extern "C" void emu_path_stats(long* iters, long* lanes) {
  for (int k = 0; k < 3; ++k) {
    iters[k] = emu::path_iters[k];
    lanes[k] = emu::path_lanes[k];
  }
}

// LDS ring capacity of the emulated kernel for the next emu_walker calls.
extern "C" void emu_set_cap(int cap) { g_cap = cap == 8 ? 8 : 16; }
