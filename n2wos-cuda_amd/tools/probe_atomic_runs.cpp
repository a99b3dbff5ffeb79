// n2wos_probe_atomic_runs -- rate of no-return FP32 atomic adds on the device
// by run length: each group of R consecutive lanes adds to R contiguous
// dwords at a random R-aligned place of a 64 MiB buffer (R = 1, 2, 4, 8, 16,
// 32, 64).  Prints runs/s and dwords/s per R: the memory-side atomic unit
// takes one request per 64-byte run, so runs/s is flat up to R = 16 and
// halves with each doubling beyond.  Used to size the learned cache's
// gradient scatter (one 2 x 2 x 2 brick = 64 bytes).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

__device__ __forceinline__ uint32_t mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// runs_mask + 1 runs of R dwords: every index stays below (runs_mask + 1) R
template <int R>
__global__ void __launch_bounds__(256) k_runs(float* buf, uint32_t runs_mask, int iters) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t grp = t / R, k = t % R;
  for (int i = 0; i < iters; ++i) {
    const uint32_t run = mix(grp * 7919u + static_cast<uint32_t>(i) * 104729u) & runs_mask;
    atomicAdd(buf + static_cast<size_t>(run) * R + k, 1.0f);
  }
}

bool check(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
  return false;
}

template <int R>
bool run(float* buf, size_t nfl, hipEvent_t a, hipEvent_t b) {
  const uint32_t runs = static_cast<uint32_t>(nfl / R);  // a power of two
  const int blocks = 256 * 32, iters = 64;
  hipLaunchKernelGGL(k_runs<R>, dim3(blocks), dim3(256), 0, 0, buf, runs - 1, iters);  // warm-up
  if (!check(hipEventRecord(a), "hipEventRecord")) return false;
  hipLaunchKernelGGL(k_runs<R>, dim3(blocks), dim3(256), 0, 0, buf, runs - 1, iters);
  if (!check(hipEventRecord(b), "hipEventRecord") || !check(hipEventSynchronize(b), "hipEventSynchronize"))
    return false;
  float ms = 0.0f;
  if (!check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime")) return false;
  const double dw = static_cast<double>(blocks) * 256 * iters;
  std::printf("{\"run_lanes\": %d, \"ms\": %.4f, \"dwords_per_s\": %.4e, \"runs_per_s\": %.4e}\n", R, ms, dw / (ms * 1e-3),
              dw / R / (ms * 1e-3));
  return true;
}

}  // namespace

int main() {
  const size_t nfl = (size_t{64} << 20) / 4;
  float* buf = nullptr;
  hipEvent_t a, b;
  if (!check(hipMalloc(&buf, nfl * 4), "hipMalloc") || !check(hipMemset(buf, 0, nfl * 4), "hipMemset") ||
      !check(hipEventCreate(&a), "hipEventCreate") || !check(hipEventCreate(&b), "hipEventCreate"))
    return 1;
  const bool ok = run<1>(buf, nfl, a, b) && run<2>(buf, nfl, a, b) && run<4>(buf, nfl, a, b) &&
                  run<8>(buf, nfl, a, b) && run<16>(buf, nfl, a, b) && run<32>(buf, nfl, a, b) &&
                  run<64>(buf, nfl, a, b);
  (void)hipFree(buf);
  return ok ? 0 : 1;
}
