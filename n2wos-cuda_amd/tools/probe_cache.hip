// n2wos_probe_tcnn_cache -- learned-cache probe over the engine's C ABI.
//
// Flags, workload and JSON schema follow the reference probe
// (src/tools/probe_tcnn_cache.cu): --samples / --batch-size / --train-steps /
// --repeat / --warmup / --seed and the hash-grid and MLP shape flags (:40-57,
// :97-147), a fresh device-generated training batch of x^2 - y^2 targets on
// unit-cube points before every Adam step (:375-380), --warmup untimed then
// --repeat timed inference launches (:384-397), the same launches again each
// followed by a device consumer that reduces prediction, exact value, squared
// and absolute error per repeat (:399-406), and JSON
// "n2wos_tcnn_cache_probe_v1" (:440-508) with truthful implementation_mode
// labels.  Timing is HIP events on the probe's stream.
//
// The cache here is this engine's hash grid + ReLU MLP on MFMA (FP16 operands,
// FP32 accumulation); the unit cube is the cache's input box (origin 0,
// extent 1).  Points come from Philox4x32-10 under the probe's seed, one
// counter per (sample, step); the consumer reduces each wave in registers and
// adds one FP64 atomic per wave and statistic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/n2wos_amd.h"
#include "../csrc/geom.hpp"

namespace {

struct Options {
  int samples = 262144;
  int batch_size = 262144;
  int train_steps = 200;
  int repeat = 20;
  int warmup = 3;
  int seed = 12345;
  int n_levels = 12;
  int n_features_per_level = 2;
  int log2_hashmap_size = 18;
  int base_resolution = 8;
  float per_level_scale = 1.5f;
  int n_neurons = 32;
  int n_hidden_layers = 2;
  float learning_rate = 1.0e-2f;
  int jit = 0;
  std::string network = "FullyFusedMLP";
  std::string output = "results/probe_tcnn_cache.json";
};

void check(int rc, const char* what) {
  if (rc != N2WOS_OK) throw std::runtime_error(std::string(what) + ": " + n2wos_last_error());
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void usage(std::ostream& o, const char* argv0) {
  o << "Usage: " << argv0 << " [options]\n\nOptions:\n"
    << "  --samples <int>                Inference sample count [262144]\n"
    << "  --batch-size <int>             Training batch size [262144]\n"
    << "  --train-steps <int>            Adam training steps [200]\n"
    << "  --repeat <int>                 Repeated inference timings [20]\n"
    << "  --warmup <int>                 Untimed inference warmups [3]\n"
    << "  --seed <int>                   GPU sample-generation seed [12345]\n"
    << "  --network FullyFusedMLP|CutlassMLP\n"
    << "                                 Accepted for compatibility; both run this engine's MFMA MLP\n"
    << "  --n-levels <int>               Hash-grid levels, 1..16 [12]\n"
    << "  --n-features-per-level <int>   Hash-grid features per level (2 only) [2]\n"
    << "  --log2-hashmap-size <int>      Hash-grid log2 table size, 8..24 [18]\n"
    << "  --base-resolution <int>        Hash-grid base resolution [8]\n"
    << "  --per-level-scale <float>      Hash-grid scale per level [1.5]\n"
    << "  --n-neurons <int>              MLP neurons per hidden layer: 16, 32, 64 [32]\n"
    << "  --n-hidden-layers <int>        MLP hidden layers, 1..4 [2]\n"
    << "  --learning-rate <float>        Adam learning rate [1e-2]\n"
    << "  --jit 0|1                      Accepted; no JIT fusion here [0]\n"
    << "  --output <path>                JSON output [results/probe_tcnn_cache.json]\n"
    << "  --help\n";
}

std::string need(int& i, int argc, char** argv) {
  if (i + 1 >= argc) throw std::runtime_error(std::string("missing value for ") + argv[i]);
  return argv[++i];
}

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--help" || a == "-h") {
      usage(std::cout, argv[0]);
      std::exit(0);
    } else if (a == "--samples") {
      o.samples = std::stoi(need(i, argc, argv));
    } else if (a == "--batch-size") {
      o.batch_size = std::stoi(need(i, argc, argv));
    } else if (a == "--train-steps") {
      o.train_steps = std::stoi(need(i, argc, argv));
    } else if (a == "--repeat") {
      o.repeat = std::stoi(need(i, argc, argv));
    } else if (a == "--warmup") {
      o.warmup = std::stoi(need(i, argc, argv));
    } else if (a == "--seed") {
      o.seed = std::stoi(need(i, argc, argv));
    } else if (a == "--network") {
      o.network = need(i, argc, argv);
    } else if (a == "--n-levels") {
      o.n_levels = std::stoi(need(i, argc, argv));
    } else if (a == "--n-features-per-level") {
      o.n_features_per_level = std::stoi(need(i, argc, argv));
    } else if (a == "--log2-hashmap-size") {
      o.log2_hashmap_size = std::stoi(need(i, argc, argv));
    } else if (a == "--base-resolution") {
      o.base_resolution = std::stoi(need(i, argc, argv));
    } else if (a == "--per-level-scale") {
      o.per_level_scale = std::stof(need(i, argc, argv));
    } else if (a == "--n-neurons") {
      o.n_neurons = std::stoi(need(i, argc, argv));
    } else if (a == "--n-hidden-layers") {
      o.n_hidden_layers = std::stoi(need(i, argc, argv));
    } else if (a == "--learning-rate") {
      o.learning_rate = std::stof(need(i, argc, argv));
    } else if (a == "--jit") {
      o.jit = std::stoi(need(i, argc, argv));
    } else if (a == "--output") {
      o.output = need(i, argc, argv);
    } else {
      throw std::runtime_error("unknown argument: " + a);
    }
  }
  if (o.samples <= 0 || o.batch_size <= 0 || o.train_steps < 0 || o.repeat <= 0 || o.warmup < 0)
    throw std::runtime_error("samples, batch-size, repeat must be positive; train-steps and warmup must be nonnegative");
  if (o.network != "FullyFusedMLP" && o.network != "CutlassMLP")
    throw std::runtime_error("--network must be FullyFusedMLP or CutlassMLP");
  if (o.n_features_per_level != 2) throw std::runtime_error("this engine's hash grid has 2 features per level");
  return o;
}

std::string quote(const std::string& s) {
  std::string r = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') {
      r += '\\';
      r += c;
    } else if (static_cast<unsigned char>(c) < 0x20) {
      char b[8];
      std::snprintf(b, sizeof(b), "\\u%04x", c);
      r += b;
    } else {
      r += c;
    }
  }
  return r + "\"";
}

std::string utc_now() {
  const std::time_t t = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
  std::tm tm{};
  gmtime_r(&t, &tm);
  char b[32];
  std::strftime(b, sizeof(b), "%Y-%m-%dT%H:%M:%SZ", &tm);
  return b;
}

float median(std::vector<float> v) {
  if (v.empty()) return 0.0f;
  std::sort(v.begin(), v.end());
  const size_t m = v.size() / 2;
  return v.size() % 2 ? v[m] : 0.5f * (v[m - 1] + v[m]);
}

std::string vec(const std::vector<float>& v) {
  std::ostringstream o;
  o << std::setprecision(9) << '[';
  for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << v[i];
  return o.str() + "]";
}

__device__ __forceinline__ float harmonic(float ux, float uy) {
  const float x = 2.0f * ux - 1.0f, y = 2.0f * uy - 1.0f;
  return x * x - y * y;
}

// Sample i of batch `step`: a point of the unit cube and its x^2 - y^2 target
// (in [-1, 1] coordinates).
__global__ void k_fill(int n, uint32_t seed, uint32_t step, float* __restrict__ pts, float* __restrict__ targets) {
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const n2::U4 r = n2::philox4x32_10(n2::U4{static_cast<uint32_t>(i), step, 0x63616368u, 0x70726f62u}, seed,
                                     0x9e3779b9u);
  const float ux = n2::u01(r.x), uy = n2::u01(r.y), uz = n2::u01(r.z);
  pts[3 * i] = ux;
  pts[3 * i + 1] = uy;
  pts[3 * i + 2] = uz;
  targets[i] = harmonic(ux, uy);
}

// Downstream consumer: sums of prediction, exact value, squared and absolute
// error into stats[4 r .. 4 r + 3] (FP64), one atomic per wave and statistic.
__global__ void __launch_bounds__(256) k_consume(int n, int r, const float* __restrict__ pts,
                                                 const float* __restrict__ out, double* __restrict__ stats) {
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    const float exact = harmonic(pts[3 * i], pts[3 * i + 1]);
    const float pred = out[i];
    const float e = pred - exact;
    s[0] = pred;
    s[1] = exact;
    s[2] = static_cast<double>(e) * e;
    s[3] = fabs(static_cast<double>(e));
  }
  for (int k = 0; k < 4; ++k)
    for (int off = 32; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off);
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 4; ++k) atomicAdd(stats + 4 * r + k, s[k]);
}

struct Timer {
  hipEvent_t a{}, b{};
  Timer() {
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventCreate(&b), "hipEventCreate");
  }
  ~Timer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  void start(hipStream_t s) { hip_check(hipEventRecord(a, s), "hipEventRecord"); }
  float stop(hipStream_t s) {
    hip_check(hipEventRecord(b, s), "hipEventRecord");
    hip_check(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    return ms;
  }
};

}  // namespace

int main(int argc, char** argv) {
  std::string cmd;
  for (int i = 0; i < argc; ++i) cmd += (i ? " " : "") + std::string(argv[i]);
  try {
    const Options opt = parse(argc, argv);
    const int n = opt.samples;
    const int nb = opt.batch_size;
    hipStream_t s{};
    hip_check(hipStreamCreate(&s), "hipStreamCreate");

    n2wos_cache_config cfg{};
    cfg.n_levels = opt.n_levels;
    cfg.log2_table_size = opt.log2_hashmap_size;
    cfg.base_resolution = opt.base_resolution;
    cfg.growth = opt.per_level_scale;
    for (int a = 0; a < 3; ++a) {
      cfg.origin[a] = 0.0f;
      cfg.extent[a] = 1.0f;
    }
    cfg.learning_rate = opt.learning_rate;
    cfg.beta1 = 0.9f;
    cfg.beta2 = 0.99f;
    cfg.epsilon = 1e-15f;
    cfg.loss_scale = 128.0f;
    cfg.seed = static_cast<uint64_t>(opt.seed);
    cfg.max_batch = nb;
    cfg.n_neurons = opt.n_neurons;
    cfg.n_hidden_layers = opt.n_hidden_layers;
    n2wos_cache* cache = nullptr;
    check(n2wos_cache_create(&cfg, &cache), "n2wos_cache_create");

    float *train_pts = nullptr, *train_t = nullptr, *inf_pts = nullptr, *inf_out = nullptr;
    double* stats = nullptr;
    hip_check(hipMalloc(&train_pts, sizeof(float) * 3 * static_cast<size_t>(nb)), "hipMalloc");
    hip_check(hipMalloc(&train_t, sizeof(float) * static_cast<size_t>(nb)), "hipMalloc");
    hip_check(hipMalloc(&inf_pts, sizeof(float) * 3 * static_cast<size_t>(n)), "hipMalloc");
    hip_check(hipMalloc(&inf_out, sizeof(float) * static_cast<size_t>(n)), "hipMalloc");
    hip_check(hipMalloc(&stats, sizeof(double) * 4 * static_cast<size_t>(opt.repeat)), "hipMalloc");
    hip_check(hipMemsetAsync(stats, 0, sizeof(double) * 4 * static_cast<size_t>(opt.repeat), s), "hipMemsetAsync");
    const uint32_t seed = static_cast<uint32_t>(opt.seed);
    // inference points: their own stream of counters (key word 1 differs)
    hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, n, seed ^ 0xa341316cu, 0u, inf_pts, inf_out);
    hip_check(hipGetLastError(), "k_fill");

    Timer timer;
    float train_ms = 0.0f;
    if (opt.train_steps > 0) {
      timer.start(s);
      for (int step = 0; step < opt.train_steps; ++step) {
        hipLaunchKernelGGL(k_fill, dim3((nb + 255) / 256), dim3(256), 0, s, nb, seed, static_cast<uint32_t>(step),
                           train_pts, train_t);
        check(n2wos_cache_train_step(cache, train_pts, train_t, nb, nullptr, s), "n2wos_cache_train_step");
      }
      train_ms = timer.stop(s);
    }
    for (int i = 0; i < opt.warmup; ++i) check(n2wos_cache_infer(cache, inf_pts, n, inf_out, s), "infer");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    std::vector<float> inf_ms, cons_ms;
    for (int r = 0; r < opt.repeat; ++r) {
      timer.start(s);
      check(n2wos_cache_infer(cache, inf_pts, n, inf_out, s), "infer");
      inf_ms.push_back(timer.stop(s));
    }
    for (int r = 0; r < opt.repeat; ++r) {
      timer.start(s);
      check(n2wos_cache_infer(cache, inf_pts, n, inf_out, s), "infer");
      hipLaunchKernelGGL(k_consume, dim3((n + 255) / 256), dim3(256), 0, s, n, r, inf_pts, inf_out, stats);
      hip_check(hipGetLastError(), "k_consume");
      cons_ms.push_back(timer.stop(s));
    }
    std::vector<double> hs(4 * static_cast<size_t>(opt.repeat));
    hip_check(hipMemcpyAsync(hs.data(), stats, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, s), "readback");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    double sp = 0, se = 0, sse = 0, sa = 0;
    for (int r = 0; r < opt.repeat; ++r) {
      sp += hs[4 * r];
      se += hs[4 * r + 1];
      sse += hs[4 * r + 2];
      sa += hs[4 * r + 3];
    }
    const double denom = static_cast<double>(n) * opt.repeat;
    const double mse = sse / denom;
    const float med = median(inf_ms), cmed = median(cons_ms);
    const float mean = std::accumulate(inf_ms.begin(), inf_ms.end(), 0.0f) / static_cast<float>(inf_ms.size());
    const float cmean = std::accumulate(cons_ms.begin(), cons_ms.end(), 0.0f) / static_cast<float>(cons_ms.size());
    const double us_per = 1000.0 * med / n;
    const double msps = n / (1000.0 * med);
    const double cus_per = 1000.0 * cmed / n;

    char dev[256] = {0};
    n2wos_device_summary(dev, sizeof(dev));
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    const std::filesystem::path parent = std::filesystem::path(opt.output).parent_path();
    if (!parent.empty()) std::filesystem::create_directories(parent);
    std::ofstream o(opt.output);
    if (!o) throw std::runtime_error("failed to open output JSON: " + opt.output);
    const int pad_b = (nb + 15) / 16 * 16;
    o << std::setprecision(9);
    o << "{\n  \"schema\": \"n2wos_tcnn_cache_probe_v1\",\n"
      << "  \"patch\": \"n2wos_amd_gfx950\",\n"
      << "  \"generated_at_utc\": " << quote(utc_now()) << ",\n"
      << "  \"command_line\": " << quote(cmd) << ",\n"
      << "  \"cuda\": {\n    \"runtime\": " << quote("HIP " + std::to_string(rt)) << ",\n    \"device\": "
      << quote(dev) << "\n  },\n"
      << "  \"implementation_mode\": {\n"
      << "    \"cache_backend\": \"hip_hashgrid_mlp_mfma_gfx950\",\n"
      << "    \"python_bindings\": false,\n"
      << "    \"csv_postprocess\": false,\n"
      << "    \"training_batch_generation\": \"hip_kernel\",\n"
      << "    \"training_inputs_location\": \"hip_device_memory\",\n"
      << "    \"training_targets_location\": \"hip_device_memory\",\n"
      << "    \"inference_inputs_location\": \"hip_device_memory\",\n"
      << "    \"inference_outputs_location\": \"hip_device_memory\",\n"
      << "    \"downstream_consumer\": \"hip_kernel\",\n"
      << "    \"host_transfer_between_tcnn_and_consumer\": false,\n"
      << "    \"timing_scope\": \"hip_events; inference timing excludes final validation readback\"\n"
      << "  },\n"
      << "  \"options\": {\n"
      << "    \"samples\": " << opt.samples << ",\n"
      << "    \"samples_padded_to_tcnn_granularity\": " << opt.samples << ",\n"
      << "    \"batch_size\": " << opt.batch_size << ",\n"
      << "    \"batch_size_padded_to_tcnn_granularity\": " << pad_b << ",\n"
      << "    \"batch_granularity\": 16,\n"
      << "    \"train_steps\": " << opt.train_steps << ",\n"
      << "    \"repeat\": " << opt.repeat << ",\n"
      << "    \"warmup\": " << opt.warmup << ",\n"
      << "    \"seed\": " << opt.seed << ",\n"
      << "    \"target_function\": \"harmonic_x2_minus_y2_on_unit_cube_coordinates\",\n"
      << "    \"network\": " << quote(opt.network) << ",\n"
      << "    \"network_backend\": " << quote("hashgrid_relu_mlp_" + std::to_string(opt.n_hidden_layers) + "x" +
                                              std::to_string(opt.n_neurons) + "_fp16_mfma")
      << ",\n"
      << "    \"jit_requested\": " << (opt.jit ? "true" : "false") << ",\n"
      << "    \"jit_supported\": false,\n"
      << "    \"jit_enabled\": false,\n"
      << "    \"n_levels\": " << opt.n_levels << ",\n"
      << "    \"n_features_per_level\": " << opt.n_features_per_level << ",\n"
      << "    \"log2_hashmap_size\": " << opt.log2_hashmap_size << ",\n"
      << "    \"base_resolution\": " << opt.base_resolution << ",\n"
      << "    \"per_level_scale\": " << opt.per_level_scale << ",\n"
      << "    \"n_neurons\": " << opt.n_neurons << ",\n"
      << "    \"n_hidden_layers\": " << opt.n_hidden_layers << ",\n"
      << "    \"learning_rate\": " << opt.learning_rate << "\n"
      << "  },\n"
      << "  \"training\": {\n"
      << "    \"total_ms\": " << train_ms << ",\n"
      << "    \"ms_per_step\": " << (opt.train_steps > 0 ? train_ms / opt.train_steps : 0.0f) << ",\n"
      << "    \"includes_gpu_batch_generation\": true,\n"
      << "    \"host_readback_in_training_loop\": false\n"
      << "  },\n"
      << "  \"inference\": {\n"
      << "    \"kernel_ms\": " << vec(inf_ms) << ",\n"
      << "    \"kernel_ms_median\": " << med << ",\n"
      << "    \"kernel_ms_mean\": " << mean << ",\n"
      << "    \"median_us_per_sample\": " << us_per << ",\n"
      << "    \"median_msamples_per_second\": " << msps << "\n"
      << "  },\n"
      << "  \"inference_plus_downstream_consumer\": {\n"
      << "    \"kernel_ms\": " << vec(cons_ms) << ",\n"
      << "    \"kernel_ms_median\": " << cmed << ",\n"
      << "    \"kernel_ms_mean\": " << cmean << ",\n"
      << "    \"median_us_per_sample\": " << cus_per << ",\n"
      << "    \"host_transfer_between_stages\": false\n"
      << "  },\n"
      << "  \"validation\": {\n"
      << "    \"checked_repeated_samples\": " << static_cast<unsigned long long>(n) * opt.repeat << ",\n"
      << "    \"mean_prediction\": " << sp / denom << ",\n"
      << "    \"mean_exact\": " << se / denom << ",\n"
      << "    \"mse\": " << mse << ",\n"
      << "    \"rmse\": " << std::sqrt(mse) << ",\n"
      << "    \"mae\": " << sa / denom << "\n"
      << "  }\n}\n";
    o.close();
    for (void* p : {static_cast<void*>(train_pts), static_cast<void*>(train_t), static_cast<void*>(inf_pts),
                    static_cast<void*>(inf_out), static_cast<void*>(stats)})
      (void)hipFree(p);
    check(n2wos_cache_destroy(cache), "n2wos_cache_destroy");
    (void)hipStreamDestroy(s);
    std::cout << "wrote " << opt.output << "\n";
    std::cout << "median inference: " << us_per << " us/sample\n";
    std::cout << "RMSE after training: " << std::sqrt(mse) << "\n";
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
}
