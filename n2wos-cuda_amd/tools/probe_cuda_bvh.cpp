// n2wos_probe_cuda_bvh -- closest-point probe over the engine's C ABI.
//
// Command line, query generation, JSON schema and exit codes follow the
// reference probe (src/tools/probe_cuda_bvh.cpp): the same flags (:42-61,
// plus the benchmark mesh kinds sin6 / multi), the same std::mt19937 +
// uniform_real_distribution<float> query points for a given seed (:124-132),
// warm-up then --repeat timed launches of the device query, validation of
// the first --validate points against the host brute force with the 2e-4
// distance threshold (:143-173), JSON "n2wos_cuda_probe_bvh_v1" (:237-316)
// with truthful implementation_mode labels, and exit code 0 (valid, no stack
// overflow), 2 (validation or overflow failure) or 1 (error) (:393-397).
// Timing is kernel-only (HIP events around the query launch), as the
// reference's kernel_milliseconds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/n2wos_amd.h"

namespace {

struct Options {
  std::string mesh = "procedural_bumpy_sphere";
  std::string mesh_path;
  int normalize = 1;
  int bumpy_stacks = 128;
  int bumpy_slices = 256;
  float bumpy_amplitude = 0.15f;
  int leaf_size = 8;
  int queries = 262144;
  int validate = 512;
  int repeat = 10;
  int block_size = 128;
  unsigned int seed = 12345;
  float query_extent = 1.35f;
  std::string output = "results/probe_cuda_bvh.json";
};

void check(int rc, const char* what) {
  if (rc != N2WOS_OK) throw std::runtime_error(std::string(what) + ": " + n2wos_last_error());
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void usage(std::ostream& o, const char* argv0) {
  o << "Usage: " << argv0 << " [options]\n\nOptions:\n"
    << "  --mesh procedural_bumpy_sphere|sin6|multi|obj|ply\n"
    << "                                     sin6: r = 1 + a sin6t sin6p (a = --bumpy-amplitude);\n"
    << "                                     multi: r = 1 + .06 sin5t sin7p + .03 sin17t cos13p\n"
    << "  --mesh-path <path>                 Required when --mesh obj or ply\n"
    << "  --normalize 0|1                    Normalize a file mesh to centered unit radius [1]\n"
    << "  --bumpy-stacks <int>               Procedural sphere stacks [128]\n"
    << "  --bumpy-slices <int>               Procedural sphere slices [256]\n"
    << "  --bumpy-amplitude <float>          Procedural bump amplitude [0.15]\n"
    << "  --leaf-size <int>                  BVH leaf triangle count [8]\n"
    << "  --queries <int>                    Number of GPU query points [262144]\n"
    << "  --validate <int>                   Number of points checked by CPU brute force [512]\n"
    << "  --repeat <int>                     Repeated timed kernel launches [10]\n"
    << "  --block-size <int>                 Block size of the query launch [128]\n"
    << "  --seed <int>                       Host query generation seed [12345]\n"
    << "  --query-extent <float>             Uniform query cube half-width [1.35]\n"
    << "  --output <path>                    JSON output [results/probe_cuda_bvh.json]\n"
    << "  --help\n";
}

Options parse(int argc, char** argv) {
  Options o;
  auto val = [&](int& i) -> std::string {
    if (i + 1 >= argc) throw std::runtime_error(std::string("missing value for ") + argv[i]);
    return argv[++i];
  };
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--help" || a == "-h") {
      usage(std::cout, argv[0]);
      std::exit(0);
    } else if (a == "--mesh") o.mesh = val(i);
    else if (a == "--mesh-path") o.mesh_path = val(i);
    else if (a == "--normalize") o.normalize = std::stoi(val(i));
    else if (a == "--bumpy-stacks") o.bumpy_stacks = std::stoi(val(i));
    else if (a == "--bumpy-slices") o.bumpy_slices = std::stoi(val(i));
    else if (a == "--bumpy-amplitude") o.bumpy_amplitude = std::stof(val(i));
    else if (a == "--leaf-size") o.leaf_size = std::stoi(val(i));
    else if (a == "--queries") o.queries = std::stoi(val(i));
    else if (a == "--validate") o.validate = std::stoi(val(i));
    else if (a == "--repeat") o.repeat = std::stoi(val(i));
    else if (a == "--block-size") o.block_size = std::stoi(val(i));
    else if (a == "--seed") o.seed = static_cast<unsigned int>(std::stoul(val(i)));
    else if (a == "--query-extent") o.query_extent = std::stof(val(i));
    else if (a == "--output") o.output = val(i);
    else throw std::runtime_error("unknown option: " + a);
  }
  const bool file = o.mesh == "obj" || o.mesh == "ply";
  if (!file && o.mesh != "procedural_bumpy_sphere" && o.mesh != "sin6" && o.mesh != "multi")
    throw std::runtime_error("--mesh must be procedural_bumpy_sphere, sin6, multi, obj, or ply");
  if (file && o.mesh_path.empty()) throw std::runtime_error("--mesh-path is required when --mesh obj or ply");
  if (o.queries <= 0) throw std::runtime_error("--queries must be positive");
  if (o.validate < 0) throw std::runtime_error("--validate must be non-negative");
  if (o.repeat <= 0) throw std::runtime_error("--repeat must be positive");
  if (o.leaf_size <= 0) throw std::runtime_error("--leaf-size must be positive");
  if (!(o.query_extent > 0.0f)) throw std::runtime_error("--query-extent must be positive");
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

std::string basename(const std::string& p) {
  const size_t s = p.find_last_of("/\\");
  return s == std::string::npos ? p : p.substr(s + 1);
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const Options opt = parse(argc, argv);
    std::string cmd;
    for (int i = 0; i < argc; ++i) cmd += (i ? " " : "") + std::string(argv[i]);

    // ------------------------------------------------------------ mesh
    std::vector<float> verts;
    std::vector<uint32_t> tris;
    std::string mesh_name;
    bool normalized = false;
    float center[3] = {0, 0, 0}, scale = 1.0f;
    if (opt.mesh == "obj" || opt.mesh == "ply") {
      float* v = nullptr;
      uint32_t* t = nullptr;
      int nv = 0, nt = 0;
      check(n2wos_mesh_load(opt.mesh_path.c_str(), opt.mesh.c_str(), &v, &nv, &t, &nt), "mesh load");
      verts.assign(v, v + 3 * static_cast<size_t>(nv));
      tris.assign(t, t + 3 * static_cast<size_t>(nt));
      n2wos_mesh_free(v, t);
      mesh_name = basename(opt.mesh_path);
      normalized = opt.normalize != 0;
    } else {
      int nv = 0, nt = 0;
      check(n2wos_uv_sphere_counts(opt.bumpy_stacks, opt.bumpy_slices, &nv, &nt), "uv sphere");
      verts.resize(3 * static_cast<size_t>(nv));
      tris.resize(3 * static_cast<size_t>(nt));
      const int kind = opt.mesh == "sin6" ? N2WOS_SPHERE_SIN6
                       : opt.mesh == "multi" ? N2WOS_SPHERE_MULTI : N2WOS_SPHERE_REFERENCE_BUMP;
      const float amp = opt.bumpy_amplitude;
      check(n2wos_uv_sphere(opt.bumpy_stacks, opt.bumpy_slices, kind, &amp, 1, verts.data(), tris.data()),
            "uv sphere");
      mesh_name = "procedural_bumpy_sphere";
      if (opt.mesh != "procedural_bumpy_sphere") mesh_name += "_" + opt.mesh;
      normalized = true;
    }
    const int nv = static_cast<int>(verts.size() / 3), nt = static_cast<int>(tris.size() / 3);
    if (normalized) check(n2wos_normalize_to_unit_radius(verts.data(), nv, center, &scale), "normalize");
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < nv; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], verts[3 * i + a]);
        hi[a] = std::max(hi[a], verts[3 * i + a]);
      }
    const int64_t degenerate = n2wos_count_degenerate_triangles(verts.data(), nv, tris.data(), nt, 0.0f);
    std::cerr << "mesh: " << mesh_name << " vertices=" << nv << " triangles=" << nt << "\n";
    std::cerr << "building host BVH and copying to GPU...\n";
    n2wos_bvh* bvh = nullptr;
    check(n2wos_bvh_create(verts.data(), nv, tris.data(), nt, opt.leaf_size, "median", &bvh), "bvh create");
    n2wos_bvh_stats bs{};
    check(n2wos_bvh_get_stats(bvh, &bs), "bvh stats");
    std::cerr << "BVH: nodes=" << bs.node_count << " max_depth=" << bs.max_depth << "\n";

    // --------------------------------------------------------- queries
    std::mt19937 rng(opt.seed);
    std::uniform_real_distribution<float> uni(-opt.query_extent, opt.query_extent);
    const size_t n = static_cast<size_t>(opt.queries);
    std::vector<float> pts(3 * n);
    for (size_t i = 0; i < n; ++i) {
      const float x = uni(rng), y = uni(rng), z = uni(rng);
      pts[3 * i] = x;
      pts[3 * i + 1] = y;
      pts[3 * i + 2] = z;
    }
    float *d_pts = nullptr, *d_d2 = nullptr, *d_cp = nullptr;
    int *d_tri = nullptr, *d_ovf = nullptr;
    hip_check(hipMalloc(&d_pts, 12 * n), "hipMalloc");
    hip_check(hipMalloc(&d_d2, 4 * n), "hipMalloc");
    hip_check(hipMalloc(&d_cp, 12 * n), "hipMalloc");
    hip_check(hipMalloc(&d_tri, 4 * n), "hipMalloc");
    hip_check(hipMalloc(&d_ovf, 4 * n), "hipMalloc");
    hip_check(hipMemcpy(d_pts, pts.data(), 12 * n, hipMemcpyHostToDevice), "upload");
    hipStream_t st;
    hipEvent_t e0, e1;
    hip_check(hipStreamCreate(&st), "stream");
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    auto launch = [&]() {
      check(n2wos_bvh_query_device(bvh, d_pts, opt.queries, d_d2, d_cp, d_tri, d_ovf, opt.block_size, st), "query");
    };
    launch();  // warm-up
    hip_check(hipStreamSynchronize(st), "sync");
    std::vector<float> ms;
    std::vector<int> ovf(n);
    int64_t overflow_count = 0;
    for (int r = 0; r < opt.repeat; ++r) {
      hip_check(hipEventRecord(e0, st), "event");
      launch();
      hip_check(hipEventRecord(e1, st), "event");
      hip_check(hipEventSynchronize(e1), "event");
      float t = 0.0f;
      hip_check(hipEventElapsedTime(&t, e0, e1), "event");
      ms.push_back(t);
      hip_check(hipMemcpy(ovf.data(), d_ovf, 4 * n, hipMemcpyDeviceToHost), "download");
      overflow_count = 0;
      for (int v : ovf) overflow_count += v != 0;
      std::cerr << "repeat " << (r + 1) << "/" << opt.repeat << ": kernel_ms=" << t << " overflow=" << overflow_count
                << "\n";
    }
    std::vector<float> d2(n), cp(3 * n);
    std::vector<int> tri(n);
    hip_check(hipMemcpy(d2.data(), d_d2, 4 * n, hipMemcpyDeviceToHost), "download");
    hip_check(hipMemcpy(cp.data(), d_cp, 12 * n, hipMemcpyDeviceToHost), "download");
    hip_check(hipMemcpy(tri.data(), d_tri, 4 * n, hipMemcpyDeviceToHost), "download");

    // ------------------------------------------------------ validation
    const int checked = std::min(opt.validate, opt.queries);
    std::cerr << "validating " << checked << " points by CPU brute force...\n";
    std::vector<float> bd2(checked), bcp(3 * static_cast<size_t>(checked));
    std::vector<int> btri(checked);
    if (checked > 0)
      check(n2wos_closest_point_bruteforce(verts.data(), nv, tris.data(), nt, pts.data(), checked, bd2.data(),
                                           bcp.data(), btri.data(), 0),
            "brute force");
    int bad = 0, tri_mismatch = 0, exact = 0;
    float max_d = 0.0f, max_cp = 0.0f;
    double s_d = 0.0, s_cp = 0.0;
    for (int i = 0; i < checked; ++i) {
      const float cd = std::sqrt(std::max(bd2[i], 0.0f)), gd = std::sqrt(std::max(d2[i], 0.0f));
      const float ed = std::fabs(cd - gd);
      const float ex = bcp[3 * i] - cp[3 * i], ey = bcp[3 * i + 1] - cp[3 * i + 1], ez = bcp[3 * i + 2] - cp[3 * i + 2];
      const float ec = std::sqrt(ex * ex + ey * ey + ez * ez);
      max_d = std::max(max_d, ed);
      max_cp = std::max(max_cp, ec);
      s_d += static_cast<double>(ed) * ed;
      s_cp += static_cast<double>(ec) * ec;
      if (ed > 2.0e-4f) ++bad;
      if (btri[i] != tri[i]) ++tri_mismatch;
      if (bd2[i] == d2[i] && ec == 0.0f && btri[i] == tri[i]) ++exact;
    }
    const float rms_d = checked ? static_cast<float>(std::sqrt(s_d / checked)) : 0.0f;
    const float rms_cp = checked ? static_cast<float>(std::sqrt(s_cp / checked)) : 0.0f;

    // ------------------------------------------------------------ JSON
    char dev[256] = {0};
    n2wos_device_summary(dev, sizeof(dev));
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    const float med = median(ms);
    const float mean = std::accumulate(ms.begin(), ms.end(), 0.0f) / static_cast<float>(ms.size());
    const std::filesystem::path parent = std::filesystem::path(opt.output).parent_path();
    if (!parent.empty()) std::filesystem::create_directories(parent);
    std::ofstream o(opt.output);
    if (!o) throw std::runtime_error("failed to open output JSON: " + opt.output);
    o << std::setprecision(9);
    o << "{\n  \"schema\": \"n2wos_cuda_probe_bvh_v1\",\n";
    o << "  \"patch\": \"n2wos_amd_gfx950\",\n";
    o << "  \"generated_at_utc\": " << quote(utc_now()) << ",\n";
    o << "  \"command_line\": " << quote(cmd) << ",\n";
    o << "  \"cuda\": {\n    \"runtime\": " << quote("HIP " + std::to_string(rt)) << ",\n    \"device\": "
      << quote(dev) << "\n  },\n";
    o << "  \"implementation_mode\": {\n"
      << "    \"gpu_resident_geometry_query\": true,\n"
      << "    \"geometry_query\": \"hip_bvh_gfx950\",\n"
      << "    \"fcpw_gpu_public_api\": false,\n"
      << "    \"host_driven_walker_loop\": false,\n"
      << "    \"cache_inference\": \"none\",\n"
      << "    \"rng\": \"host_query_generation_only\",\n"
      << "    \"accumulation\": \"not_applicable\",\n"
      << "    \"timing_scope\": \"hip_kernel_only_excludes_h2d_d2h_and_allocation\"\n  },\n";
    o << "  \"options\": {\n    \"mesh\": " << quote(opt.mesh) << ",\n    \"mesh_path\": " << quote(opt.mesh_path)
      << ",\n    \"normalize\": " << opt.normalize << ",\n    \"bumpy_stacks\": " << opt.bumpy_stacks
      << ",\n    \"bumpy_slices\": " << opt.bumpy_slices << ",\n    \"bumpy_amplitude\": " << opt.bumpy_amplitude
      << ",\n    \"leaf_size\": " << opt.leaf_size << ",\n    \"queries\": " << opt.queries
      << ",\n    \"validate\": " << opt.validate << ",\n    \"repeat\": " << opt.repeat
      << ",\n    \"block_size\": " << opt.block_size << ",\n    \"seed\": " << opt.seed
      << ",\n    \"query_extent\": " << opt.query_extent << "\n  },\n";
    o << "  \"mesh_stats\": {\n    \"name\": " << quote(mesh_name) << ",\n    \"vertices\": " << nv
      << ",\n    \"triangles\": " << nt << ",\n    \"degenerate_triangles\": " << degenerate
      << ",\n    \"bounds_min\": [" << lo[0] << ", " << lo[1] << ", " << lo[2] << "],\n    \"bounds_max\": [" << hi[0]
      << ", " << hi[1] << ", " << hi[2] << "]";
    if (normalized)
      o << ",\n    \"normalization\": {\n      \"center\": [" << center[0] << ", " << center[1] << ", " << center[2]
        << "],\n      \"scale\": " << scale << "\n    }\n";
    else
      o << "\n";
    o << "  },\n";
    o << "  \"bvh_stats\": {\n    \"nodes\": " << bs.node_count << ",\n    \"triangle_indices\": " << bs.index_count
      << ",\n    \"leaf_size\": " << bs.leaf_size << ",\n    \"max_depth\": " << bs.max_depth
      << ",\n    \"binary_nodes\": " << bs.binary_nodes << ",\n    \"branching\": 4"
      << ",\n    \"build_milliseconds\": " << bs.build_milliseconds << ",\n    \"device_bytes\": " << bs.device_bytes
      << "\n  },\n";
    o << "  \"timing\": {\n    \"kernel_ms\": [";
    for (size_t i = 0; i < ms.size(); ++i) o << (i ? ", " : "") << ms[i];
    o << "],\n    \"kernel_ms_median\": " << med << ",\n    \"kernel_ms_mean\": " << mean
      << ",\n    \"median_us_per_query\": " << static_cast<double>(med) * 1000.0 / opt.queries
      << ",\n    \"mean_us_per_query\": " << static_cast<double>(mean) * 1000.0 / opt.queries
      << ",\n    \"median_mqueries_per_second\": " << opt.queries / (static_cast<double>(med) * 1000.0) << "\n  },\n";
    o << "  \"validation\": {\n    \"checked\": " << checked << ",\n    \"bad_distance_count_threshold_2e_4\": " << bad
      << ",\n    \"max_abs_distance_error\": " << max_d << ",\n    \"rms_distance_error\": " << rms_d
      << ",\n    \"max_closest_point_error\": " << max_cp << ",\n    \"rms_closest_point_error\": " << rms_cp
      << ",\n    \"triangle_id_mismatches\": " << tri_mismatch << ",\n    \"bit_exact_matches\": " << exact
      << ",\n    \"cuda_stack_overflow_count\": " << overflow_count << "\n  }\n}\n";
    o.close();

    std::cout << "wrote " << opt.output << "\n";
    std::cout << "median_us_per_query=" << static_cast<double>(med) * 1000.0 / opt.queries
              << " validation_max_abs_distance_error=" << max_d << " overflow=" << overflow_count << "\n";
    (void)hipFree(d_pts);
    (void)hipFree(d_d2);
    (void)hipFree(d_cp);
    (void)hipFree(d_tri);
    (void)hipFree(d_ovf);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    n2wos_bvh_destroy(bvh);
    return bad == 0 && overflow_count == 0 ? 0 : 2;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
}
