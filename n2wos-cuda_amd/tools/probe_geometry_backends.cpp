// n2wos_probe_geometry_backends -- builder / leaf-size sweep of the engine's
// closest-point query over the C ABI.
//
// The counterpart of the reference's backend probe
// (src/tools/probe_geometry_backends.cu): the same flags (:156-181, with this
// engine's builders in place of cuBQL's), the same five query distributions
// from the same std::mt19937 streams (:321-412: uniform box, interior z-slice,
// surface points, near-boundary shell with log-uniform signed normal offsets,
// and WoS-like prefixes = shell points with every 4th replaced by a slice
// point), one device-resident query buffer shared by every backend, warm-up
// then --repeat timed launches (HIP events, kernel only), validation of the
// first --validate points against the host brute force with the 2e-4
// distance threshold (:577-615), and the "n2wos_geometry_backend_probe_v2"
// JSON with its backend_policy labels and a recommendation: the fastest
// validated production candidate (:720-743, :768-775).  Exit code 0 when
// every built backend validates without stack overflow, 2 otherwise, 1 on
// error or when nothing was built (:985-1002).
//
// Backends: hip_bvh4_<builder>_leaf<N> for builder in {median, sah} (the
// host builders of n2wos_bvh_create: object-median split, binned surface-area
// heuristic) and each leaf size, each queried with the four-lanes-per-query
// kernel ("quad", the default) and/or one lane per query ("lane").
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
#include <limits>
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
  std::vector<int> leaf_sizes = {8};
  bool leaf_sizes_explicit = false;
  std::vector<std::string> build_methods = {"median", "sah"};
  std::vector<std::string> query_kernels = {"quad"};
  int queries = 262144;
  int validate = 2048;
  int repeat = 10;
  int block_size = 128;
  unsigned int seed = 12345;
  float query_extent = 1.35f;
  std::string query_mode = "uniform_box";
  float shell_min = 1.0e-4f;
  float shell_max = 2.5e-1f;
  float plane_z = 0.0f;
  std::string output = "results/probe_geometry_backends.json";
};

void check(int rc, const char* what) {
  if (rc != N2WOS_OK) throw std::runtime_error(std::string(what) + ": " + n2wos_last_error());
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

std::string trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\n\r");
  if (a == std::string::npos) return std::string();
  return s.substr(a, s.find_last_not_of(" \t\n\r") - a + 1);
}

std::vector<std::string> split_csv(const std::string& text) {
  std::vector<std::string> out;
  std::stringstream ss(text);
  std::string item;
  while (std::getline(ss, item, ',')) {
    item = trim(item);
    if (!item.empty()) out.push_back(item);
  }
  return out;
}

// The reference's builder names map to this engine's builders: its in-tree
// median split and cuBQL's spatial median to the object-median split, SAH to
// the binned SAH.  cuBQL's radix and ELH builders have no counterpart here.
std::string canonical_builder(std::string m) {
  m = trim(m);
  if (m == "spatial_median" || m == "sm" || m == "spatial-median" || m == "custom" || m == "custom_cuda_bvh")
    return "median";
  if (m == "surface_area_heuristic" || m == "cubql_cuda_sah") return "sah";
  if (m != "median" && m != "sah")
    throw std::runtime_error("build method must be median (spatial_median) or sah; " + m +
                             " is a cuBQL builder this engine does not have");
  return m;
}

std::string canonical_query_mode(std::string m) {
  m = trim(m);
  if (m == "plane_slice") return "interior_slice";
  if (m == "near_surface_shell") return "near_boundary_shell";
  if (m == "mixed_wos_like") return "wos_like_prefix";
  if (m != "uniform_box" && m != "interior_slice" && m != "surface_exact" && m != "near_boundary_shell" &&
      m != "wos_like_prefix")
    throw std::runtime_error(
        "--query-mode must be uniform_box, interior_slice, surface_exact, near_boundary_shell, or wos_like_prefix");
  return m;
}

void usage(std::ostream& o, const char* argv0) {
  o << "Usage: " << argv0 << " [options]\n\nOptions:\n"
    << "  --mesh procedural_bumpy_sphere|sin6|multi|obj|ply\n"
    << "  --mesh-path <path>                 Required when --mesh obj or ply\n"
    << "  --normalize 0|1                    Normalize a file mesh to centered unit radius [1]\n"
    << "  --bumpy-stacks <int>               Procedural sphere stacks [128]\n"
    << "  --bumpy-slices <int>               Procedural sphere slices [256]\n"
    << "  --bumpy-amplitude <float>          Procedural bump amplitude [0.15]\n"
    << "  --leaf-size <int>                  Leaf triangle count of every backend [8]\n"
    << "  --leaf-sizes <csv>                 Leaf sizes to sweep, e.g. 1,2,4,8 (alias --cubql-leaf-sizes)\n"
    << "  --build-methods <csv>              median,sah or sweep/all [median,sah] (alias --cubql-build-methods;\n"
    << "                                     spatial_median = median; radix and elh are cuBQL-only)\n"
    << "  --query-kernels <csv>              quad (four lanes per query), lane (one lane per query) [quad]\n"
    << "  --queries <int>                    Number of device-resident query points [262144]\n"
    << "  --validate <int>                   Number of points checked by CPU brute force [2048]\n"
    << "  --repeat <int>                     Repeated timed device-query launches [10]\n"
    << "  --block-size <int>                 Block size of the query launch [128]\n"
    << "  --seed <int>                       Host query generation seed [12345]\n"
    << "  --query-extent <float>             Uniform cube/slice half-width [1.35]\n"
    << "  --query-mode <mode>                uniform_box|interior_slice|surface_exact|near_boundary_shell|"
       "wos_like_prefix [uniform_box]\n"
    << "  --shell-min <float>                near_boundary_shell minimum offset [1e-4]\n"
    << "  --shell-max <float>                near_boundary_shell maximum offset [0.25]\n"
    << "  --plane-z <float>                  z coordinate for interior_slice [0]\n"
    << "  --output <path>                    JSON output [results/probe_geometry_backends.json]\n"
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
    else if (a == "--leaf-size") {
      o.leaf_size = std::stoi(val(i));
      if (!o.leaf_sizes_explicit) o.leaf_sizes = {o.leaf_size};
    } else if (a == "--leaf-sizes" || a == "--cubql-leaf-sizes") {
      o.leaf_sizes.clear();
      for (const std::string& s : split_csv(val(i))) o.leaf_sizes.push_back(std::stoi(s));
      o.leaf_sizes_explicit = true;
    } else if (a == "--build-methods" || a == "--cubql-build-methods" || a == "--cubql-build-method") {
      const std::string v = trim(val(i));
      o.build_methods.clear();
      if (v == "sweep" || v == "all") {
        o.build_methods = {"median", "sah"};
      } else {
        for (const std::string& m : split_csv(v)) {
          const std::string c = canonical_builder(m);
          if (std::find(o.build_methods.begin(), o.build_methods.end(), c) == o.build_methods.end())
            o.build_methods.push_back(c);
        }
      }
    } else if (a == "--query-kernels") {
      o.query_kernels = split_csv(val(i));
    } else if (a == "--queries") o.queries = std::stoi(val(i));
    else if (a == "--validate") o.validate = std::stoi(val(i));
    else if (a == "--repeat") o.repeat = std::stoi(val(i));
    else if (a == "--block-size") o.block_size = std::stoi(val(i));
    else if (a == "--seed") o.seed = static_cast<unsigned int>(std::stoul(val(i)));
    else if (a == "--query-extent") o.query_extent = std::stof(val(i));
    else if (a == "--query-mode") o.query_mode = canonical_query_mode(val(i));
    else if (a == "--shell-min") o.shell_min = std::stof(val(i));
    else if (a == "--shell-max") o.shell_max = std::stof(val(i));
    else if (a == "--plane-z") o.plane_z = std::stof(val(i));
    else if (a == "--output") o.output = val(i);
    else throw std::runtime_error("unknown option: " + a);
  }
  const bool file = o.mesh == "obj" || o.mesh == "ply";
  if (!file && o.mesh != "procedural_bumpy_sphere" && o.mesh != "sin6" && o.mesh != "multi")
    throw std::runtime_error("--mesh must be procedural_bumpy_sphere, sin6, multi, obj, or ply");
  if (file && o.mesh_path.empty()) throw std::runtime_error("--mesh-path is required when --mesh obj or ply");
  if (o.build_methods.empty()) throw std::runtime_error("--build-methods cannot be empty");
  if (o.leaf_sizes.empty()) throw std::runtime_error("--leaf-sizes cannot be empty");
  for (int l : o.leaf_sizes)
    if (l <= 0 || l > 31) throw std::runtime_error("leaf sizes must be in [1, 31]");
  if (o.query_kernels.empty()) throw std::runtime_error("--query-kernels cannot be empty");
  for (const std::string& k : o.query_kernels)
    if (k != "quad" && k != "lane") throw std::runtime_error("--query-kernels entries must be quad or lane");
  if (o.queries <= 0) throw std::runtime_error("--queries must be positive");
  if (o.validate < 0) throw std::runtime_error("--validate must be non-negative");
  if (o.repeat <= 0) throw std::runtime_error("--repeat must be positive");
  if (!(o.query_extent > 0.0f)) throw std::runtime_error("--query-extent must be positive");
  if (!(o.shell_min >= 0.0f) || !(o.shell_max >= o.shell_min))
    throw std::runtime_error("require 0 <= --shell-min <= --shell-max");
  return o;
}

// ---------------------------------------------------------------- queries
struct F3 {
  float x, y, z;
};
F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float length(F3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

struct MeshView {
  const std::vector<float>& v;
  const std::vector<uint32_t>& t;
  F3 vert(uint32_t i) const { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }
};

std::vector<F3> uniform_box(int n, unsigned seed, float ext) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(-ext, ext);
  std::vector<F3> p(static_cast<size_t>(n));
  for (F3& q : p) {
    const float x = u(rng), y = u(rng), z = u(rng);
    q = {x, y, z};
  }
  return p;
}

std::vector<F3> plane_slice(int n, unsigned seed, float ext, float z) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(-ext, ext);
  std::vector<F3> p(static_cast<size_t>(n));
  for (F3& q : p) {
    const float x = u(rng), y = u(rng);
    q = {x, y, z};
  }
  return p;
}

// Area-weighted surface points (uniform in each triangle), moved along the
// face normal by a log-uniform offset in [shell_min, shell_max] of random sign.
std::vector<F3> surface_shell(const MeshView& m, int n, unsigned seed, float smin, float smax, bool zero) {
  std::mt19937 rng(seed);
  const size_t nt = m.t.size() / 3;
  std::vector<double> w(nt);
  for (size_t k = 0; k < nt; ++k) {
    const F3 a = m.vert(m.t[3 * k]), b = m.vert(m.t[3 * k + 1]), c = m.vert(m.t[3 * k + 2]);
    w[k] = std::max<double>(0.5f * length(cross(b - a, c - a)), 1.0e-30);
  }
  std::discrete_distribution<size_t> pick(w.begin(), w.end());
  std::uniform_real_distribution<float> unit(0.0f, 1.0f);
  const float minp = std::max(smin, 1.0e-8f);
  const bool logs = smax > minp;
  std::vector<F3> p(static_cast<size_t>(n));
  for (F3& q : p) {
    const size_t k = pick(rng);
    const F3 a = m.vert(m.t[3 * k]), b = m.vert(m.t[3 * k + 1]), c = m.vert(m.t[3 * k + 2]);
    const float su = std::sqrt(unit(rng));
    const float v = unit(rng);
    const F3 s = a * (1.0f - su) + b * (su * (1.0f - v)) + c * (su * v);
    F3 nrm = cross(b - a, c - a);
    const float len = length(nrm);
    nrm = len > 0.0f ? nrm * (1.0f / len) : F3{1.0f, 0.0f, 0.0f};
    float off = 0.0f;
    if (!zero) {
      off = logs ? minp * std::exp(std::log(smax / minp) * unit(rng)) : smax;
      if (unit(rng) < 0.5f) off = -off;
    }
    q = s + nrm * off;
  }
  return p;
}

std::vector<F3> make_queries(const Options& o, const MeshView& m) {
  if (o.query_mode == "uniform_box") return uniform_box(o.queries, o.seed, o.query_extent);
  if (o.query_mode == "interior_slice") return plane_slice(o.queries, o.seed, o.query_extent, o.plane_z);
  if (o.query_mode == "surface_exact") return surface_shell(m, o.queries, o.seed, 0.0f, 0.0f, true);
  if (o.query_mode == "near_boundary_shell")
    return surface_shell(m, o.queries, o.seed, o.shell_min, o.shell_max, false);
  std::vector<F3> sh = surface_shell(m, o.queries, o.seed, o.shell_min, o.shell_max, false);
  const std::vector<F3> sl = plane_slice(o.queries, o.seed ^ 0x9e3779b9u, o.query_extent, o.plane_z);
  for (int i = 0; i < o.queries; i += 4) sh[static_cast<size_t>(i)] = sl[static_cast<size_t>(i)];
  return sh;
}

// ------------------------------------------------------------------- JSON
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

struct Validation {
  int checked = 0, bad = 0, tri_mismatch = 0, exact = 0;
  float max_d = 0.0f, rms_d = 0.0f, max_cp = 0.0f, rms_cp = 0.0f, max_delta = 0.0f;
};

struct Run {
  std::string name, builder, kernel, error;
  bool built = false;
  int leaf = 0;
  n2wos_bvh_stats stats{};
  std::vector<float> ms;
  int64_t overflow = 0;
  std::vector<float> d2;
  Validation val;
  bool passed() const { return built && val.bad == 0 && overflow == 0 && !ms.empty(); }
  double us_per_query(int q) const {
    return ms.empty() ? std::numeric_limits<double>::infinity() : static_cast<double>(median(ms)) * 1000.0 / q;
  }
};

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
      const size_t s = opt.mesh_path.find_last_of("/\\");
      mesh_name = s == std::string::npos ? opt.mesh_path : opt.mesh_path.substr(s + 1);
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
    const MeshView mv{verts, tris};
    const std::vector<F3> q = make_queries(opt, mv);
    const size_t n = q.size();
    std::vector<float> pts(3 * n);
    for (size_t i = 0; i < n; ++i) {
      pts[3 * i] = q[i].x;
      pts[3 * i + 1] = q[i].y;
      pts[3 * i + 2] = q[i].z;
    }
    std::cerr << "mesh: " << mesh_name << " vertices=" << nv << " triangles=" << nt << "\n";
    std::cerr << "query_mode=" << opt.query_mode
              << " query points uploaded once; timed backend queries do not include H2D/D2H transfers\n";

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

    // host brute force of the validated prefix, shared by every backend
    const int checked = std::min(opt.validate, opt.queries);
    std::vector<float> bd2(static_cast<size_t>(checked)), bcp(3 * static_cast<size_t>(checked));
    std::vector<int> btri(static_cast<size_t>(checked));
    if (checked > 0) {
      std::cerr << "CPU brute force of " << checked << " validation points...\n";
      check(n2wos_closest_point_bruteforce(verts.data(), nv, tris.data(), nt, pts.data(), checked, bd2.data(),
                                           bcp.data(), btri.data(), 0),
            "brute force");
    }

    std::vector<Run> runs;
    std::vector<float> comparison;  // distances of the first built backend
    for (int leaf : opt.leaf_sizes) {
      for (const std::string& builder : opt.build_methods) {
        n2wos_bvh* bvh = nullptr;
        Run base;
        base.builder = builder == "sah" ? "host_binned_sah_collapsed_to_bvh4" : "host_object_median_split_bvh4";
        base.leaf = leaf;
        const int rc = n2wos_bvh_create(verts.data(), nv, tris.data(), nt, leaf, builder.c_str(), &bvh);
        for (const std::string& kernel : opt.query_kernels) {
          Run run = base;
          run.kernel = kernel;
          run.name = "hip_bvh4_" + builder + "_leaf" + std::to_string(leaf) + (kernel == "quad" ? "" : "_lane");
          if (rc != N2WOS_OK) {
            run.error = n2wos_last_error();
            runs.push_back(run);
            std::cerr << "backend failed: " << run.name << ": " << run.error << "\n";
            continue;
          }
          run.built = true;
          check(n2wos_bvh_get_stats(bvh, &run.stats), "bvh stats");
          check(n2wos_set_tuning("query_quad", kernel == "quad" ? 1 : 0), "tuning");
          auto launch = [&]() {
            check(n2wos_bvh_query_device(bvh, d_pts, opt.queries, d_d2, d_cp, d_tri, d_ovf, opt.block_size, st),
                  "query");
          };
          launch();  // warm-up
          hip_check(hipStreamSynchronize(st), "sync");
          for (int r = 0; r < opt.repeat; ++r) {
            hip_check(hipEventRecord(e0, st), "event");
            launch();
            hip_check(hipEventRecord(e1, st), "event");
            hip_check(hipEventSynchronize(e1), "event");
            float t = 0.0f;
            hip_check(hipEventElapsedTime(&t, e0, e1), "event");
            run.ms.push_back(t);
            std::cerr << run.name << " repeat " << (r + 1) << "/" << opt.repeat << ": kernel_ms=" << t << "\n";
          }
          std::vector<int> ovf(n), tri(n);
          std::vector<float> cp(3 * n);
          run.d2.resize(n);
          hip_check(hipMemcpy(run.d2.data(), d_d2, 4 * n, hipMemcpyDeviceToHost), "download");
          hip_check(hipMemcpy(cp.data(), d_cp, 12 * n, hipMemcpyDeviceToHost), "download");
          hip_check(hipMemcpy(tri.data(), d_tri, 4 * n, hipMemcpyDeviceToHost), "download");
          hip_check(hipMemcpy(ovf.data(), d_ovf, 4 * n, hipMemcpyDeviceToHost), "download");
          for (int v : ovf) run.overflow += v != 0;
          Validation& V = run.val;
          V.checked = checked;
          double sd = 0.0, sc = 0.0;
          for (int i = 0; i < checked; ++i) {
            const float cd = std::sqrt(std::max(bd2[i], 0.0f)), gd = std::sqrt(std::max(run.d2[i], 0.0f));
            const float ed = std::fabs(cd - gd);
            const float ex = bcp[3 * i] - cp[3 * i], ey = bcp[3 * i + 1] - cp[3 * i + 1],
                        ez = bcp[3 * i + 2] - cp[3 * i + 2];
            const float ec = std::sqrt(ex * ex + ey * ey + ez * ez);
            V.max_d = std::max(V.max_d, ed);
            V.max_cp = std::max(V.max_cp, ec);
            sd += static_cast<double>(ed) * ed;
            sc += static_cast<double>(ec) * ec;
            if (ed > 2.0e-4f) ++V.bad;
            if (btri[i] != tri[i]) ++V.tri_mismatch;
            if (bd2[i] == run.d2[i] && ec == 0.0f && btri[i] == tri[i]) ++V.exact;
            if (!comparison.empty()) {
              const float od = std::sqrt(std::max(comparison[static_cast<size_t>(i)], 0.0f));
              V.max_delta = std::max(V.max_delta, std::fabs(od - gd));
            }
          }
          if (checked > 0) {
            V.rms_d = static_cast<float>(std::sqrt(sd / checked));
            V.rms_cp = static_cast<float>(std::sqrt(sc / checked));
          }
          if (comparison.empty()) comparison = run.d2;
          runs.push_back(std::move(run));
        }
        if (bvh) n2wos_bvh_destroy(bvh);
      }
    }
    check(n2wos_set_tuning("query_quad", 1), "tuning");

    // ------------------------------------------------------------ JSON
    const Run* best = nullptr;
    for (const Run& r : runs)
      if (r.passed() && (!best || r.us_per_query(opt.queries) < best->us_per_query(opt.queries))) best = &r;
    char dev[256] = {0};
    n2wos_device_summary(dev, sizeof(dev));
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    const std::filesystem::path parent = std::filesystem::path(opt.output).parent_path();
    if (!parent.empty()) std::filesystem::create_directories(parent);
    std::ofstream o(opt.output);
    if (!o) throw std::runtime_error("failed to open output JSON: " + opt.output);
    auto ints = [&](const std::vector<int>& v) {
      std::string s = "[";
      for (size_t i = 0; i < v.size(); ++i) s += (i ? ", " : "") + std::to_string(v[i]);
      return s + "]";
    };
    auto strs = [&](const std::vector<std::string>& v) {
      std::string s = "[";
      for (size_t i = 0; i < v.size(); ++i) s += (i ? ", " : "") + quote(v[i]);
      return s + "]";
    };
    o << std::setprecision(9);
    o << "{\n  \"schema\": \"n2wos_geometry_backend_probe_v2\",\n";
    o << "  \"patch\": \"n2wos_amd_gfx950_builder_sweep\",\n";
    o << "  \"generated_at_utc\": " << quote(utc_now()) << ",\n";
    o << "  \"command_line\": " << quote(cmd) << ",\n";
    o << "  \"cuda\": {\n    \"runtime\": " << quote("HIP " + std::to_string(rt)) << ",\n    \"device\": " << quote(dev)
      << "\n  },\n";
    o << "  \"backend_policy\": {\n"
      << "    \"production_requires_gpu_resident_geometry_query\": true,\n"
      << "    \"production_forbids_per_step_host_transfer\": true,\n"
      << "    \"production_forbids_fcpw_public_host_vector_loop\": true,\n"
      << "    \"probe_timing_scope\": \"hip_kernel_only_excludes_query_upload_result_readback; build time is "
         "separately reported\",\n"
      << "    \"hip_bvh4_role\": \"production_candidate_when_it_validates\",\n"
      << "    \"cubql_cuda_role\": \"not_available_on_gfx950_replaced_by_hip_bvh4_builders\"\n  },\n";
    o << "  \"options\": {\n    \"mesh\": " << quote(opt.mesh) << ",\n    \"mesh_path\": " << quote(opt.mesh_path)
      << ",\n    \"normalize\": " << opt.normalize << ",\n    \"bumpy_stacks\": " << opt.bumpy_stacks
      << ",\n    \"bumpy_slices\": " << opt.bumpy_slices << ",\n    \"bumpy_amplitude\": " << opt.bumpy_amplitude
      << ",\n    \"leaf_size\": " << opt.leaf_size << ",\n    \"leaf_sizes\": " << ints(opt.leaf_sizes)
      << ",\n    \"queries\": " << opt.queries << ",\n    \"validate\": " << opt.validate
      << ",\n    \"repeat\": " << opt.repeat << ",\n    \"block_size\": " << opt.block_size
      << ",\n    \"seed\": " << opt.seed << ",\n    \"query_extent\": " << opt.query_extent
      << ",\n    \"query_mode\": " << quote(opt.query_mode) << ",\n    \"shell_min\": " << opt.shell_min
      << ",\n    \"shell_max\": " << opt.shell_max << ",\n    \"plane_z\": " << opt.plane_z
      << ",\n    \"build_methods\": " << strs(opt.build_methods) << ",\n    \"query_kernels\": "
      << strs(opt.query_kernels) << "\n  },\n";
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
    o << "  \"build_features\": {\n    \"N2WOS_HAS_CUBQL\": false,\n    \"hip_bvh4\": true\n  },\n";
    o << "  \"recommendation\": {\n"
      << "    \"selection_policy\": \"choose the fastest validated device-resident production candidate for the "
         "tested mesh/query mode\",\n"
      << "    \"preferred_sweep_order\": [\"sah\", \"median\"],\n"
      << "    \"default_before_measurement\": \"sah for static meshes when build cost is amortized; median as the "
         "fast builder\",\n";
    if (best)
      o << "    \"selected_production_candidate\": " << quote(best->name)
        << ",\n    \"selected_median_us_per_query\": " << best->us_per_query(opt.queries)
        << ",\n    \"fastest_validated_overall\": " << quote(best->name)
        << ",\n    \"fastest_validated_overall_median_us_per_query\": " << best->us_per_query(opt.queries) << "\n";
    else
      o << "    \"selected_production_candidate\": null,\n    \"selected_median_us_per_query\": null,\n"
        << "    \"fastest_validated_overall\": null,\n    \"fastest_validated_overall_median_us_per_query\": null\n";
    o << "  },\n  \"backends\": {\n";
    for (size_t k = 0; k < runs.size(); ++k) {
      const Run& r = runs[k];
      o << "    " << quote(r.name) << ": {\n      \"enabled\": true,\n      \"built\": " << (r.built ? "true" : "false")
        << ",\n      \"builder\": " << quote(r.builder) << ",\n      \"query_kernel\": "
        << quote(r.kernel == "quad" ? "four_lanes_per_query" : "one_lane_per_query") << ",\n";
      if (!r.error.empty()) o << "      \"error\": " << quote(r.error) << ",\n";
      o << "      \"production_candidate\": " << (r.built ? "true" : "false")
        << ",\n      \"device_resident_query_api\": true,\n      \"host_transfer_in_timing\": false,\n";
      o << "      \"bvh_stats\": {\n        \"triangles\": " << r.stats.triangle_count << ",\n        \"nodes\": "
        << r.stats.node_count << ",\n        \"prim_ids\": " << r.stats.index_count << ",\n        \"leaf_size\": "
        << r.leaf << ",\n        \"max_depth\": " << r.stats.max_depth << ",\n        \"branching\": 4"
        << ",\n        \"build_milliseconds\": " << r.stats.build_milliseconds << "\n      },\n";
      const float med = median(r.ms);
      const float mean = r.ms.empty() ? 0.0f : std::accumulate(r.ms.begin(), r.ms.end(), 0.0f) / r.ms.size();
      o << "      \"timing\": {\n        \"kernel_ms\": [";
      for (size_t i = 0; i < r.ms.size(); ++i) o << (i ? ", " : "") << r.ms[i];
      o << "],\n        \"kernel_ms_median\": " << med << ",\n        \"kernel_ms_mean\": " << mean
        << ",\n        \"median_us_per_query\": " << (r.ms.empty() ? 0.0 : r.us_per_query(opt.queries))
        << ",\n        \"mean_us_per_query\": " << static_cast<double>(mean) * 1000.0 / opt.queries
        << ",\n        \"median_mqueries_per_second\": "
        << (r.ms.empty() ? 0.0 : opt.queries / (static_cast<double>(med) * 1000.0)) << "\n      },\n";
      const Validation& V = r.val;
      o << "      \"validation\": {\n        \"checked\": " << V.checked
        << ",\n        \"bad_distance_count_threshold_2e_4\": " << V.bad << ",\n        \"max_abs_distance_error\": "
        << V.max_d << ",\n        \"rms_distance_error\": " << V.rms_d << ",\n        \"max_closest_point_error\": "
        << V.max_cp << ",\n        \"rms_closest_point_error\": " << V.rms_cp
        << ",\n        \"max_backend_distance_delta\": " << V.max_delta
        << ",\n        \"triangle_id_disagreement_count\": " << V.tri_mismatch
        << ",\n        \"bit_exact_matches\": " << V.exact << "\n      },\n";
      o << "      \"validation_passed\": " << (r.passed() ? "true" : "false") << ",\n      \"overflow_count\": "
        << r.overflow << "\n    },\n";
    }
    o << "    \"cubql_cuda\": {\n      \"enabled\": false,\n      \"built\": false,\n      \"reason\": "
      << quote("cuBQL is a CUDA library; on gfx950 the engine's own BVH4 builders take its place")
      << ",\n      \"production_candidate\": false,\n      \"device_resident_query_api\": true,\n"
      << "      \"host_transfer_in_timing\": false\n    }\n  }\n}\n";
    o.close();

    int built = 0, bad = 0;
    int64_t overflow = 0;
    for (const Run& r : runs) {
      if (!r.built) continue;
      ++built;
      bad += r.val.bad;
      overflow += r.overflow;
      std::cout << r.name << " median_us_per_query=" << r.us_per_query(opt.queries)
                << " build_ms=" << r.stats.build_milliseconds
                << " validation_max_abs_distance_error=" << r.val.max_d << " overflow=" << r.overflow << "\n";
    }
    if (best)
      std::cout << "selected_production_candidate=" << best->name
                << " median_us_per_query=" << best->us_per_query(opt.queries) << "\n";
    std::cout << "wrote " << opt.output << "\n";
    (void)hipFree(d_pts);
    (void)hipFree(d_d2);
    (void)hipFree(d_cp);
    (void)hipFree(d_tri);
    (void)hipFree(d_ovf);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    if (built == 0) return 1;
    return bad == 0 && overflow == 0 ? 0 : 2;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
}
