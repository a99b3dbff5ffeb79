// mesh.cpp -- procedural meshes, normalisation and misc host entry points.
//
// UV-sphere topology and vertex order follow the reference generator
// (src/mesh/procedural_bumpy_sphere.cpp:32-90): north pole, (stacks-1) rings
// of `slices` vertices, south pole; caps {N,a,b} / {a,S,b}, bands {a,c,b},
// {b,c,d}.  Positions are computed in FP32 exactly as the reference does
// (r * sin(theta) * cos(phi), ...), with the benchmark radius functions beside
// the reference bump_radius (:12-17).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "api_internal.hpp"

namespace n2 {

thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {
std::mutex g_tuning_mu;
std::map<std::string, int> g_tuning;
}  // namespace

int tuning_value(const char* key, int fallback) {
  std::lock_guard<std::mutex> lock(g_tuning_mu);
  const auto it = g_tuning.find(key);
  return it == g_tuning.end() ? fallback : it->second;
}

namespace {

constexpr float kPi = 3.14159265358979323846f;

// Real orthonormal spherical harmonic Y_lm (no Condon-Shortley phase),
// associated Legendre functions by the standard upward recurrence in l.
double sh_real(int l, int m, double theta, double phi) {
  const int am = m < 0 ? -m : m;
  const double x = std::cos(theta), s = std::sin(theta);
  double pmm = 1.0;
  for (int i = 1; i <= am; ++i) pmm *= (2.0 * i - 1.0) * s;
  double plm = pmm;
  if (l > am) {
    double a = pmm;
    double b = x * (2.0 * am + 1.0) * pmm;
    for (int ll = am + 2; ll <= l; ++ll) {
      const double c = ((2.0 * ll - 1.0) * x * b - (ll + am - 1.0) * a) / static_cast<double>(ll - am);
      a = b;
      b = c;
    }
    plm = b;
  }
  double ratio = 1.0;  // (l - |m|)! / (l + |m|)!
  for (int i = l - am + 1; i <= l + am; ++i) ratio /= static_cast<double>(i);
  const double n = std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * ratio);
  if (m == 0) return n * plm;
  if (m > 0) return std::sqrt(2.0) * n * plm * std::cos(am * phi);
  return std::sqrt(2.0) * n * plm * std::sin(am * phi);
}

float radius(int kind, const float* prm, float theta, float phi) {
  switch (kind) {
    case N2WOS_SPHERE_REFERENCE_BUMP: {
      const float amplitude = prm[0];
      const float low = std::sin(3.0f * theta + 0.35f * std::cos(5.0f * phi));
      const float high = std::sin(7.0f * phi + 0.5f * std::cos(2.0f * theta));
      const float mid = std::cos(5.0f * theta - 2.0f * phi);
      return 1.0f + amplitude * (0.45f * low * high + 0.35f * mid * std::sin(theta) + 0.20f * std::sin(4.0f * theta));
    }
    case N2WOS_SPHERE_SIN6:
      return 1.0f + prm[0] * std::sin(6.0f * theta) * std::sin(6.0f * phi);
    case N2WOS_SPHERE_MULTI:
      return 1.0f + 0.06f * std::sin(5.0f * theta) * std::sin(7.0f * phi) +
             0.03f * std::sin(17.0f * theta) * std::cos(13.0f * phi);
    default: {
      double r = 1.0;
      for (int l = 0; l <= 8; ++l)
        for (int m = -l; m <= l; ++m) r += static_cast<double>(prm[l * l + l + m]) * sh_real(l, m, theta, phi);
      return static_cast<float>(r);
    }
  }
}

void sphere_point(int kind, const float* prm, float theta, float phi, float* out) {
  const float r = radius(kind, prm, theta, phi);
  const float st = std::sin(theta);
  out[0] = r * st * std::cos(phi);
  out[1] = r * st * std::sin(phi);
  out[2] = r * std::cos(theta);
}

uint32_t ring(int i, int j, int slices) { return static_cast<uint32_t>(1 + (i - 1) * slices + (j % slices)); }

}  // namespace
}  // namespace n2

using namespace n2;

extern "C" const char* n2wos_last_error(void) { return g_last_error.c_str(); }

extern "C" int n2wos_abi_version(void) { return N2WOS_ABI_VERSION; }

extern "C" int n2wos_set_tuning(const char* key, int value) {
  return guarded([&]() -> int {
    N2_REQUIRE(key != nullptr, N2WOS_ERR_NULL, "NULL key");
    const std::string k(key);
    N2_REQUIRE(k == "ring" || k == "refill_min" || k == "weight_node" || k == "weight_tri" || k == "weight_update" ||
                   k == "query_quad" || k == "unit_ranges" || k == "spatial_order" || k == "query_cap" ||
                   k == "query_finish_slots",
               N2WOS_ERR_INVALID, "unknown tuning key: " + k);
    if (k == "query_cap") N2_REQUIRE(value >= 0, N2WOS_ERR_INVALID, "query_cap must be >= 0 (0: no second pass)");
    if (k == "query_finish_slots")
      N2_REQUIRE(value >= 1 && value <= 64, N2WOS_ERR_INVALID, "query_finish_slots must be in [1, 64]");
    if (k == "spatial_order") N2_REQUIRE(value == 0 || value == 1, N2WOS_ERR_INVALID, "spatial_order must be 0 or 1");
    if (k == "unit_ranges") N2_REQUIRE(value >= 1 && value <= 8, N2WOS_ERR_INVALID, "unit_ranges must be in [1, 8]");
    if (k == "ring") N2_REQUIRE(value == 8 || value == 16, N2WOS_ERR_INVALID, "ring must be 8 or 16");
    if (k == "query_quad") N2_REQUIRE(value == 0 || value == 1, N2WOS_ERR_INVALID, "query_quad must be 0 or 1");
    if (k == "refill_min") N2_REQUIRE(value >= 1 && value <= 64, N2WOS_ERR_INVALID, "refill_min must be in [1, 64]");
    if (k.rfind("weight_", 0) == 0) N2_REQUIRE(value >= 1 && value <= 1000, N2WOS_ERR_INVALID, "weight in [1, 1000]");
    std::lock_guard<std::mutex> lock(g_tuning_mu);
    g_tuning[k] = value;
    return N2WOS_OK;
  });
}

extern "C" int n2wos_device_summary(char* buf, int buf_len) {
  return guarded([&]() -> int {
    N2_REQUIRE(buf && buf_len > 0, N2WOS_ERR_NULL, "NULL buffer");
    int dev = 0, rt = 0, drv = 0;
    N2_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    N2_HIP(hipGetDeviceProperties(&prop, dev));
    N2_HIP(hipRuntimeGetVersion(&rt));
    N2_HIP(hipDriverGetVersion(&drv));
    std::snprintf(buf, static_cast<size_t>(buf_len), "%s %s cus=%d global_mem=%llu runtime=%d driver=%d",
                  prop.name, prop.gcnArchName, prop.multiProcessorCount,
                  static_cast<unsigned long long>(prop.totalGlobalMem), rt, drv);
    return N2WOS_OK;
  });
}

extern "C" int n2wos_uv_sphere_counts(int stacks, int slices, int* nv, int* nt) {
  return guarded([&]() -> int {
    N2_REQUIRE(nv && nt, N2WOS_ERR_NULL, "NULL output");
    N2_REQUIRE(stacks >= 4, N2WOS_ERR_INVALID, "bumpy sphere requires stacks >= 4");
    N2_REQUIRE(slices >= 8, N2WOS_ERR_INVALID, "bumpy sphere requires slices >= 8");
    N2_REQUIRE(static_cast<int64_t>(stacks) * slices < (1 << 26), N2WOS_ERR_INVALID, "sphere too large");
    *nv = 2 + (stacks - 1) * slices;
    *nt = 2 * slices * (stacks - 1);
    return N2WOS_OK;
  });
}

extern "C" int n2wos_uv_sphere(int stacks, int slices, int kind, const float* params, int n_params, float* verts,
                               uint32_t* tris) {
  return guarded([&]() -> int {
    int nv = 0, nt = 0;
    const int rc = n2wos_uv_sphere_counts(stacks, slices, &nv, &nt);
    if (rc != N2WOS_OK) return rc;
    N2_REQUIRE(verts && tris, N2WOS_ERR_NULL, "NULL output arrays");
    N2_REQUIRE(kind >= 0 && kind <= 3, N2WOS_ERR_INVALID, "unknown sphere kind");
    float prm[81] = {0.0f};
    const int need = kind == N2WOS_SPHERE_SH_BLOB ? 81 : (kind == N2WOS_SPHERE_MULTI ? 0 : 1);
    N2_REQUIRE(n_params >= need && (need == 0 || params), N2WOS_ERR_INVALID, "too few sphere parameters");
    for (int i = 0; i < need; ++i) prm[i] = params[i];
    if (kind == N2WOS_SPHERE_REFERENCE_BUMP)
      N2_REQUIRE(prm[0] >= 0.0f && prm[0] <= 0.5f, N2WOS_ERR_INVALID, "bump amplitude must be in [0, 0.5]");
    int v = 0;
    sphere_point(kind, prm, 0.0f, 0.0f, verts + 3 * v++);
    for (int i = 1; i <= stacks - 1; ++i) {
      const float theta = kPi * static_cast<float>(i) / static_cast<float>(stacks);
      for (int j = 0; j < slices; ++j) {
        const float phi = 2.0f * kPi * static_cast<float>(j) / static_cast<float>(slices);
        sphere_point(kind, prm, theta, phi, verts + 3 * v++);
      }
    }
    const uint32_t south = static_cast<uint32_t>(v);
    sphere_point(kind, prm, kPi, 0.0f, verts + 3 * v++);
    int t = 0;
    auto put = [&](uint32_t a, uint32_t b, uint32_t c) {
      tris[3 * t] = a; tris[3 * t + 1] = b; tris[3 * t + 2] = c; ++t;
    };
    for (int j = 0; j < slices; ++j) put(0u, ring(1, j, slices), ring(1, j + 1, slices));
    for (int i = 1; i < stacks - 1; ++i) {
      for (int j = 0; j < slices; ++j) {
        const uint32_t a = ring(i, j, slices), b = ring(i, j + 1, slices);
        const uint32_t c = ring(i + 1, j, slices), d = ring(i + 1, j + 1, slices);
        put(a, c, b);
        put(b, c, d);
      }
    }
    for (int j = 0; j < slices; ++j) put(ring(stacks - 1, j, slices), south, ring(stacks - 1, j + 1, slices));
    return N2WOS_OK;
  });
}

extern "C" int n2wos_normalize_to_unit_radius(float* verts, int nv, float* center, float* scale) {
  return guarded([&]() -> int {
    N2_REQUIRE(verts && center && scale, N2WOS_ERR_NULL, "NULL argument");
    N2_REQUIRE(nv > 0, N2WOS_ERR_EMPTY, "mesh has no vertices");
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < nv; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], verts[3 * i + a]);
        hi[a] = std::max(hi[a], verts[3 * i + a]);
      }
    float c[3];
    for (int a = 0; a < 3; ++a) c[a] = (lo[a] + hi[a]) * 0.5f;
    float r = 0.0f;
    for (int i = 0; i < nv; ++i) {
      const float dx = verts[3 * i] - c[0], dy = verts[3 * i + 1] - c[1], dz = verts[3 * i + 2] - c[2];
      r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    N2_REQUIRE(r > 0.0f, N2WOS_ERR_INVALID, "cannot normalize mesh with zero radius");
    for (int i = 0; i < nv; ++i)
      for (int a = 0; a < 3; ++a) verts[3 * i + a] = (verts[3 * i + a] - c[a]) / r;
    for (int a = 0; a < 3; ++a) center[a] = c[a];
    *scale = 1.0f / r;
    return N2WOS_OK;
  });
}

extern "C" int64_t n2wos_count_degenerate_triangles(const float* verts, int nv, const uint32_t* tris, int nt,
                                                    float area2_eps) {
  if (!verts || !tris || nv <= 0 || nt <= 0) return N2WOS_ERR_INVALID;
  int64_t count = 0;
  for (int i = 0; i < nt; ++i) {
    const float* a = verts + 3 * static_cast<size_t>(tris[3 * i]);
    const float* b = verts + 3 * static_cast<size_t>(tris[3 * i + 1]);
    const float* c = verts + 3 * static_cast<size_t>(tris[3 * i + 2]);
    const float ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    const float vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    const float cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
    if (cx * cx + cy * cy + cz * cz <= area2_eps) ++count;
  }
  return count;
}
