// ref_shim.cpp -- C entry points over the reference's own CPU geometry core.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
// reference sources where they lie (src/mesh/*.cpp and
// src/geometry/closest_point.cpp of the reference repository) into
// oracle/_ref/libn2wos_ref.so, which the tests use to pin the oracle's
// restatement.  Nothing here is part of the product, and the library is only
// built when the reference tree is present.
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "n2wos/closest_point.hpp"
#include "n2wos/mesh.hpp"
#include "n2wos/types.hpp"

namespace {

n2wos::Mesh make_mesh(const float* verts, int nv, const uint32_t* tris, int nt) {
  n2wos::Mesh m;
  m.vertices.resize(static_cast<std::size_t>(nv));
  for (int i = 0; i < nv; ++i) m.vertices[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
  m.triangles.resize(static_cast<std::size_t>(nt));
  for (int i = 0; i < nt; ++i) m.triangles[i] = {tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]};
  return m;
}

void export_mesh(const n2wos::Mesh& m, float* verts, uint32_t* tris) {
  for (std::size_t i = 0; i < m.vertices.size(); ++i) {
    verts[3 * i] = m.vertices[i].x;
    verts[3 * i + 1] = m.vertices[i].y;
    verts[3 * i + 2] = m.vertices[i].z;
  }
  for (std::size_t i = 0; i < m.triangles.size(); ++i) {
    tris[3 * i] = m.triangles[i].v0;
    tris[3 * i + 1] = m.triangles[i].v1;
    tris[3 * i + 2] = m.triangles[i].v2;
  }
}

}  // namespace

extern "C" {

// make_procedural_bumpy_sphere (+ normalize_to_unit_radius when normalize != 0).
// Call once with verts == nullptr to get the counts.
int ref_bumpy_sphere(int stacks, int slices, float amplitude, int normalize, int* nv, int* nt, float* verts,
                     uint32_t* tris, float* center, float* scale) {
  try {
    n2wos::Mesh m = n2wos::make_procedural_bumpy_sphere(stacks, slices, amplitude);
    *nv = static_cast<int>(m.vertices.size());
    *nt = static_cast<int>(m.triangles.size());
    if (!verts) return 0;
    if (normalize) {
      const n2wos::NormalizeTransform t = n2wos::normalize_to_unit_radius(m);
      center[0] = t.center.x; center[1] = t.center.y; center[2] = t.center.z;
      *scale = t.scale;
    }
    export_mesh(m, verts, tris);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

int ref_normalize(float* verts, int nv, const uint32_t* tris, int nt, float* center, float* scale) {
  try {
    n2wos::Mesh m = make_mesh(verts, nv, tris, nt);
    const n2wos::NormalizeTransform t = n2wos::normalize_to_unit_radius(m);
    center[0] = t.center.x; center[1] = t.center.y; center[2] = t.center.z;
    *scale = t.scale;
    for (int i = 0; i < nv; ++i) {
      verts[3 * i] = m.vertices[i].x;
      verts[3 * i + 1] = m.vertices[i].y;
      verts[3 * i + 2] = m.vertices[i].z;
    }
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

void ref_closest_point_on_triangle(const float* p, const float* a, const float* b, const float* c, float* out) {
  const n2wos::Vec3f r = n2wos::closest_point_on_triangle({p[0], p[1], p[2]}, {a[0], a[1], a[2]},
                                                          {b[0], b[1], b[2]}, {c[0], c[1], c[2]});
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

int ref_bruteforce(const float* verts, int nv, const uint32_t* tris, int nt, const float* pts, long n, float* d2,
                   float* cp, int* tri) {
  try {
    const n2wos::Mesh m = make_mesh(verts, nv, tris, nt);
    for (long q = 0; q < n; ++q) {
      const n2wos::ClosestPointResult r = n2wos::closest_point_bruteforce(m, {pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]});
      d2[q] = r.distance2;
      cp[3 * q] = r.closest.x; cp[3 * q + 1] = r.closest.y; cp[3 * q + 2] = r.closest.z;
      tri[q] = r.triangle_id;
    }
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

long ref_count_degenerate(const float* verts, int nv, const uint32_t* tris, int nt, float eps) {
  try {
    return static_cast<long>(n2wos::count_degenerate_triangles(make_mesh(verts, nv, tris, nt), eps));
  } catch (const std::exception&) {
    return -1;
  }
}

// load_obj_mesh / load_ply_mesh: two calls, counts first (verts == nullptr).
int ref_load_mesh(const char* path, int is_ply, int* nv, int* nt, float* verts, uint32_t* tris) {
  try {
    const n2wos::Mesh m = is_ply ? n2wos::load_ply_mesh(path) : n2wos::load_obj_mesh(path);
    *nv = static_cast<int>(m.vertices.size());
    *nt = static_cast<int>(m.triangles.size());
    if (verts) export_mesh(m, verts, tris);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

}  // extern "C"
