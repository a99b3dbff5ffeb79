// mesh_io.cpp -- mesh files (OBJ, PLY) and the host brute-force closest point.
//
// Loaders accept what the reference's loaders accept and fail on what they
// reject:
//   OBJ (src/mesh/obj_loader.cpp:63-107): "v x y z" and "f i j k ..." lines,
//     1-based or negative (relative) indices, "i/vt/vn" tokens, '#' comments,
//     other tags ignored, polygons fan-triangulated (i0, i_k, i_k+1).
//   PLY (src/mesh/ply_loader.cpp): ascii or binary_little_endian, any element
//     order, vertex x/y/z of any scalar type, the face index list named
//     vertex_indices / vertex_index / vertices (else the first list
//     property), other elements and properties skipped, fan triangulation.
// A mesh must end up with vertices, triangles and in-range indices
// (include/n2wos/types.hpp:91-103).
//
// The brute force keeps the first strict minimum of the squared distance
// (src/geometry/closest_point.cpp:50-68) and uses the product's FP32 closest
// point function, so it returns exactly what the GPU traversal returns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cfloat>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "api_internal.hpp"
#include "geom.hpp"

namespace n2 {
namespace {

struct HostMesh {
  std::vector<float> v;     // xyz
  std::vector<uint32_t> t;  // i j k
};

[[noreturn]] void fail(const std::string& m) { throw Error(N2WOS_ERR_INVALID, m); }

void fan(HostMesh& m, const std::vector<uint32_t>& face, const std::string& where) {
  if (face.size() < 3) fail("face with fewer than 3 vertices (" + where + ")");
  for (size_t k = 1; k + 1 < face.size(); ++k) {
    m.t.push_back(face[0]);
    m.t.push_back(face[k]);
    m.t.push_back(face[k + 1]);
  }
}

void finish(const HostMesh& m) {
  if (m.v.empty()) fail("mesh has no vertices");
  if (m.t.empty()) fail("mesh has no triangles");
  const uint64_t nv = m.v.size() / 3;
  for (uint32_t i : m.t)
    if (i >= nv) fail("mesh triangle contains out-of-range vertex index");
}

HostMesh load_obj(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw Error(N2WOS_ERR_INVALID, "cannot open OBJ file: " + path);
  HostMesh m;
  std::string line;
  size_t ln = 0;
  std::vector<uint32_t> face;
  while (std::getline(in, line)) {
    ++ln;
    std::istringstream ss(line);
    std::string tag;
    if (!(ss >> tag) || tag[0] == '#') continue;
    if (tag == "v") {
      float x, y, z;
      if (!(ss >> x >> y >> z)) fail("bad OBJ vertex at line " + std::to_string(ln));
      m.v.push_back(x);
      m.v.push_back(y);
      m.v.push_back(z);
    } else if (tag == "f") {
      face.clear();
      std::string tok;
      const int64_t nv = static_cast<int64_t>(m.v.size() / 3);
      while (ss >> tok) {
        const std::string head = tok.substr(0, tok.find('/'));
        char* end = nullptr;
        const long long raw = head.empty() ? 0 : std::strtoll(head.c_str(), &end, 10);
        if (head.empty() || *end != '\0') fail("bad OBJ face index '" + tok + "' at line " + std::to_string(ln));
        if (raw == 0) fail("OBJ indices are 1-based; index 0 at line " + std::to_string(ln));
        const int64_t idx = raw > 0 ? raw - 1 : nv + raw;
        if (idx < 0 || idx >= nv) fail("OBJ face index out of range at line " + std::to_string(ln));
        face.push_back(static_cast<uint32_t>(idx));
      }
      fan(m, face, "OBJ line " + std::to_string(ln));
    }
  }
  finish(m);
  return m;
}

// ------------------------------------------------------------------ PLY
enum class PType { I8, U8, I16, U16, I32, U32, F32, F64 };

PType ptype(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
  if (s == "char" || s == "int8") return PType::I8;
  if (s == "uchar" || s == "uint8") return PType::U8;
  if (s == "short" || s == "int16") return PType::I16;
  if (s == "ushort" || s == "uint16") return PType::U16;
  if (s == "int" || s == "int32") return PType::I32;
  if (s == "uint" || s == "uint32") return PType::U32;
  if (s == "float" || s == "float32") return PType::F32;
  if (s == "double" || s == "float64") return PType::F64;
  fail("unsupported PLY scalar type: " + s);
}

size_t psize(PType t) {
  switch (t) {
    case PType::I8: case PType::U8: return 1;
    case PType::I16: case PType::U16: return 2;
    case PType::I32: case PType::U32: case PType::F32: return 4;
    default: return 8;
  }
}

bool pfloat(PType t) { return t == PType::F32 || t == PType::F64; }

struct PProp {
  bool list = false;
  PType type = PType::F32;   // scalar type, or item type of a list
  PType count = PType::U8;   // count type of a list
  std::string name;
};
struct PElem {
  std::string name;
  uint64_t n = 0;
  std::vector<PProp> props;
};

// Binary little-endian cursor over the file body.
struct Bin {
  const unsigned char* p;
  const unsigned char* end;
  double scalar(PType t) {
    const size_t s = psize(t);
    if (static_cast<size_t>(end - p) < s) fail("unexpected end of binary PLY data");
    double v = 0.0;
    switch (t) {
      case PType::I8: { int8_t x; std::memcpy(&x, p, 1); v = x; break; }
      case PType::U8: { uint8_t x; std::memcpy(&x, p, 1); v = x; break; }
      case PType::I16: { int16_t x; std::memcpy(&x, p, 2); v = x; break; }
      case PType::U16: { uint16_t x; std::memcpy(&x, p, 2); v = x; break; }
      case PType::I32: { int32_t x; std::memcpy(&x, p, 4); v = x; break; }
      case PType::U32: { uint32_t x; std::memcpy(&x, p, 4); v = x; break; }
      case PType::F32: { float x; std::memcpy(&x, p, 4); v = x; break; }
      case PType::F64: { std::memcpy(&v, p, 8); break; }
    }
    p += s;
    return v;
  }
};

int face_list_index(const PElem& e) {
  for (size_t i = 0; i < e.props.size(); ++i)
    if (e.props[i].list && (e.props[i].name == "vertex_indices" || e.props[i].name == "vertex_index" ||
                            e.props[i].name == "vertices"))
      return static_cast<int>(i);
  for (size_t i = 0; i < e.props.size(); ++i)
    if (e.props[i].list) return static_cast<int>(i);
  return -1;
}

uint32_t checked_index(double v, uint64_t nv, uint64_t face) {
  if (!(v >= 0.0) || v >= static_cast<double>(nv) || v != static_cast<double>(static_cast<int64_t>(v)))
    fail("PLY face vertex index out of range in face " + std::to_string(face));
  return static_cast<uint32_t>(v);
}

HostMesh load_ply(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Error(N2WOS_ERR_INVALID, "cannot open PLY file: " + path);
  std::vector<unsigned char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  size_t pos = 0;
  auto next_line = [&](std::string& out) -> bool {
    if (pos >= buf.size()) return false;
    size_t e = pos;
    while (e < buf.size() && buf[e] != '\n') ++e;
    out.assign(reinterpret_cast<const char*>(buf.data() + pos), e - pos);
    if (!out.empty() && out.back() == '\r') out.pop_back();
    pos = e < buf.size() ? e + 1 : e;
    return true;
  };
  std::string line;
  if (!next_line(line) || line.substr(0, 3) != "ply") fail("file does not start with the PLY magic line");
  bool ascii = true, have_format = false, have_end = false;
  std::vector<PElem> elems;
  while (next_line(line)) {
    std::istringstream ss(line);
    std::string tag;
    if (!(ss >> tag) || tag == "comment" || tag == "obj_info") continue;
    if (tag == "end_header") {
      have_end = true;
      break;
    }
    if (tag == "format") {
      std::string fmt;
      ss >> fmt;
      if (fmt == "ascii") ascii = true;
      else if (fmt == "binary_little_endian") ascii = false;
      else if (fmt == "binary_big_endian") throw Error(N2WOS_ERR_UNSUPPORTED, "binary_big_endian PLY is not supported");
      else fail("unsupported PLY format: " + fmt);
      have_format = true;
    } else if (tag == "element") {
      PElem e;
      if (!(ss >> e.name >> e.n)) fail("bad PLY element line: " + line);
      elems.push_back(e);
    } else if (tag == "property") {
      if (elems.empty()) fail("PLY property before any element");
      PProp p;
      std::string a;
      ss >> a;
      if (a == "list") {
        std::string ct, it;
        if (!(ss >> ct >> it >> p.name)) fail("bad PLY list property: " + line);
        p.list = true;
        p.count = ptype(ct);
        p.type = ptype(it);
      } else {
        if (!(ss >> p.name)) fail("bad PLY property: " + line);
        p.type = ptype(a);
      }
      elems.back().props.push_back(p);
    }
  }
  if (!have_format) fail("PLY header has no format line");
  if (!have_end) fail("PLY header has no end_header");
  HostMesh m;
  uint64_t nv_total = 0;
  for (const PElem& e : elems)
    if (e.name == "vertex") nv_total = e.n;
  std::istringstream text;
  if (ascii) text.str(std::string(reinterpret_cast<const char*>(buf.data() + pos), buf.size() - pos));
  Bin bin{buf.data() + pos, buf.data() + buf.size()};
  auto read = [&](PType t) -> double {
    if (!ascii) return bin.scalar(t);
    std::string tok;
    if (!(text >> tok)) fail("unexpected end of ASCII PLY data");
    char* end = nullptr;
    const double v = std::strtod(tok.c_str(), &end);
    if (*end != '\0') fail("bad ASCII PLY number: " + tok);
    return v;
  };
  std::vector<uint32_t> face;
  for (const PElem& e : elems) {
    int ix = -1, iy = -1, iz = -1, il = -1;
    if (e.name == "vertex") {
      for (size_t i = 0; i < e.props.size(); ++i) {
        if (e.props[i].list) continue;
        if (e.props[i].name == "x") ix = static_cast<int>(i);
        if (e.props[i].name == "y") iy = static_cast<int>(i);
        if (e.props[i].name == "z") iz = static_cast<int>(i);
      }
      if (ix < 0 || iy < 0 || iz < 0) fail("PLY vertex element needs scalar x, y, z properties");
      if (!pfloat(e.props[ix].type)) fail("PLY x property must be floating point");
    } else if (e.name == "face") {
      il = face_list_index(e);
      if (il < 0) fail("PLY face element has no vertex index list");
    }
    for (uint64_t r = 0; r < e.n; ++r) {
      float xyz[3] = {0.0f, 0.0f, 0.0f};
      face.clear();
      for (size_t i = 0; i < e.props.size(); ++i) {
        const PProp& p = e.props[i];
        if (p.list) {
          const double c = read(p.count);
          if (!(c >= 0.0)) fail("negative PLY list count");
          const uint64_t cnt = static_cast<uint64_t>(c);
          for (uint64_t k = 0; k < cnt; ++k) {
            const double v = read(p.type);
            if (static_cast<int>(i) == il) face.push_back(checked_index(v, nv_total, r));
          }
        } else {
          const double v = read(p.type);
          if (static_cast<int>(i) == ix) xyz[0] = static_cast<float>(v);
          if (static_cast<int>(i) == iy) xyz[1] = static_cast<float>(v);
          if (static_cast<int>(i) == iz) xyz[2] = static_cast<float>(v);
        }
      }
      if (e.name == "vertex") {
        m.v.insert(m.v.end(), xyz, xyz + 3);
      } else if (e.name == "face") {
        fan(m, face, "PLY face " + std::to_string(r));
      }
    }
  }
  finish(m);
  return m;
}

bool ends_with(const std::string& s, const std::string& suf) {
  if (s.size() < suf.size()) return false;
  for (size_t i = 0; i < suf.size(); ++i)
    if (std::tolower(static_cast<unsigned char>(s[s.size() - suf.size() + i])) != suf[i]) return false;
  return true;
}

}  // namespace
}  // namespace n2

using namespace n2;

extern "C" int n2wos_mesh_load(const char* path, const char* kind, float** verts, int* n_verts, uint32_t** tris,
                               int* n_tris) {
  return guarded([&]() -> int {
    N2_REQUIRE(path && verts && n_verts && tris && n_tris, N2WOS_ERR_NULL, "NULL argument");
    *verts = nullptr;
    *tris = nullptr;
    std::string k = kind ? kind : "";
    const std::string p = path;
    if (k.empty()) k = ends_with(p, ".ply") ? "ply" : ends_with(p, ".obj") ? "obj" : "";
    N2_REQUIRE(k == "obj" || k == "ply", N2WOS_ERR_INVALID, "mesh kind must be obj or ply");
    const HostMesh m = k == "obj" ? load_obj(p) : load_ply(p);
    N2_REQUIRE(m.v.size() / 3 < (size_t{1} << 31) && m.t.size() / 3 < (size_t{1} << 31), N2WOS_ERR_INVALID,
               "mesh too large");
    float* v = static_cast<float*>(std::malloc(m.v.size() * sizeof(float)));
    uint32_t* t = static_cast<uint32_t*>(std::malloc(m.t.size() * sizeof(uint32_t)));
    if (!v || !t) {
      std::free(v);
      std::free(t);
      throw Error(N2WOS_ERR_NOMEM, "host allocation failed");
    }
    std::memcpy(v, m.v.data(), m.v.size() * sizeof(float));
    std::memcpy(t, m.t.data(), m.t.size() * sizeof(uint32_t));
    *verts = v;
    *tris = t;
    *n_verts = static_cast<int>(m.v.size() / 3);
    *n_tris = static_cast<int>(m.t.size() / 3);
    return N2WOS_OK;
  });
}

extern "C" int n2wos_mesh_free(float* verts, uint32_t* tris) {
  std::free(verts);
  std::free(tris);
  return N2WOS_OK;
}

extern "C" int n2wos_closest_point_bruteforce(const float* verts, int nv, const uint32_t* tris, int nt,
                                              const float* points, int n, float* d2, float* closest, int* tri,
                                              int nthreads) {
  return guarded([&]() -> int {
    N2_REQUIRE(n >= 0 && nv >= 0 && nt >= 0, N2WOS_ERR_INVALID, "negative count");
    if (n == 0) return N2WOS_OK;
    N2_REQUIRE(verts && tris && points && d2 && closest && tri, N2WOS_ERR_NULL, "NULL argument");
    N2_REQUIRE(nt > 0, N2WOS_ERR_EMPTY, "mesh has no triangles");
    for (int64_t i = 0; i < 3 * static_cast<int64_t>(nt); ++i)
      N2_REQUIRE(tris[i] < static_cast<uint32_t>(nv), N2WOS_ERR_INVALID, "triangle index out of range");
    int T = nthreads > 0 ? nthreads : static_cast<int>(std::thread::hardware_concurrency());
    T = std::max(1, std::min(T, n));
    auto work = [&](int w) {
      for (int q = w; q < n; q += T) {
        const V3 p = v3(points[3 * q], points[3 * q + 1], points[3 * q + 2]);
        float best = FLT_MAX;
        V3 bp = v3(0.0f, 0.0f, 0.0f);
        int bi = -1;
        for (int k = 0; k < nt; ++k) {
          const uint32_t* t = tris + 3 * static_cast<int64_t>(k);
          const V3 a = v3(verts[3 * t[0]], verts[3 * t[0] + 1], verts[3 * t[0] + 2]);
          const V3 b = v3(verts[3 * t[1]], verts[3 * t[1] + 1], verts[3 * t[1] + 2]);
          const V3 c = v3(verts[3 * t[2]], verts[3 * t[2] + 1], verts[3 * t[2] + 2]);
          const V3 cp = closest_on_triangle(p, a, b, c);
          const V3 e = cp - p;
          const float dd = dot(e, e);
          if (dd < best) {
            best = dd;
            bp = cp;
            bi = k;
          }
        }
        d2[q] = best;
        closest[3 * q] = bp.x;
        closest[3 * q + 1] = bp.y;
        closest[3 * q + 2] = bp.z;
        tri[q] = bi;
      }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < T; ++w) th.emplace_back(work, w);
    work(0);
    for (auto& x : th) x.join();
    return N2WOS_OK;
  });
}
