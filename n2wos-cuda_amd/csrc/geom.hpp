// geom.hpp -- FP32 geometry, RNG and boundary math shared by the host code and
// the gfx950 kernels.
//
// Everything here is compiled with -ffp-contract=off on both sides, so each
// expression rounds identically on the CPU and the GPU; where a fused
// multiply-add is wanted it is written as fmaf explicitly.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#define N2_HD __host__ __device__ __forceinline__

namespace n2 {

struct V3 {
  float x, y, z;
};

N2_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
N2_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
N2_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
N2_HD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
N2_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
N2_HD bool finite3(V3 p) { return isfinite(p.x) && isfinite(p.y) && isfinite(p.z); }

// Ericson, Real-Time Collision Detection 5.1.5: closest point of triangle abc
// to p.  Same region tests and the same operation order as the reference CPU
// oracle (src/geometry/closest_point.cpp:7-48), so the FP32 result is
// bit-identical to it.
N2_HD V3 closest_on_triangle(V3 p, V3 a, V3 b, V3 c) {
  const V3 ab = b - a;
  const V3 ac = c - a;
  const V3 ap = p - a;
  const float d1 = dot(ab, ap);
  const float d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  const V3 bp = p - b;
  const float d3 = dot(ab, bp);
  const float d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float v = d1 / (d1 - d3);
    return a + ab * v;
  }
  const V3 cp = p - c;
  const float d5 = dot(ab, cp);
  const float d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = d2 / (d2 - d6);
    return a + ac * w;
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return b + (c - b) * w;
  }
  const float denom = 1.0f / (va + vb + vc);
  const float v = vb * denom;
  const float w = vc * denom;
  return a + ab * v + ac * w;
}

// The same closest point, computed without branches for SIMD execution: all
// six dot products and the three barycentric numerators are evaluated, the
// Voronoi region is chosen by the same tests in the same order as above, and
// the one division of that region (1 / (va + vb + vc) for the face) is done
// once on the selected operands.  Every value returned is produced by the same
// IEEE operations on the same operands as closest_on_triangle, so the two are
// bit-identical; a wave whose lanes fall in different regions then executes
// one division instead of up to four.
N2_HD V3 closest_on_triangle_flat(V3 p, V3 a, V3 b, V3 c) {
  const V3 ab = b - a;
  const V3 ac = c - a;
  const V3 ap = p - a;
  const float d1 = dot(ab, ap);
  const float d2 = dot(ac, ap);
  const V3 bp = p - b;
  const float d3 = dot(ab, bp);
  const float d4 = dot(ac, bp);
  const float vc = d1 * d4 - d3 * d2;
  const V3 cp = p - c;
  const float d5 = dot(ab, cp);
  const float d6 = dot(ac, cp);
  const float vb = d5 * d2 - d1 * d6;
  const float va = d3 * d6 - d5 * d4;
  const float e43 = d4 - d3;
  const float e56 = d5 - d6;
  const bool rA = d1 <= 0.0f && d2 <= 0.0f;
  const bool rB = !rA && d3 >= 0.0f && d4 <= d3;
  const bool vtxAB = rA || rB;
  const bool rAB = !vtxAB && vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f;
  const bool rC = !vtxAB && !rAB && d6 >= 0.0f && d5 <= d6;
  const bool done4 = vtxAB || rAB || rC;
  const bool rAC = !done4 && vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f;
  const bool rBC = !done4 && !rAC && va <= 0.0f && e43 >= 0.0f && e56 >= 0.0f;
  const bool rFace = !done4 && !rAC && !rBC;
  // the region's division.  Every candidate denominator is computed and one
  // is selected (the regions exclude each other), so the compiler emits
  // selects instead of branches around each candidate.
  const float den_ab = d1 - d3;
  const float den_ac = d2 - d6;
  const float den_bc = e43 + e56;
  const float den_face = va + vb + vc;
  float num = 1.0f, den = den_face;
  num = rBC ? e43 : num;
  den = rBC ? den_bc : den;
  num = rAC ? d2 : num;
  den = rAC ? den_ac : den;
  num = rAB ? d1 : num;
  den = rAB ? den_ab : den;
  const float q = num / den;
  // edge: base + dir * q; face: (a + ab * (vb q)) + ac * (vc q).  Selects
  // are written per component so they stay register selects.
  const V3 cb = c - b;
  const float s = rFace ? vb * q : q;
  const float w = vc * q;
  const int axis_src = rAC ? 1 : (rBC ? 2 : 0);
  V3 r;
#define N2_FLAT_AXIS(X)                                                              \
  {                                                                                \
    const float base = rBC ? b.X : a.X;                                            \
    const float dir = axis_src == 1 ? ac.X : (axis_src == 2 ? cb.X : ab.X);        \
    const float e1 = base + dir * s;                                               \
    const float face = e1 + ac.X * w;                                              \
    float v = rFace ? face : e1;                                                   \
    v = rA ? a.X : v;                                                              \
    v = rB ? b.X : v;                                                              \
    v = rC ? c.X : v;                                                              \
    r.X = v;                                                                       \
  }
  N2_FLAT_AXIS(x)
  N2_FLAT_AXIS(y)
  N2_FLAT_AXIS(z)
#undef N2_FLAT_AXIS
  return r;
}

// Squared distance from p to an axis-aligned box (0 inside).  An empty box
// (lo = +inf, hi = -inf) yields +inf.
N2_HD float box_dist2(V3 p, float lox, float loy, float loz, float hix, float hiy, float hiz) {
  const float dx = fmaxf(fmaxf(lox - p.x, p.x - hix), 0.0f);
  const float dy = fmaxf(fmaxf(loy - p.y, p.y - hiy), 0.0f);
  const float dz = fmaxf(fmaxf(loz - p.z, p.z - hiz), 0.0f);
  return dx * dx + dy * dy + dz * dz;
}

// A box is visited when its squared distance is <= prune_threshold(best).
// (sqrt(best)(1 + 2^-20) + delta)^2 exceeds best by more than the FP32
// rounding of any point-triangle or point-box distance, so no triangle whose
// computed distance could tie or beat the current best is ever skipped, and
// the result equals the brute-force scan bit for bit.
N2_HD float prune_threshold(float best, float delta) {
  if (best >= FLT_MAX) return FLT_MAX;
  const float t = sqrtf(best) * 1.00000095367431640625f + delta;
  return t * t;
}

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Counter-based: (key, counter) -> 4 x u32.
struct U4 {
  uint32_t x, y, z, w;
};

N2_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c.x;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c.z;
    c = U4{static_cast<uint32_t>(p1 >> 32) ^ c.y ^ k0, static_cast<uint32_t>(p1),
           static_cast<uint32_t>(p0 >> 32) ^ c.w ^ k1, static_cast<uint32_t>(p0)};
  }
  return c;
}

// u32 -> (0,1): ((u >> 9) + 0.5) 2^-23, exact in FP32.
N2_HD float u01(uint32_t u) { return (static_cast<float>(u >> 9) + 0.5f) * 1.1920928955078125e-07f; }

// sin(2 pi v), cos(2 pi v).  Exact quarter-turn reduction, then Taylor
// polynomials on [0, pi/4] (to x^9 for sin, x^10 for cos; truncation below
// 3e-10).  Written without FMA so host and device agree bit for bit.
N2_HD void sincos2pi(float v, float* s_out, float* c_out) {
  float t = v - floorf(v);
  t = t * 4.0f;
  const float q = floorf(t);
  float f = t - q;
  const int quad = static_cast<int>(q) & 3;
  const bool swap = f > 0.5f;
  if (swap) f = 1.0f - f;
  const float x = f * 1.57079632679489662f;
  const float x2 = x * x;
  float s = 1.0f - x2 * (1.0f / 72.0f);
  s = 1.0f - x2 * (1.0f / 42.0f) * s;
  s = 1.0f - x2 * (1.0f / 20.0f) * s;
  s = 1.0f - x2 * (1.0f / 6.0f) * s;
  s = x * s;
  float c = 1.0f - x2 * (1.0f / 90.0f);
  c = 1.0f - x2 * (1.0f / 56.0f) * c;
  c = 1.0f - x2 * (1.0f / 30.0f) * c;
  c = 1.0f - x2 * (1.0f / 12.0f) * c;
  c = 1.0f - x2 * 0.5f * c;
  if (swap) {
    const float tmp = s;
    s = c;
    c = tmp;
  }
  float so, co;
  if (quad == 0) {
    co = c; so = s;
  } else if (quad == 1) {
    co = -s; so = c;
  } else if (quad == 2) {
    co = -c; so = -s;
  } else {
    co = s; so = -c;
  }
  *s_out = so;
  *c_out = co;
}

N2_HD float sin2pi(float v) {
  float s, c;
  sincos2pi(v, &s, &c);
  return s;
}

// Uniform direction on the unit sphere from two u32: z = 1 - 2u, phi = 2 pi v.
N2_HD V3 sphere_dir(uint32_t a, uint32_t b) {
  const float z = 1.0f - 2.0f * u01(a);
  const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
  float s, c;
  sincos2pi(u01(b), &s, &c);
  return v3(r * c, r * s, z);
}

N2_HD float median3(float a, float b, float c) {
  const float mx = fmaxf(fmaxf(a, b), c);
  const float mn = fminf(fminf(a, b), c);
  return (a + b + c) - mx - mn;
}

// ---------------------------------------------------------------- boundary data
// Mode ids 0, 3, 4, 5 and 12 are the reference's NcBoundaryMode values
// (include/n2wos/tcnn_nc_wos.hpp:15-36) evaluated as in
// src/cuda/cubql_bvh.cu:1196-1262; 100 = point charges sum q_i / |x - c_i|,
// 101 = manufactured Poisson solution sin(2 pi x) sin(2 pi y) sin(2 pi z) + |x|^2/2.
enum BoundaryMode : int {
  kBcX2MinusY2 = 0,
  kBcZebraK4 = 3,
  kBcZebraK8 = 4,
  kBcZebraK12 = 5,
  kBcConstantOne = 12,
  kBcCharges = 100,
  kBcPoissonSines = 101,
};

constexpr int kMaxCharges = 16;

// Boundary description passed by value to kernels.
struct Boundary {
  int mode;
  int n_charges;
  float charge[kMaxCharges][4];  // cx, cy, cz, q
};

N2_HD float complex_power_real(float x, float y, int k) {
  float ar = 1.0f, ai = 0.0f;
  for (int i = 0; i < k; ++i) {
    const float nr = ar * x - ai * y;
    const float ni = ar * y + ai * x;
    ar = nr;
    ai = ni;
  }
  return ar;
}

// Sum over point charges of q / |p - c|, r^2 floored at 1e-12 (the
// reference's charges boundary, src/cuda/cubql_bvh.cu:1264-1270); c holds
// (cx, cy, cz, q) per charge.
N2_HD float charges_value(const float* c, int n, V3 p) {
  float value = 0.0f;
  for (int i = 0; i < n; ++i) {
    const float dx = p.x - c[4 * i];
    const float dy = p.y - c[4 * i + 1];
    const float dz = p.z - c[4 * i + 2];
    value += c[4 * i + 3] / sqrtf(fmaxf(dx * dx + dy * dy + dz * dz, 1.0e-12f));
  }
  return value;
}

N2_HD float boundary_value(const Boundary& b, V3 p) {
  switch (b.mode) {
    case kBcX2MinusY2: return p.x * p.x - p.y * p.y;
    case kBcZebraK4: return complex_power_real(p.x, p.y, 4);
    case kBcZebraK8: return complex_power_real(p.x, p.y, 8);
    case kBcZebraK12: return complex_power_real(p.x, p.y, 12);
    case kBcConstantOne: return 1.0f;
    case kBcCharges: return charges_value(&b.charge[0][0], b.n_charges, p);
    case kBcPoissonSines: {
      const float s = sin2pi(p.x) * sin2pi(p.y) * sin2pi(p.z);
      return s + 0.5f * (p.x * p.x + p.y * p.y + p.z * p.z);
    }
    default: return 0.0f;
  }
}

// f = -Laplacian(u*) for the manufactured Poisson solution: 12 pi^2 sss - 3.
N2_HD float poisson_source(V3 p) {
  return 118.4352528130723f * sin2pi(p.x) * sin2pi(p.y) * sin2pi(p.z) - 3.0f;
}

}  // namespace n2
