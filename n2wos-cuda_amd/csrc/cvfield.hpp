// cvfield.hpp -- the along-walk control variate of the extended walker:
// a field C given at the vertices of a box grid, read by trilinear
// interpolation, and one walk step's zero-mean correction.
//
// A Walk-on-Spheres step moves from p to p + R d with d uniform on the unit
// sphere.  For any field C, each point p + R s of a spherical design whose
// every point s is itself uniform on the sphere has E[C(p + R s)] =
// E[C(p + R d)], so
//     t = C(p + R d) - (1/m) sum_j C(p + R s_j)
// has zero mean given p, and a walk's value minus beta times the sum of its
// steps' t is an unbiased estimate of the solution for any C and beta.  With
// C close to the solution u, the sum cancels most of the walk's variation:
// over the antipodal pair {d, -d} (m = 2) every odd term of u's expansion
// around p; over the octahedron {+-d, +-e1, +-e2} (m = 6), turned about d by
// a uniform angle so that e1 and e2 are uniform too, every term of degree
// <= 3 as well (the octahedron is a spherical 3-design).  The same arithmetic
// in the same order runs in the CPU oracle, so walks agree bit for bit.
#pragma once

#include "geom.hpp"

namespace n2 {

N2_HD float lerp_f32(float a, float b, float t) { return a + t * (b - a); }

// C at q: vertex (x, y, z) of the G^3-cell grid over the box starting at
// origin (inv_h cells per unit length) is vals[x + (G+1) (y + (G+1) z)];
// coordinates are clamped to the box.
N2_HD float cv_lookup(const float* __restrict__ vals, int G, float ox, float oy, float oz, float ihx, float ihy,
                      float ihz, V3 q) {
  const float gmax = static_cast<float>(G);
  const float sx = fminf(fmaxf((q.x - ox) * ihx, 0.0f), gmax);
  const float sy = fminf(fmaxf((q.y - oy) * ihy, 0.0f), gmax);
  const float sz = fminf(fmaxf((q.z - oz) * ihz, 0.0f), gmax);
  int ix = static_cast<int>(floorf(sx));
  int iy = static_cast<int>(floorf(sy));
  int iz = static_cast<int>(floorf(sz));
  ix = ix < G - 1 ? ix : G - 1;
  iy = iy < G - 1 ? iy : G - 1;
  iz = iz < G - 1 ? iz : G - 1;
  const float fx = sx - static_cast<float>(ix);
  const float fy = sy - static_cast<float>(iy);
  const float fz = sz - static_cast<float>(iz);
  const int stride_y = G + 1, stride_z = (G + 1) * (G + 1);
  const float* v = vals + ix + stride_y * iy + stride_z * iz;
  const float c00 = lerp_f32(v[0], v[1], fx);
  const float c10 = lerp_f32(v[stride_y], v[stride_y + 1], fx);
  const float c01 = lerp_f32(v[stride_z], v[stride_z + 1], fx);
  const float c11 = lerp_f32(v[stride_z + stride_y], v[stride_z + stride_y + 1], fx);
  const float c0 = lerp_f32(c00, c10, fy);
  const float c1 = lerp_f32(c01, c11, fy);
  return lerp_f32(c0, c1, fz);
}

// Orthonormal frame (d, e1, e2) with e1 turned by the angle 2 pi psi about
// d: the branch-free basis of Duff et al., "Building an Orthonormal Basis,
// Revisited" (JCGT 6(1), 2017), then the rotation.
N2_HD void cv_frame(V3 d, float psi, V3& e1, V3& e2) {
  const float sg = copysignf(1.0f, d.z);
  const float a = -1.0f / (sg + d.z);
  const float b = d.x * d.y * a;
  const V3 b1 = v3(1.0f + sg * d.x * d.x * a, sg * b, -sg * d.x);
  const V3 b2 = v3(b, sg + d.y * d.y * a, -d.y);
  float s, c;
  sincos2pi(psi, &s, &c);
  e1 = v3(c * b1.x + s * b2.x, c * b1.y + s * b2.y, c * b1.z + s * b2.z);
  e2 = v3(c * b2.x - s * b1.x, c * b2.y - s * b1.y, c * b2.z - s * b1.z);
}

// One step's term C(p + R d) - (1/m) sum_j C(p + R s_j), m = 2 or 6.
N2_HD float cv_term(const float* __restrict__ vals, int G, float ox, float oy, float oz, float ihx, float ihy,
                    float ihz, int m, V3 p, float R, V3 d, uint32_t psi_word) {
  const float a = cv_lookup(vals, G, ox, oy, oz, ihx, ihy, ihz, v3(p.x + R * d.x, p.y + R * d.y, p.z + R * d.z));
  float sum = a + cv_lookup(vals, G, ox, oy, oz, ihx, ihy, ihz, v3(p.x - R * d.x, p.y - R * d.y, p.z - R * d.z));
  if (m == 6) {
    V3 e1, e2;
    cv_frame(d, u01(psi_word), e1, e2);
    sum = sum + cv_lookup(vals, G, ox, oy, oz, ihx, ihy, ihz, v3(p.x + R * e1.x, p.y + R * e1.y, p.z + R * e1.z));
    sum = sum + cv_lookup(vals, G, ox, oy, oz, ihx, ihy, ihz, v3(p.x - R * e1.x, p.y - R * e1.y, p.z - R * e1.z));
    sum = sum + cv_lookup(vals, G, ox, oy, oz, ihx, ihy, ihz, v3(p.x + R * e2.x, p.y + R * e2.y, p.z + R * e2.z));
    sum = sum + cv_lookup(vals, G, ox, oy, oz, ihx, ihy, ihz, v3(p.x - R * e2.x, p.y - R * e2.y, p.z - R * e2.z));
  }
  return a - sum / static_cast<float>(m);
}

}  // namespace n2
