// tri_flat.cpp -- host build of the product's two closest-point-on-triangle
// forms (n2wos-cuda_amd/csrc/geom.hpp) for the CPU tests: the branchy form
// that follows the reference order and the branch-free form the kernels use.
// Test infrastructure only.
#include "hip/hip_runtime.h"
#include "geom.hpp"

extern "C" void tri_closest_both(const float* p, const float* a, const float* b, const float* c, long n,
                                 float* out_branchy, float* out_flat) {
  for (long i = 0; i < n; ++i) {
    const n2::V3 P = n2::v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    const n2::V3 A = n2::v3(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
    const n2::V3 B = n2::v3(b[3 * i], b[3 * i + 1], b[3 * i + 2]);
    const n2::V3 Cc = n2::v3(c[3 * i], c[3 * i + 1], c[3 * i + 2]);
    const n2::V3 r0 = n2::closest_on_triangle(P, A, B, Cc);
    const n2::V3 r1 = n2::closest_on_triangle_flat(P, A, B, Cc);
    out_branchy[3 * i] = r0.x; out_branchy[3 * i + 1] = r0.y; out_branchy[3 * i + 2] = r0.z;
    out_flat[3 * i] = r1.x; out_flat[3 * i + 1] = r1.y; out_flat[3 * i + 2] = r1.z;
  }
}
