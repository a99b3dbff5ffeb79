/*
 * oracle.c -- CPU restatement of the reference's Walk-on-Spheres hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by the smoke entry and by the CPU-baseline leg of the benchmark, and
 * never by the product library.  The product path runs on the GPU and fails
 * loudly when its HIP library is missing.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math -fPIC -shared
 * (no FMA contraction, IEEE single precision, so that every FP32 expression
 * below rounds exactly like the same expression in the reference's C++ and in
 * the HIP kernels, which are compiled with the same contraction setting).
 *
 * Contents and the reference code each part restates (paths relative to the
 * reference repository root):
 *   orc_closest_point_on_triangle  src/geometry/closest_point.cpp:7-48
 *   orc_bruteforce                 src/geometry/closest_point.cpp:50-68
 *   orc_count_degenerate           src/geometry/closest_point.cpp:70-83
 *   orc_uv_sphere                  src/mesh/procedural_bumpy_sphere.cpp:12-90
 *                                  (radius functions of the benchmark meshes
 *                                  added beside the reference bump_radius)
 *   orc_normalize                  src/mesh/normalize.cpp:9-37
 *   orc_boundary_f32               src/cuda/cubql_bvh.cu:1196-1300 (modes kept:
 *                                  x^2-y^2, zebra k, constant one, charges)
 *   orc_walk                       src/cuda/cubql_bvh.cu:1323-1357 (walk loop),
 *                                  with Philox4x32-10 replacing PCG32
 *   orc_welford / orc_chan_merge   scripts/run_nc_slice_reference_time_to_mse.py:67-139
 *   orc_bvh2_*                     src/cuda/bvh_build_host.cpp:74-120 and
 *                                  src/cuda/bvh_query.cu:75-160 (median split
 *                                  build, near-first stack traversal), with a
 *                                  lowest-index tie-break and a conservative
 *                                  pruning margin so results equal brute force
 *   orc_bvh4_traverse              restatement of the product's BVH4 traversal
 *                                  (node/triangle visit counts for the bytes
 *                                  model); not a reference function
 *   orc_philox4x32_10              Salmon et al., SC'11 (Random123)
 *   orc_sincos2pi                  polynomial sin/cos of 2*pi*v used by both
 *                                  the product and this oracle
 *   orc_cache_*                    the learned cache (hash grid + ReLU MLP) that
 *                                  replaces tiny-cuda-nn's HashGrid +
 *                                  FullyFusedMLP (src/tools/eval_tcnn_nc_wos.cu:
 *                                  747-769); FP64, plus a variant with the
 *                                  kernels' binary16 rounding points
 *   orc_walks_cv, orc_cv_*         walks with the along-walk control variate (no
 *                                  reference counterpart: a variance reduction
 *                                  of this engine)
 *
 * How each part is pinned (tests/test_oracle.py and the other CPU tests): the
 * compiled reference core and its golden vectors (closest point, brute force,
 * meshes, normalisation), known-answer vectors (Philox), closed forms
 * (triangle regions, radius formulas at the vertices, trilinear reproduction
 * of polynomials by the hash grid and by the control-variate field, the
 * control-variate term of quadratic fields, frame orthonormality), invariants
 * and statistics (zero-mean control variate, calibrated walk estimates against
 * FP64 solutions), finite differences (cache gradients) and two-pass FP64
 * (Welford / Chan).
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vld(const float* p) { return mk(p[0], p[1], p[2]); }
static inline void vst(float* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline v3 vcross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* ------------------------------------------------------------------ */
/* Closest point on a triangle (Ericson, RTCD 5.1.5), closest_point.cpp:7-48.
 * The order of every operation follows the reference's C++ expression trees:
 * Vec3f operators are component-wise and left-associative, and dot() is
 * (x*x + y*y) + z*z. */
static v3 cp_tri(v3 p, v3 a, v3 b, v3 c) {
  const v3 ab = vsub(b, a);
  const v3 ac = vsub(c, a);
  const v3 ap = vsub(p, a);
  const float d1 = vdot(ab, ap);
  const float d2 = vdot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;

  const v3 bp = vsub(p, b);
  const float d3 = vdot(ab, bp);
  const float d4 = vdot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;

  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float v = d1 / (d1 - d3);
    return vadd(a, vmul(ab, v));
  }

  const v3 cp = vsub(p, c);
  const float d5 = vdot(ab, cp);
  const float d6 = vdot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;

  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = d2 / (d2 - d6);
    return vadd(a, vmul(ac, w));
  }

  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return vadd(b, vmul(vsub(c, b), w));
  }

  const float denom = 1.0f / (va + vb + vc);
  const float v = vb * denom;
  const float w = vc * denom;
  /* a + ab * v + ac * w parses as (a + ab*v) + ac*w */
  return vadd(vadd(a, vmul(ab, v)), vmul(ac, w));
}

void orc_closest_point_on_triangle(const float* p, const float* a, const float* b,
                                   const float* c, float* out) {
  vst(out, cp_tri(vld(p), vld(a), vld(b), vld(c)));
}

/* ------------------------------------------------------------------ */
/* Brute force, closest_point.cpp:50-68: first minimum in index order wins. */
typedef struct {
  const float* verts;
  const uint32_t* tris;
  int nt;
  const float* pts;
  float* d2;
  float* cp;
  int* tri;
  long begin, end;
} bf_job;

static void bf_one(const float* verts, const uint32_t* tris, int nt, v3 p,
                   float* out_d2, v3* out_cp, int* out_tri) {
  float best = INFINITY;
  v3 bcp = mk(0.0f, 0.0f, 0.0f);
  int bid = -1;
  for (int i = 0; i < nt; ++i) {
    const uint32_t* t = tris + 3 * (size_t)i;
    const v3 c = cp_tri(p, vld(verts + 3 * (size_t)t[0]), vld(verts + 3 * (size_t)t[1]),
                        vld(verts + 3 * (size_t)t[2]));
    const v3 d = vsub(c, p);
    const float dd = vdot(d, d);
    if (dd < best) {
      best = dd;
      bcp = c;
      bid = i;
    }
  }
  *out_d2 = best;
  *out_cp = bcp;
  *out_tri = bid;
}

static void* bf_worker(void* arg) {
  bf_job* j = (bf_job*)arg;
  for (long q = j->begin; q < j->end; ++q) {
    v3 c;
    bf_one(j->verts, j->tris, j->nt, vld(j->pts + 3 * q), &j->d2[q], &c, &j->tri[q]);
    vst(j->cp + 3 * q, c);
  }
  return NULL;
}

typedef void* (*worker_fn)(void*);

static void run_threads(worker_fn fn, void* jobs, size_t job_size, int nthreads) {
  pthread_t th[256];
  if (nthreads > 256) nthreads = 256;
  for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, fn, (char*)jobs + t * job_size);
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

int orc_bruteforce(const float* verts, const uint32_t* tris, int nt, const float* pts, long n,
                   float* d2, float* cp, int* tri, int nthreads) {
  if (nt <= 0 || n < 0) return -1;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  bf_job jobs[256];
  for (int t = 0; t < nthreads; ++t) {
    jobs[t].verts = verts; jobs[t].tris = tris; jobs[t].nt = nt; jobs[t].pts = pts;
    jobs[t].d2 = d2; jobs[t].cp = cp; jobs[t].tri = tri;
    jobs[t].begin = n * t / nthreads; jobs[t].end = n * (t + 1) / nthreads;
  }
  run_threads(bf_worker, jobs, sizeof(bf_job), nthreads);
  return 0;
}

/* closest_point.cpp:70-83 */
long orc_count_degenerate(const float* verts, const uint32_t* tris, int nt, float area2_eps) {
  long count = 0;
  for (int i = 0; i < nt; ++i) {
    const v3 a = vld(verts + 3 * (size_t)tris[3 * i]);
    const v3 b = vld(verts + 3 * (size_t)tris[3 * i + 1]);
    const v3 c = vld(verts + 3 * (size_t)tris[3 * i + 2]);
    const v3 x = vcross(vsub(b, a), vsub(c, a));
    if (vdot(x, x) <= area2_eps) ++count;
  }
  return count;
}

/* ------------------------------------------------------------------ */
/* UV sphere, procedural_bumpy_sphere.cpp:32-90.  Radius kinds:
 *   0  reference bump_radius (procedural_bumpy_sphere.cpp:12-17), p[0]=amplitude
 *   1  1 + p[0]*sin(6 theta)*sin(6 phi)                      (benchmark meshes 1, 2)
 *   2  1 + 0.06 sin(5 theta) sin(7 phi) + 0.03 sin(17 theta) cos(13 phi)  (mesh 3)
 *   3  1 + sum_{l<=8,m} p[l*l+l+m] * Y_lm(theta, phi), real orthonormal SH (blob)
 */
#define ORC_PI_F 3.14159265358979323846f

static double sh_real(int l, int m, double theta, double phi);

static float radius_of(int kind, const float* prm, float theta, float phi) {
  if (kind == 0) {
    const float amplitude = prm[0];
    const float low = sinf(3.0f * theta + 0.35f * cosf(5.0f * phi));
    const float high = sinf(7.0f * phi + 0.5f * cosf(2.0f * theta));
    const float mid = cosf(5.0f * theta - 2.0f * phi);
    return 1.0f + amplitude * (0.45f * low * high + 0.35f * mid * sinf(theta) + 0.20f * sinf(4.0f * theta));
  }
  if (kind == 1) return 1.0f + prm[0] * sinf(6.0f * theta) * sinf(6.0f * phi);
  if (kind == 2)
    return 1.0f + 0.06f * sinf(5.0f * theta) * sinf(7.0f * phi) +
           0.03f * sinf(17.0f * theta) * cosf(13.0f * phi);
  /* kind 3: double-precision sum, rounded once */
  double r = 1.0;
  for (int l = 0; l <= 8; ++l)
    for (int m = -l; m <= l; ++m) r += (double)prm[l * l + l + m] * sh_real(l, m, (double)theta, (double)phi);
  return (float)r;
}

static v3 sphere_point(int kind, const float* prm, float theta, float phi) {
  const float r = radius_of(kind, prm, theta, phi);
  const float st = sinf(theta);
  return mk(r * st * cosf(phi), r * st * sinf(phi), r * cosf(theta));
}

static uint32_t ring_index(int i, int j, int slices) {
  return (uint32_t)(1 + (i - 1) * slices + (j % slices));
}

int orc_uv_sphere_counts(int stacks, int slices, int* nv, int* nt) {
  if (stacks < 4 || slices < 8) return -1;
  *nv = 2 + (stacks - 1) * slices;
  *nt = 2 * slices * (stacks - 1);
  return 0;
}

int orc_uv_sphere(int stacks, int slices, int kind, const float* prm, float* verts, uint32_t* tris) {
  if (stacks < 4 || slices < 8) return -1;
  int v = 0;
  vst(verts + 3 * v++, sphere_point(kind, prm, 0.0f, 0.0f));
  for (int i = 1; i <= stacks - 1; ++i) {
    const float theta = ORC_PI_F * (float)i / (float)stacks;
    for (int j = 0; j < slices; ++j) {
      const float phi = 2.0f * ORC_PI_F * (float)j / (float)slices;
      vst(verts + 3 * v++, sphere_point(kind, prm, theta, phi));
    }
  }
  const uint32_t south = (uint32_t)v;
  vst(verts + 3 * v++, sphere_point(kind, prm, ORC_PI_F, 0.0f));
  int t = 0;
  for (int j = 0; j < slices; ++j) {
    tris[3 * t] = 0; tris[3 * t + 1] = ring_index(1, j, slices); tris[3 * t + 2] = ring_index(1, j + 1, slices);
    ++t;
  }
  for (int i = 1; i < stacks - 1; ++i) {
    for (int j = 0; j < slices; ++j) {
      const uint32_t a = ring_index(i, j, slices), b = ring_index(i, j + 1, slices);
      const uint32_t c = ring_index(i + 1, j, slices), d = ring_index(i + 1, j + 1, slices);
      tris[3 * t] = a; tris[3 * t + 1] = c; tris[3 * t + 2] = b; ++t;
      tris[3 * t] = b; tris[3 * t + 1] = c; tris[3 * t + 2] = d; ++t;
    }
  }
  for (int j = 0; j < slices; ++j) {
    tris[3 * t] = ring_index(stacks - 1, j, slices);
    tris[3 * t + 1] = south;
    tris[3 * t + 2] = ring_index(stacks - 1, j + 1, slices);
    ++t;
  }
  return 0;
}

/* Real orthonormal spherical harmonics:
 *   Y_l0 = N_l0 P_l^0(cos t),  Y_lm = sqrt2 N_lm P_l^m(cos t) cos(m phi)  (m>0),
 *   Y_l,-m = sqrt2 N_lm P_l^m(cos t) sin(m phi),  N_lm = sqrt((2l+1)/(4pi) (l-m)!/(l+m)!),
 * P_l^m without the Condon-Shortley phase, by the standard upward recurrence. */
static double sh_real(int l, int m, double theta, double phi) {
  const int am = m < 0 ? -m : m;
  const double x = cos(theta), s = sin(theta);
  double pmm = 1.0;
  for (int i = 1; i <= am; ++i) pmm *= (2.0 * i - 1.0) * s;
  double plm;
  if (l == am) {
    plm = pmm;
  } else {
    double pm1 = x * (2.0 * am + 1.0) * pmm;
    if (l == am + 1) {
      plm = pm1;
    } else {
      double a = pmm, b = pm1;
      for (int ll = am + 2; ll <= l; ++ll) {
        const double c = ((2.0 * ll - 1.0) * x * b - (ll + am - 1.0) * a) / (double)(ll - am);
        a = b;
        b = c;
      }
      plm = b;
    }
  }
  double fact = 1.0; /* (l-m)!/(l+m)! */
  for (int i = l - am + 1; i <= l + am; ++i) fact /= (double)i;
  const double n = sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * fact);
  if (m == 0) return n * plm;
  if (m > 0) return sqrt(2.0) * n * plm * cos(am * phi);
  return sqrt(2.0) * n * plm * sin(am * phi);
}

double orc_sh_real(int l, int m, double theta, double phi) { return sh_real(l, m, theta, phi); }

/* normalize.cpp:9-37: AABB midpoint centre, divide by max radius. */
int orc_normalize(float* verts, int nv, float* center_out, float* scale_out) {
  if (nv <= 0) return -1;
  v3 lo = mk(INFINITY, INFINITY, INFINITY), hi = mk(-INFINITY, -INFINITY, -INFINITY);
  for (int i = 0; i < nv; ++i) {
    const v3 p = vld(verts + 3 * i);
    /* std::min(a, b) = (b < a) ? b : a ; std::max(a, b) = (a < b) ? b : a */
    lo = mk(p.x < lo.x ? p.x : lo.x, p.y < lo.y ? p.y : lo.y, p.z < lo.z ? p.z : lo.z);
    hi = mk(hi.x < p.x ? p.x : hi.x, hi.y < p.y ? p.y : hi.y, hi.z < p.z ? p.z : hi.z);
  }
  const v3 c = vmul(vadd(lo, hi), 0.5f);
  float radius = 0.0f;
  for (int i = 0; i < nv; ++i) {
    const v3 d = vsub(vld(verts + 3 * i), c);
    const float l = sqrtf(vdot(d, d));
    radius = radius < l ? l : radius;
  }
  if (!(radius > 0.0f)) return -2;
  for (int i = 0; i < nv; ++i) {
    const v3 d = vsub(vld(verts + 3 * i), c);
    vst(verts + 3 * i, mk(d.x / radius, d.y / radius, d.z / radius));
  }
  vst(center_out, c);
  *scale_out = 1.0f / radius;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t* ctr_in, const uint32_t* key_in, uint32_t* out) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* u32 -> float in the open interval (0,1): ((u >> 9) + 0.5) * 2^-23, exact. */
static inline float u01(uint32_t u) { return ((float)(u >> 9) + 0.5f) * 1.1920928955078125e-07f; }
float orc_u01(uint32_t u) { return u01(u); }

/* sin(2 pi v), cos(2 pi v) for any finite v.  Reduction by quarter turns is
 * exact in FP32; the kernel on [0, pi/4] is the Taylor series to x^9 (sin)
 * and x^10 (cos), truncation error < 3e-10.  Bit-identical on CPU and GPU
 * when built without FMA contraction. */
static inline void sincos2pi(float v, float* s_out, float* c_out) {
  float t = v - floorf(v);              /* [0,1) exact */
  t = t * 4.0f;                         /* exact */
  float q = floorf(t);
  float f = t - q;                      /* [0,1) exact */
  int quad = (int)q & 3;
  int swap = f > 0.5f;
  if (swap) f = 1.0f - f;               /* exact */
  const float x = f * 1.57079632679489662f;
  const float x2 = x * x;
  /* Horner: sin x = x (1 - x2/6 (1 - x2/20 (1 - x2/42 (1 - x2/72)))) */
  float s = 1.0f - x2 * (1.0f / 72.0f);
  s = 1.0f - x2 * (1.0f / 42.0f) * s;
  s = 1.0f - x2 * (1.0f / 20.0f) * s;
  s = 1.0f - x2 * (1.0f / 6.0f) * s;
  s = x * s;
  /* cos x = 1 - x2/2 (1 - x2/12 (1 - x2/30 (1 - x2/56 (1 - x2/90)))) */
  float c = 1.0f - x2 * (1.0f / 90.0f);
  c = 1.0f - x2 * (1.0f / 56.0f) * c;
  c = 1.0f - x2 * (1.0f / 30.0f) * c;
  c = 1.0f - x2 * (1.0f / 12.0f) * c;
  c = 1.0f - x2 * 0.5f * c;
  if (swap) { const float tmp = s; s = c; c = tmp; }
  float so, co;
  switch (quad) {
    case 0: co = c; so = s; break;
    case 1: co = -s; so = c; break;
    case 2: co = -c; so = -s; break;
    default: co = s; so = -c; break;
  }
  *s_out = so;
  *c_out = co;
}

void orc_sincos2pi(float v, float* s, float* c) { sincos2pi(v, s, c); }

/* ------------------------------------------------------------------ */
/* Boundary data.  Mode ids 0, 3, 4, 5, 12 are the reference's own
 * (include/n2wos/tcnn_nc_wos.hpp:15-36, evaluated as in cubql_bvh.cu:1196-1262);
 * 100 and 101 are the benchmark problems (random point charges; manufactured
 * Poisson solution u* = sin(2 pi x) sin(2 pi y) sin(2 pi z) + |x|^2 / 2). */
#define ORC_BC_X2_MINUS_Y2 0
#define ORC_BC_ZEBRA_K4 3
#define ORC_BC_ZEBRA_K8 4
#define ORC_BC_ZEBRA_K12 5
#define ORC_BC_CONSTANT_ONE 12
#define ORC_BC_CHARGES 100
#define ORC_BC_POISSON_SINES 101

static inline float sin2pi(float v) { float s, c; sincos2pi(v, &s, &c); return s; }

static float complex_power_real(float x, float y, int k) {
  float ar = 1.0f, ai = 0.0f;
  for (int i = 0; i < k; ++i) {
    const float nr = ar * x - ai * y;
    const float ni = ar * y + ai * x;
    ar = nr;
    ai = ni;
  }
  return ar;
}

/* params for ORC_BC_CHARGES: params[0] = n, then n records (cx, cy, cz, q). */
static float boundary_f32(int mode, const float* prm, v3 p) {
  switch (mode) {
    case ORC_BC_X2_MINUS_Y2: return p.x * p.x - p.y * p.y;
    case ORC_BC_ZEBRA_K4: return complex_power_real(p.x, p.y, 4);
    case ORC_BC_ZEBRA_K8: return complex_power_real(p.x, p.y, 8);
    case ORC_BC_ZEBRA_K12: return complex_power_real(p.x, p.y, 12);
    case ORC_BC_CONSTANT_ONE: return 1.0f;
    case ORC_BC_CHARGES: {
      const int n = (int)prm[0];
      float value = 0.0f;
      for (int i = 0; i < n; ++i) {
        const float* c = prm + 1 + 4 * i;
        const float dx = p.x - c[0], dy = p.y - c[1], dz = p.z - c[2];
        value += c[3] / sqrtf(fmaxf(dx * dx + dy * dy + dz * dz, 1.0e-12f));
      }
      return value;
    }
    case ORC_BC_POISSON_SINES: {
      const float s = sin2pi(p.x) * sin2pi(p.y) * sin2pi(p.z);
      return s + 0.5f * (p.x * p.x + p.y * p.y + p.z * p.z);
    }
    default: return 0.0f;
  }
}

/* f = -Laplacian(u*) = 12 pi^2 sin sin sin - 3 */
static float poisson_source_f32(v3 p) {
  return 118.4352528130723f * sin2pi(p.x) * sin2pi(p.y) * sin2pi(p.z) - 3.0f;
}

float orc_boundary_f32(int mode, const float* prm, const float* p) { return boundary_f32(mode, prm, vld(p)); }
float orc_poisson_source_f32(const float* p) { return poisson_source_f32(vld(p)); }

/* FP64 analytic solution inside the domain (all modes are harmonic except the
 * manufactured Poisson one, whose solution is u* itself). */
double orc_solution_f64(int mode, const float* prm, const double* p) {
  const double x = p[0], y = p[1], z = p[2];
  switch (mode) {
    case ORC_BC_X2_MINUS_Y2: return x * x - y * y;
    case ORC_BC_ZEBRA_K4: case ORC_BC_ZEBRA_K8: case ORC_BC_ZEBRA_K12: {
      const int k = mode == ORC_BC_ZEBRA_K4 ? 4 : (mode == ORC_BC_ZEBRA_K8 ? 8 : 12);
      double ar = 1.0, ai = 0.0;
      for (int i = 0; i < k; ++i) { const double nr = ar * x - ai * y; ai = ar * y + ai * x; ar = nr; }
      return ar;
    }
    case ORC_BC_CONSTANT_ONE: return 1.0;
    case ORC_BC_CHARGES: {
      const int n = (int)prm[0];
      double v = 0.0;
      for (int i = 0; i < n; ++i) {
        const float* c = prm + 1 + 4 * i;
        const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
        v += (double)c[3] / sqrt(dx * dx + dy * dy + dz * dz);
      }
      return v;
    }
    case ORC_BC_POISSON_SINES:
      return sin(2.0 * M_PI * x) * sin(2.0 * M_PI * y) * sin(2.0 * M_PI * z) + 0.5 * (x * x + y * y + z * z);
    default: return 0.0;
  }
}

double orc_poisson_source_f64(const double* p) {
  return 12.0 * M_PI * M_PI * sin(2.0 * M_PI * p[0]) * sin(2.0 * M_PI * p[1]) * sin(2.0 * M_PI * p[2]) - 3.0;
}

/* ------------------------------------------------------------------ */
/* Binary median-split BVH (bvh_build_host.cpp:74-120) and its near-first
 * stack traversal (bvh_query.cu:75-160).  Two changes make the result equal to
 * the brute force bit for bit: exact ties go to the lowest triangle index, and
 * a box is pruned only when it is farther than the current best by a margin
 * that covers FP32 rounding of both distances (orc_prune_threshold). */
typedef struct {
  float lo[3], hi[3];
  int left, right, first, count;
} orc_node2;

typedef struct {
  const float* verts;
  const uint32_t* tris;
  int nt;
  orc_node2* nodes;
  int nnodes;
  int* tri_index;
  float delta;  /* absolute pruning margin */
  int max_depth;
} orc_bvh2;

typedef struct {
  float lo[3], hi[3], cen[3];
} bprim;

static float cen_axis(const bprim* p, int axis) { return p->cen[axis]; }

/* quickselect: afterwards idx[k] holds the k-th smallest key, keys left of k
 * are <= it and keys right are >= it (the contract of std::nth_element). */
static void select_kth(int* idx, int n, int k, const bprim* prims, int axis) {
  int lo = 0, hi = n - 1;
  while (hi > lo) {
    const float pivot = cen_axis(&prims[idx[(lo + hi) / 2]], axis);
    int i = lo, j = hi;
    while (i <= j) {
      while (cen_axis(&prims[idx[i]], axis) < pivot) ++i;
      while (cen_axis(&prims[idx[j]], axis) > pivot) --j;
      if (i <= j) { const int t = idx[i]; idx[i] = idx[j]; idx[j] = t; ++i; --j; }
    }
    if (k <= j) hi = j;
    else if (k >= i) lo = i;
    else break;
  }
}

static int build2(orc_bvh2* b, const bprim* prims, int* idx, int begin, int end, int depth, int leaf_size) {
  if (depth > b->max_depth) b->max_depth = depth;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = begin; i < end; ++i) {
    const bprim* p = &prims[idx[i]];
    for (int a = 0; a < 3; ++a) {
      if (p->lo[a] < lo[a]) lo[a] = p->lo[a];
      if (hi[a] < p->hi[a]) hi[a] = p->hi[a];
      if (p->cen[a] < clo[a]) clo[a] = p->cen[a];
      if (chi[a] < p->cen[a]) chi[a] = p->cen[a];
    }
  }
  const int ni = b->nnodes++;
  orc_node2* nd = &b->nodes[ni];
  for (int a = 0; a < 3; ++a) { nd->lo[a] = lo[a]; nd->hi[a] = hi[a]; }
  nd->left = nd->right = nd->first = -1;
  nd->count = 0;
  const int count = end - begin;
  if (count <= leaf_size) {
    b->nodes[ni].first = begin;
    b->nodes[ni].count = count;
    return ni;
  }
  const float ex = chi[0] - clo[0], ey = chi[1] - clo[1], ez = chi[2] - clo[2];
  const int axis = (ex >= ey && ex >= ez) ? 0 : ((ey >= ex && ey >= ez) ? 1 : 2);
  const int mid = count / 2;
  select_kth(idx + begin, count, mid, prims, axis);
  const int l = build2(b, prims, idx, begin, begin + mid, depth + 1, leaf_size);
  const int r = build2(b, prims, idx, begin + mid, end, depth + 1, leaf_size);
  b->nodes[ni].left = l;
  b->nodes[ni].right = r;
  return ni;
}

static float mesh_scale(const float* verts, long nv) {
  float s = 1.0f;
  for (long i = 0; i < 3 * nv; ++i) { const float a = fabsf(verts[i]); if (a > s) s = a; }
  return s;
}

orc_bvh2* orc_bvh2_build(const float* verts, int nv, const uint32_t* tris, int nt, int leaf_size) {
  if (nt <= 0 || leaf_size < 1) return NULL;
  orc_bvh2* b = (orc_bvh2*)calloc(1, sizeof(orc_bvh2));
  bprim* prims = (bprim*)malloc(sizeof(bprim) * (size_t)nt);
  int* idx = (int*)malloc(sizeof(int) * (size_t)nt);
  for (int i = 0; i < nt; ++i) {
    const v3 a = vld(verts + 3 * (size_t)tris[3 * i]);
    const v3 bb = vld(verts + 3 * (size_t)tris[3 * i + 1]);
    const v3 c = vld(verts + 3 * (size_t)tris[3 * i + 2]);
    const float xs[3][3] = {{a.x, a.y, a.z}, {bb.x, bb.y, bb.z}, {c.x, c.y, c.z}};
    for (int k = 0; k < 3; ++k) {
      prims[i].lo[k] = fminf(fminf(xs[0][k], xs[1][k]), xs[2][k]);
      prims[i].hi[k] = fmaxf(fmaxf(xs[0][k], xs[1][k]), xs[2][k]);
    }
    const v3 s3 = vadd(vadd(a, bb), c);
    const v3 cen = mk(s3.x / 3.0f, s3.y / 3.0f, s3.z / 3.0f);
    prims[i].cen[0] = cen.x; prims[i].cen[1] = cen.y; prims[i].cen[2] = cen.z;
    idx[i] = i;
  }
  b->verts = verts; b->tris = tris; b->nt = nt;
  b->nodes = (orc_node2*)malloc(sizeof(orc_node2) * (size_t)(2 * nt + 1));
  b->nnodes = 0;
  b->max_depth = 0;
  build2(b, prims, idx, 0, nt, 0, leaf_size);
  b->tri_index = idx;
  b->delta = 3.814697265625e-06f * mesh_scale(verts, nv); /* 2^-18 * scale */
  free(prims);
  return b;
}

void orc_bvh2_free(orc_bvh2* b) {
  if (!b) return;
  free(b->nodes);
  free(b->tri_index);
  free(b);
}

int orc_bvh2_info(const orc_bvh2* b, int* nnodes, int* max_depth) {
  *nnodes = b->nnodes;
  *max_depth = b->max_depth;
  return 0;
}

/* Pruning threshold for a current best squared distance: a box is visited when
 * its squared distance is <= this value.  (sqrt(best) (1 + 2^-20) + delta)^2
 * exceeds best by more than the FP32 rounding of a point-triangle distance. */
static inline float prune_threshold(float best, float delta) {
  if (best >= FLT_MAX) return FLT_MAX;
  const float t = sqrtf(best) * 1.00000095367431640625f + delta;
  return t * t;
}
float orc_prune_threshold(float best, float delta) { return prune_threshold(best, delta); }

static inline float box_d2(v3 p, const float* lo, const float* hi) {
  const float dx = fmaxf(fmaxf(lo[0] - p.x, p.x - hi[0]), 0.0f);
  const float dy = fmaxf(fmaxf(lo[1] - p.y, p.y - hi[1]), 0.0f);
  const float dz = fmaxf(fmaxf(lo[2] - p.z, p.z - hi[2]), 0.0f);
  return dx * dx + dy * dy + dz * dz;
}

static inline int finite3(v3 p) { return isfinite(p.x) && isfinite(p.y) && isfinite(p.z); }

static void bvh2_query(const orc_bvh2* b, v3 p, float* out_d2, v3* out_cp, int* out_tri) {
  float best = FLT_MAX;
  v3 bcp = mk(0.0f, 0.0f, 0.0f);
  int bid = -1;
  if (finite3(p)) {
    float thr = FLT_MAX;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const orc_node2* nd = &b->nodes[stack[--sp]];
      if (box_d2(p, nd->lo, nd->hi) > thr) continue;
      if (nd->count > 0) {
        for (int i = 0; i < nd->count; ++i) {
          const int id = b->tri_index[nd->first + i];
          const uint32_t* t = b->tris + 3 * (size_t)id;
          const v3 c = cp_tri(p, vld(b->verts + 3 * (size_t)t[0]), vld(b->verts + 3 * (size_t)t[1]),
                              vld(b->verts + 3 * (size_t)t[2]));
          const v3 d = vsub(c, p);
          const float dd = vdot(d, d);
          if (dd < best || (dd == best && id < bid)) {
            best = dd; bcp = c; bid = id;
            thr = prune_threshold(best, b->delta);
          }
        }
      } else {
        const orc_node2* l = &b->nodes[nd->left];
        const orc_node2* r = &b->nodes[nd->right];
        const float ld = box_d2(p, l->lo, l->hi), rd = box_d2(p, r->lo, r->hi);
        int near = nd->left, far = nd->right;
        float nd2 = ld, fd2 = rd;
        if (rd < ld) { near = nd->right; far = nd->left; nd2 = rd; fd2 = ld; }
        if (fd2 <= thr && sp < 128) stack[sp++] = far;
        if (nd2 <= thr && sp < 128) stack[sp++] = near;
      }
    }
  }
  *out_d2 = best;
  *out_cp = bcp;
  *out_tri = bid;
}

typedef struct {
  const orc_bvh2* b;
  const float* pts;
  float* d2;
  float* cp;
  int* tri;
  long begin, end;
} bq_job;

static void* bq_worker(void* arg) {
  bq_job* j = (bq_job*)arg;
  for (long q = j->begin; q < j->end; ++q) {
    v3 c;
    bvh2_query(j->b, vld(j->pts + 3 * q), &j->d2[q], &c, &j->tri[q]);
    vst(j->cp + 3 * q, c);
  }
  return NULL;
}

int orc_bvh2_query(const orc_bvh2* b, const float* pts, long n, float* d2, float* cp, int* tri, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  bq_job jobs[256];
  for (int t = 0; t < nthreads; ++t) {
    jobs[t].b = b; jobs[t].pts = pts; jobs[t].d2 = d2; jobs[t].cp = cp; jobs[t].tri = tri;
    jobs[t].begin = n * t / nthreads; jobs[t].end = n * (t + 1) / nthreads;
  }
  run_threads(bq_worker, jobs, sizeof(bq_job), nthreads);
  return 0;
}

/* ------------------------------------------------------------------ */
/* One Walk-on-Spheres walk, following nc_wos_from_point (cubql_bvh.cu:1323-1357)
 * and the depth-m prefix of nc_hybrid_prefix_kernel (:1451-1516):
 *   for step = 0..max_steps: query; d = sqrt(max(d2, 0));
 *     stop with src + g(P) if d <= eps, step == max_steps or d is not finite;
 *     stop with the prefix state if step >= depth (depth >= 0);
 *     R = step_scale * d; add the ball's source term; p += R * dir.
 * Randomness: Philox4x32-10 with key = seed and counter
 * (step, walk, point, 2*stream) for the direction (words 0, 1) and, for the
 * Poisson source, words 2, 3 plus counter (step, walk, point, 2*stream + 1)
 * for the Green's-function sample: radius fraction rho ~ Beta(2,2) as the
 * median of three uniforms, uniform direction, contribution R^2/6 f(y). */
typedef struct {
  int max_steps;
  float eps;
  float step_scale;
  int bmode;
  const float* bparams;
  int poisson;
  uint32_t seed_lo, seed_hi;
  uint32_t stream;
  int antithetic; /* walks 2j+1 take the negated step-0 draws of walk 2j */
  /* along-walk control variate (plain walks only; cv_points 0 = off) */
  int cv_points;         /* 2: antipodal pair; 6: rotated octahedron */
  int cv_res;            /* G cells per axis, (G+1)^3 vertex values */
  const float* cv_grid;
  float cv_origin[3], cv_inv_h[3];
  float cv_beta;
} orc_walk_opts;

/* Control-variate field: C at the (G+1)^3 vertices of a box (origin, inverse
 * cell size per axis), vertex (x, y, z) at index x + (G+1) (y + (G+1) z),
 * read by trilinear interpolation in FP32 (coordinates clamped to the box). */
static inline float lerp_f32(float a, float b, float t) { return a + t * (b - a); }

static float cv_lookup(const orc_walk_opts* o, float qx, float qy, float qz) {
  const int G = o->cv_res;
  const float q[3] = {qx, qy, qz};
  int i[3];
  float f[3];
  for (int a = 0; a < 3; ++a) {
    const float sa = fminf(fmaxf((q[a] - o->cv_origin[a]) * o->cv_inv_h[a], 0.0f), (float)G);
    int ia = (int)floorf(sa);
    ia = ia < G - 1 ? ia : G - 1;
    i[a] = ia;
    f[a] = sa - (float)ia;
  }
  const long sy = G + 1, sz = (long)(G + 1) * (G + 1);
  const float* v = o->cv_grid + i[0] + sy * i[1] + sz * i[2];
  const float c00 = lerp_f32(v[0], v[1], f[0]);
  const float c10 = lerp_f32(v[sy], v[sy + 1], f[0]);
  const float c01 = lerp_f32(v[sz], v[sz + 1], f[0]);
  const float c11 = lerp_f32(v[sz + sy], v[sz + sy + 1], f[0]);
  const float c0 = lerp_f32(c00, c10, f[1]);
  const float c1 = lerp_f32(c01, c11, f[1]);
  return lerp_f32(c0, c1, f[2]);
}

/* Orthonormal frame (d, e1, e2), e1 turned by the angle 2 pi psi about d:
 * the branch-free basis of Duff et al., "Building an Orthonormal Basis,
 * Revisited" (JCGT 6(1), 2017), then the rotation.  With d uniform on the
 * sphere and psi uniform, each of +-d, +-e1, +-e2 is uniform on the sphere. */
static void cv_frame(v3 d, float psi, v3* e1, v3* e2) {
  const float sg = copysignf(1.0f, d.z);
  const float a = -1.0f / (sg + d.z);
  const float b = d.x * d.y * a;
  const v3 b1 = mk(1.0f + sg * d.x * d.x * a, sg * b, -sg * d.x);
  const v3 b2 = mk(b, sg + d.y * d.y * a, -d.y);
  float s, c;
  sincos2pi(psi, &s, &c);
  *e1 = mk(c * b1.x + s * b2.x, c * b1.y + s * b2.y, c * b1.z + s * b2.z);
  *e2 = mk(c * b2.x - s * b1.x, c * b2.y - s * b1.y, c * b2.z - s * b1.z);
}

/* One step's control-variate term C(p + R d) - (1/m) sum_j C(p + R s_j) over
 * the design {d, -d} (m = 2) or {d, -d, e1, -e1, e2, -e2} (m = 6): zero mean
 * for any C, because every design point is uniform on the sphere of radius R
 * (Walk-on-Spheres' next position is p + R d). */
static float cv_term(const orc_walk_opts* o, v3 p, float R, v3 d, uint32_t psi_word) {
  const float a = cv_lookup(o, p.x + R * d.x, p.y + R * d.y, p.z + R * d.z);
  float sum = a + cv_lookup(o, p.x - R * d.x, p.y - R * d.y, p.z - R * d.z);
  if (o->cv_points == 6) {
    v3 e1, e2;
    cv_frame(d, u01(psi_word), &e1, &e2);
    sum = sum + cv_lookup(o, p.x + R * e1.x, p.y + R * e1.y, p.z + R * e1.z);
    sum = sum + cv_lookup(o, p.x - R * e1.x, p.y - R * e1.y, p.z - R * e1.z);
    sum = sum + cv_lookup(o, p.x + R * e2.x, p.y + R * e2.y, p.z + R * e2.z);
    sum = sum + cv_lookup(o, p.x - R * e2.x, p.y - R * e2.y, p.z - R * e2.z);
  }
  return a - sum / (float)o->cv_points;
}

/* Test hooks: one control-variate term and one frame, for the closed-form
 * checks of the CPU tests. */
float orc_cv_term(const float* grid, int res, const float* origin, const float* inv_h, int points, const float* p,
                  float R, const float* d, uint32_t psi_word) {
  orc_walk_opts o = {0, 0.0f, 0.0f, 0, NULL, 0, 0u, 0u, 0u, 0, 0, 0, NULL, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 0.0f};
  o.cv_points = points;
  o.cv_res = res;
  o.cv_grid = grid;
  for (int a = 0; a < 3; ++a) {
    o.cv_origin[a] = origin[a];
    o.cv_inv_h[a] = inv_h[a];
  }
  return cv_term(&o, mk(p[0], p[1], p[2]), R, mk(d[0], d[1], d[2]), psi_word);
}

float orc_cv_lookup(const float* grid, int res, const float* origin, const float* inv_h, const float* q) {
  orc_walk_opts o = {0, 0.0f, 0.0f, 0, NULL, 0, 0u, 0u, 0u, 0, 0, 0, NULL, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 0.0f};
  o.cv_res = res;
  o.cv_grid = grid;
  for (int a = 0; a < 3; ++a) {
    o.cv_origin[a] = origin[a];
    o.cv_inv_h[a] = inv_h[a];
  }
  return cv_lookup(&o, q[0], q[1], q[2]);
}

void orc_cv_frame(const float* d, float psi, float* e1, float* e2) {
  v3 a, b;
  cv_frame(mk(d[0], d[1], d[2]), psi, &a, &b);
  e1[0] = a.x; e1[1] = a.y; e1[2] = a.z;
  e2[0] = b.x; e2[1] = b.y; e2[2] = b.z;
}

static inline float median3(float a, float b, float c) {
  const float mx = fmaxf(fmaxf(a, b), c), mn = fminf(fminf(a, b), c);
  return (a + b + c) - mx - mn;
}

static inline v3 dir_from(uint32_t u0, uint32_t u1) {
  const float z = 1.0f - 2.0f * u01(u0);
  const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
  float s, c;
  sincos2pi(u01(u1), &s, &c);
  return mk(r * c, r * s, z);
}

/* returns 0 boundary, 1 forced (max steps), 2 stopped at depth.  first_*
 * optionally supplies step 0's query (identical for every walk of a point). */
static int walk_one(const orc_bvh2* b, const orc_walk_opts* o, v3 p, uint32_t point, uint32_t walk, int depth,
                    const float* first_d2, const v3* first_cp, float* value, int* steps, v3* xm) {
  float src = 0.0f;
  float cvs = 0.0f; /* control-variate sum (plain walks with cv_points) */
  const int cv = o->cv_points > 0 && depth < 0;
  for (int step = 0; step <= o->max_steps; ++step) {
    float d2; v3 P; int tri;
    if (step == 0 && first_d2) {
      d2 = *first_d2;
      P = *first_cp;
    } else if (!finite3(p)) {
      d2 = INFINITY; /* a non-finite position ends the walk */
      P = mk(0.0f, 0.0f, 0.0f);
    } else {
      bvh2_query(b, p, &d2, &P, &tri);
    }
    const float d = sqrtf(fmaxf(d2, 0.0f));
    const int forced = step >= o->max_steps;
    if (d <= o->eps || forced || !isfinite(d)) {
      *value = src + boundary_f32(o->bmode, o->bparams, P);
      if (cv) *value = *value - o->cv_beta * cvs;
      *steps = step;
      *xm = p;
      return forced ? 1 : 0;
    }
    if (depth >= 0 && step >= depth) {
      *value = src;
      *steps = step;
      *xm = p;
      return 2;
    }
    const uint32_t key[2] = {o->seed_lo, o->seed_hi};
    /* antithetic pair: the odd walk mirrors its even partner's first step */
    const int anti = o->antithetic && step == 0 && (walk & 1u);
    const uint32_t cw = anti ? walk - 1u : walk;
    const float sgn = anti ? -1.0f : 1.0f;
    const uint32_t ctr[4] = {(uint32_t)step, cw, point, 2u * o->stream};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    v3 dir = dir_from(r[0], r[1]);
    dir = mk(sgn * dir.x, sgn * dir.y, sgn * dir.z);
    const float R = o->step_scale * d;
    uint32_t psi_word = r[2]; /* the frame angle: word 2 (Laplace) or word 7 (Poisson) */
    if (o->poisson) {
      const uint32_t ctr2[4] = {(uint32_t)step, cw, point, 2u * o->stream + 1u};
      uint32_t t[4];
      orc_philox4x32_10(ctr2, key, t);
      psi_word = t[3];
      const float rho = median3(u01(r[2]), u01(r[3]), u01(t[0]));
      v3 e = dir_from(t[1], t[2]);
      e = mk(sgn * e.x, sgn * e.y, sgn * e.z);
      const float Rr = R * rho;
      const v3 y = mk(p.x + Rr * e.x, p.y + Rr * e.y, p.z + Rr * e.z);
      src = src + (R * R) * poisson_source_f32(y) / 6.0f;
    }
    if (cv) cvs = cvs + cv_term(o, p, R, dir, psi_word);
    p = mk(p.x + R * dir.x, p.y + R * dir.y, p.z + R * dir.z);
  }
  *value = cv ? src - o->cv_beta * cvs : src;
  *steps = o->max_steps;
  *xm = p;
  return 1;
}

/* Shared state of one multi-threaded walk job: samples are handed out in
 * chunks from an atomic counter (dynamic scheduling), after a first pass that
 * computes each point's step-0 query once. */
typedef struct {
  const orc_bvh2* b;
  const orc_walk_opts* o;
  const float* pts;
  const uint32_t* point_ids;
  long npts;
  int wpp;
  int depth;
  float* values;
  int* steps;
  int* status;
  float* xm;
  float* first_d2;
  v3* first_cp;
  long total;
  long next_point;
  long next_sample;
} wk_shared;

static void* wk_worker(void* arg) {
  wk_shared* j = *(wk_shared**)arg;
  for (;;) {
    const long p0 = __atomic_fetch_add(&j->next_point, 64, __ATOMIC_RELAXED);
    if (p0 >= j->npts) break;
    const long p1 = p0 + 64 < j->npts ? p0 + 64 : j->npts;
    for (long pi = p0; pi < p1; ++pi) {
      int tri;
      bvh2_query(j->b, vld(j->pts + 3 * pi), &j->first_d2[pi], &j->first_cp[pi], &tri);
    }
  }
  return NULL;
}

static void* wk_worker2(void* arg) {
  wk_shared* j = *(wk_shared**)arg;
  for (;;) {
    const long s0 = __atomic_fetch_add(&j->next_sample, 16, __ATOMIC_RELAXED);
    if (s0 >= j->total) break;
    const long s1 = s0 + 16 < j->total ? s0 + 16 : j->total;
    for (long s = s0; s < s1; ++s) {
      const long pi = s / j->wpp;
      const uint32_t w = (uint32_t)(s - pi * j->wpp);
      const uint32_t pid = j->point_ids ? j->point_ids[pi] : (uint32_t)pi;
      v3 xm;
      const int st = walk_one(j->b, j->o, vld(j->pts + 3 * pi), pid, w, j->depth, &j->first_d2[pi],
                              &j->first_cp[pi], &j->values[s], &j->steps[s], &xm);
      if (j->status) j->status[s] = st;
      if (j->xm) vst(j->xm + 3 * s, xm);
    }
  }
  return NULL;
}

/* Runs npts * wpp walks; sample s = point * wpp + walk.  point_ids (optional)
 * gives the RNG point index of each point (default: its position).  With
 * antithetic != 0, walk 2j+1 takes the negated step-0 draws of walk 2j. */
int orc_walks(const orc_bvh2* b, int max_steps, float eps, float step_scale, int bmode, const float* bparams,
              int poisson, uint64_t seed, uint32_t stream, const float* pts, const uint32_t* point_ids, long npts,
              int wpp, int depth, float* values, int* steps, int* status, float* xm, int nthreads, int antithetic) {
  if (!b || wpp < 1 || npts < 0) return -1;
  orc_walk_opts o = {max_steps, eps, step_scale, bmode, bparams, poisson, (uint32_t)seed, (uint32_t)(seed >> 32), stream,
                     antithetic, 0, 0, NULL, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 0.0f};
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  wk_shared sh;
  sh.b = b; sh.o = &o; sh.pts = pts; sh.point_ids = point_ids; sh.npts = npts; sh.wpp = wpp; sh.depth = depth;
  sh.values = values; sh.steps = steps; sh.status = status; sh.xm = xm;
  sh.first_d2 = (float*)malloc(sizeof(float) * (size_t)(npts > 0 ? npts : 1));
  sh.first_cp = (v3*)malloc(sizeof(v3) * (size_t)(npts > 0 ? npts : 1));
  sh.total = npts * wpp;
  sh.next_point = 0;
  sh.next_sample = 0;
  wk_shared* ptrs[256];
  for (int t = 0; t < nthreads; ++t) ptrs[t] = &sh;
  run_threads(wk_worker, ptrs, sizeof(wk_shared*), nthreads);
  run_threads(wk_worker2, ptrs, sizeof(wk_shared*), nthreads);
  free(sh.first_d2);
  free(sh.first_cp);
  return 0;
}

/* Plain walks with the along-walk control variate (cv_points 2 or 6, field
 * grid[(G+1)^3] over the box origin + [0, G / inv_h)): each walk's value is
 * W - beta sum_k [C(X_k + R_k d_k) - (1/m) sum_j C(X_k + R_k s_j)]. */
int orc_walks_cv(const orc_bvh2* b, int max_steps, float eps, float step_scale, int bmode, const float* bparams,
                 int poisson, uint64_t seed, uint32_t stream, const float* pts, long npts, int wpp, float* values,
                 int* steps, int nthreads, int cv_points, int cv_res, const float* grid, const float* origin,
                 const float* inv_h, float beta) {
  if (!b || wpp < 1 || npts < 0 || (cv_points != 2 && cv_points != 6) || cv_res < 1 || !grid) return -1;
  orc_walk_opts o = {max_steps, eps, step_scale, bmode, bparams, poisson, (uint32_t)seed, (uint32_t)(seed >> 32), stream,
                     0, 0, 0, NULL, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 0.0f};
  o.cv_points = cv_points;
  o.cv_res = cv_res;
  o.cv_grid = grid;
  for (int a = 0; a < 3; ++a) {
    o.cv_origin[a] = origin[a];
    o.cv_inv_h[a] = inv_h[a];
  }
  o.cv_beta = beta;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  wk_shared sh;
  sh.b = b; sh.o = &o; sh.pts = pts; sh.point_ids = NULL; sh.npts = npts; sh.wpp = wpp; sh.depth = -1;
  sh.values = values; sh.steps = steps; sh.status = NULL; sh.xm = NULL;
  sh.first_d2 = (float*)malloc(sizeof(float) * (size_t)(npts > 0 ? npts : 1));
  sh.first_cp = (v3*)malloc(sizeof(v3) * (size_t)(npts > 0 ? npts : 1));
  sh.total = npts * wpp;
  sh.next_point = 0;
  sh.next_sample = 0;
  wk_shared* ptrs[256];
  for (int t = 0; t < nthreads; ++t) ptrs[t] = &sh;
  run_threads(wk_worker, ptrs, sizeof(wk_shared*), nthreads);
  run_threads(wk_worker2, ptrs, sizeof(wk_shared*), nthreads);
  free(sh.first_d2);
  free(sh.first_cp);
  return 0;
}

/* ------------------------------------------------------------------ */
/* FP64 statistics.  Welford update and Chan et al. pairwise merge (the merge
 * the reference applies to chunked reference runs,
 * scripts/run_nc_slice_reference_time_to_mse.py:85-112). */
void orc_welford(const float* v, long n, double* mean, double* m2) {
  double m = 0.0, s = 0.0;
  for (long i = 0; i < n; ++i) {
    const double x = (double)v[i];
    const double d = x - m;
    m += d / (double)(i + 1);
    s += d * (x - m);
  }
  *mean = m;
  *m2 = s;
}

static inline void chan(double na, double ma, double sa, double nb, double mb, double sb,
                        double* n, double* m, double* s) {
  if (nb == 0.0) { *n = na; *m = ma; *s = sa; return; }
  if (na == 0.0) { *n = nb; *m = mb; *s = sb; return; }
  const double nn = na + nb;
  const double d = mb - ma;
  *n = nn;
  *m = ma + d * (nb / nn);
  *s = sa + sb + d * d * (na * nb / nn);
}

void orc_chan_merge(double na, double ma, double sa, double nb, double mb, double sb, double* out) {
  chan(na, ma, sa, nb, mb, sb, &out[0], &out[1], &out[2]);
}

/* Per-point statistics in the product's fixed reduction order: walks are
 * grouped in units of 64 consecutive walk indices; inside a unit, pairs are
 * merged with strides 1, 2, 4, ..., 32 (lower index keeps the result); units
 * are then merged left to right.  values[p * wpp + w]. */
void orc_point_stats(const float* values, long npts, int wpp, double* out_n, double* out_mean, double* out_m2) {
  for (long p = 0; p < npts; ++p) {
    double N = 0.0, M = 0.0, S = 0.0;
    for (int u0 = 0; u0 < wpp; u0 += 64) {
      double n[64], m[64], s[64];
      const int cnt = (wpp - u0) < 64 ? (wpp - u0) : 64;
      for (int i = 0; i < 64; ++i) {
        n[i] = i < cnt ? 1.0 : 0.0;
        m[i] = i < cnt ? (double)values[p * wpp + u0 + i] : 0.0;
        s[i] = 0.0;
      }
      for (int stride = 1; stride < 64; stride *= 2)
        for (int i = 0; i < 64; i += 2 * stride)
          chan(n[i], m[i], s[i], n[i + stride], m[i + stride], s[i + stride], &n[i], &m[i], &s[i]);
      chan(N, M, S, n[0], m[0], s[0], &N, &M, &S);
    }
    out_n[p] = N;
    out_mean[p] = M;
    out_m2[p] = S;
  }
}

/* ------------------------------------------------------------------ */
/* Restatement of the product's 4-wide BVH traversal (node 128 B: six float4
 * planes lox, loy, loz, hix, hiy, hiz, then four child refs; triangle record
 * 48 B: a, b, c, original id).  Same operation order and the same visit order
 * as the kernel, whose LDS ring plus spill array behaves as one stack of
 * cap + 48 entries (beyond that the newest spill is dropped and the query is
 * flagged), so results AND node/triangle visit counts equal the kernel's. */
typedef struct {
  float lox[4], loy[4], loz[4], hix[4], hiy[4], hiz[4];
  uint32_t child[4];
  uint32_t pad[4];
} orc_node4;

typedef struct {
  float ax, ay, az, bx, by, bz, cx, cy, cz;
  int32_t id;
  int32_t pad0, pad1;
} orc_tri4;

static inline void cswap4(float* ka, uint32_t* ra, int* sa, float* kb, uint32_t* rb, int* sb) {
  if ((*kb < *ka) || (*kb == *ka && *sb < *sa)) {
    float tk = *ka; *ka = *kb; *kb = tk;
    uint32_t tr = *ra; *ra = *rb; *rb = tr;
    int ts = *sa; *sa = *sb; *sb = ts;
  }
}

static int g_bvh4_maxdepth; /* deepest stack of the last query (diagnostics) */
static int g_bvh4_pushes;   /* pushes of the last query */

static void bvh4_query(const orc_node4* nodes, const orc_tri4* tris, float delta, v3 p, float thr_init, int cap,
                       float* out_d2, v3* out_cp, int* out_id, int* out_ovf, uint32_t* out_nodes,
                       uint32_t* out_tris) {
  g_bvh4_maxdepth = 0;
  g_bvh4_pushes = 0;
  float best = FLT_MAX;
  v3 bcp = mk(0.0f, 0.0f, 0.0f);
  int bid = -1, ovf = 0;
  uint32_t nn = 0, nt = 0;
  if (finite3(p)) {
    /* entries [0, sp): the bottom `spill` ones model the spill array, the top
       min(sp, cap) ones the LDS ring */
    uint32_t sref[64 + 48];
    float sd2[64 + 48];
    const int limit = cap + 48;
    float thr = thr_init;
    uint32_t cur = 0;
    int have = 1, sp = 0;
    for (;;) {
      if (have) {
        if (!(cur & 0x80000000u)) {
          const orc_node4* nd = &nodes[cur];
          ++nn;
          float k[4];
          for (int i = 0; i < 4; ++i) {
            const float lo[3] = {nd->lox[i], nd->loy[i], nd->loz[i]};
            const float hi[3] = {nd->hix[i], nd->hiy[i], nd->hiz[i]};
            k[i] = box_d2(p, lo, hi);
          }
          const int nvalid = (k[0] <= thr) + (k[1] <= thr) + (k[2] <= thr) + (k[3] <= thr);
          if (nvalid == 0) {
            have = 0;
          } else {
            uint32_t r[4] = {nd->child[0], nd->child[1], nd->child[2], nd->child[3]};
            int s[4] = {0, 1, 2, 3};
            for (int i = 0; i < 4; ++i) k[i] = k[i] <= thr ? k[i] : INFINITY;
            cswap4(&k[0], &r[0], &s[0], &k[1], &r[1], &s[1]);
            cswap4(&k[2], &r[2], &s[2], &k[3], &r[3], &s[3]);
            cswap4(&k[0], &r[0], &s[0], &k[2], &r[2], &s[2]);
            cswap4(&k[1], &r[1], &s[1], &k[3], &r[3], &s[3]);
            cswap4(&k[1], &r[1], &s[1], &k[2], &r[2], &s[2]);
            for (int i = nvalid - 1; i >= 1; --i) {
              if (sp == limit) {
                /* the kernel loses the oldest ring entry, which sits at
                   position limit - cap in this model */
                ovf = 1;
                for (int j = limit - cap; j < limit - 1; ++j) { sref[j] = sref[j + 1]; sd2[j] = sd2[j + 1]; }
                --sp;
              }
              sref[sp] = r[i];
              sd2[sp] = k[i];
              ++sp;
              ++g_bvh4_pushes;
              if (sp > g_bvh4_maxdepth) g_bvh4_maxdepth = sp;
            }
            cur = r[0];
            continue;
          }
        } else {
          const int first = (int)((cur >> 5) & 0x3ffffffu), count = (int)(cur & 31u);
          for (int i = 0; i < count; ++i) {
            const orc_tri4* t = &tris[first + i];
            ++nt;
            const v3 c = cp_tri(p, mk(t->ax, t->ay, t->az), mk(t->bx, t->by, t->bz), mk(t->cx, t->cy, t->cz));
            const v3 e = vsub(c, p);
            const float dd = vdot(e, e);
            if (dd < best || (dd == best && t->id < bid)) {
              best = dd; bcp = c; bid = t->id;
              thr = fminf(thr, prune_threshold(dd, delta));
            }
          }
          have = 0;
        }
      }
      while (sp > 0) {
        --sp;
        if (sd2[sp] <= thr) { cur = sref[sp]; have = 1; break; }
      }
      if (!have) break;
    }
  }
  *out_d2 = best; *out_cp = bcp; *out_id = bid; *out_ovf = ovf;
  if (out_nodes) *out_nodes = nn;
  if (out_tris) *out_tris = nt;
}

int orc_bvh4_query(const void* nodes, const void* tris, float delta, int cap, const float* pts, long n,
                   const float* thr_init, float* d2, float* cp, int* id, int* ovf, uint32_t* nnodes,
                   uint32_t* ntris) {
  if (cap < 1 || cap > 64 || (cap & (cap - 1))) return -1;
  for (long q = 0; q < n; ++q) {
    v3 c;
    bvh4_query((const orc_node4*)nodes, (const orc_tri4*)tris, delta, vld(pts + 3 * q),
               thr_init ? thr_init[q] : FLT_MAX, cap, &d2[q], &c, &id[q], &ovf[q],
               nnodes ? &nnodes[q] : NULL, ntris ? &ntris[q] : NULL);
    vst(cp + 3 * q, c);
  }
  return 0;
}

/* Diagnostics: deepest stack and number of pushes of each query. */
int orc_bvh4_stack_profile(const void* nodes, const void* tris, float delta, const float* pts, long n,
                           const float* thr_init, int* max_depth, int* pushes) {
  for (long q = 0; q < n; ++q) {
    float d2; v3 c; int id, ovf; uint32_t nn, nt;
    bvh4_query((const orc_node4*)nodes, (const orc_tri4*)tris, delta, vld(pts + 3 * q),
               thr_init ? thr_init[q] : FLT_MAX, 16, &d2, &c, &id, &ovf, &nn, &nt);
    max_depth[q] = g_bvh4_maxdepth;
    pushes[q] = g_bvh4_pushes;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Learned-cache oracle in FP64: multiresolution hash-grid encoding (Mueller et
 * al. 2022: N_l = floor(N_min b^l), trilinear weights) with the vertices
 * grouped in 2 x 2 x 2 bricks of 8 entries: vertex (x, y, z) is entry
 * 8 brick + (x & 1 | (y & 1) << 1 | (z & 1) << 2) of brick (x/2, y/2, z/2),
 * indexed densely (B = N_l/2 + 1 bricks per axis) when 8 B^3 <= T, else by
 * the spatial hash bx*1 ^ by*2654435761 ^ bz*805459861 mod T/8, and
 * a bias-free ReLU MLP 2L -> H -> ... -> H -> 1 (`layers` hidden layers of
 * width H <= 64), with the gradients of the mean squared error by reverse
 * accumulation.  Weight buffer layout (orc_w_off): A1 [H x 2L] (32 columns
 * reserved), the hidden layers [H x H], the output row 0 of a 16 x H block
 * (rows 1..15 unused); for H = 64 and 4 layers: offsets 0, 2048, 6144,
 * 10240, 14336. */
typedef struct {
  int n_levels;
  uint32_t table_size;
  uint32_t offset[16], size[16];
  int res[16], dense[16];
  uint32_t total;
  double in_min[3], in_scale[3];
  int hidden, layers; /* MLP: `layers` hidden layers of width `hidden` */
} orc_grid;

/* Weight-buffer layout of the MLP with L hidden layers of width H: A1 [H x 32]
 * at 0, hidden layer k (k = 1 .. L-1) [H x H] at 32 H + (k - 1) H^2, the
 * output [16 x H] (row 0 live) at 32 H + (L - 1) H^2. */
static int orc_w_off(const orc_grid* g, int layer) {
  return layer == 0 ? 0 : 32 * g->hidden + (layer - 1) * g->hidden * g->hidden;
}
static int orc_w_out(const orc_grid* g) { return 32 * g->hidden + (g->layers - 1) * g->hidden * g->hidden; }

int orc_cache_layout(int n_levels, int log2_t, int base_res, double growth, uint32_t* offsets, uint32_t* sizes,
                     int* res, int* dense, uint32_t* total) {
  const uint32_t T = 1u << log2_t;
  uint64_t off = 0;
  for (int l = 0; l < 16; ++l) {
    const double n = floor((double)base_res * pow(growth, (double)l));
    const double nb = floor(n / 2.0) + 1.0;
    const double entries = 8.0 * nb * nb * nb;
    res[l] = (int)n;
    dense[l] = entries <= (double)T;
    sizes[l] = dense[l] ? (uint32_t)entries : T;
    offsets[l] = (uint32_t)off;
    if (l < n_levels) off += sizes[l];
  }
  *total = (uint32_t)off;
  return 0;
}

/* Entry of vertex (x, y, z) within level l (brick layout above). */
static uint32_t orc_entry(const orc_grid* g, int l, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t bx = x >> 1, by = y >> 1, bz = z >> 1;
  const uint32_t corner = (x & 1u) | ((y & 1u) << 1) | ((z & 1u) << 2);
  uint32_t brick;
  if (g->dense[l]) {
    const uint32_t nb = (uint32_t)(g->res[l] / 2) + 1u;
    brick = bx + by * nb + bz * nb * nb;
  } else {
    brick = (bx ^ (by * 2654435761u) ^ (bz * 805459861u)) & (g->table_size / 8u - 1u);
  }
  return 8u * brick + corner;
}

static void orc_corners(const orc_grid* g, int l, const double u[3], uint32_t idx[8], double w[8]) {
  const int n = g->res[l];
  int c[3];
  double f[3];
  for (int a = 0; a < 3; ++a) {
    const double s = u[a] * (double)n;
    int ci = (int)floor(s);
    ci = ci < 0 ? 0 : (ci > n - 1 ? n - 1 : ci);
    c[a] = ci;
    f[a] = s - (double)ci;
  }
  for (int k = 0; k < 8; ++k) {
    const uint32_t x = (uint32_t)(c[0] + (k & 1)), y = (uint32_t)(c[1] + ((k >> 1) & 1)),
                   z = (uint32_t)(c[2] + ((k >> 2) & 1));
    idx[k] = g->offset[l] + orc_entry(g, l, x, y, z);
    w[k] = ((k & 1) ? f[0] : 1.0 - f[0]) * (((k >> 1) & 1) ? f[1] : 1.0 - f[1]) * (((k >> 2) & 1) ? f[2] : 1.0 - f[2]);
  }
}

static void orc_unit(const orc_grid* g, const float* p, double u[3]) {
  for (int a = 0; a < 3; ++a) {
    double v = ((double)p[a] - g->in_min[a]) * g->in_scale[a];
    u[a] = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
  }
}

/* FP32 versions of the unit-cube mapping and the cell corners, with the
 * product kernels' operation order (used by the FP16-rounding variant, whose
 * encoding must round like the product's FP32 encoding). */
static void orc_unit_f32(const orc_grid* g, const float* p, float u[3]) {
  for (int a = 0; a < 3; ++a) {
    const float v = (p[a] - (float)g->in_min[a]) * (float)g->in_scale[a];
    u[a] = fminf(1.0f, fmaxf(0.0f, v));
  }
}

static void orc_corners_f32(const orc_grid* g, int l, const float u[3], uint32_t idx[8], float w[8]) {
  const int n = g->res[l];
  int c[3];
  float f[3];
  for (int a = 0; a < 3; ++a) {
    const float sa = u[a] * (float)n;
    int ci = (int)floorf(sa);
    ci = ci < 0 ? 0 : (ci > n - 1 ? n - 1 : ci);
    c[a] = ci;
    f[a] = sa - (float)ci;
  }
  for (int k = 0; k < 8; ++k) {
    const uint32_t x = (uint32_t)(c[0] + (k & 1)), y = (uint32_t)(c[1] + ((k >> 1) & 1)),
                   z = (uint32_t)(c[2] + ((k >> 2) & 1));
    idx[k] = g->offset[l] + orc_entry(g, l, x, y, z);
    const float wx = (k & 1) ? f[0] : 1.0f - f[0];
    const float wy = ((k >> 1) & 1) ? f[1] : 1.0f - f[1];
    const float wz = ((k >> 2) & 1) ? f[2] : 1.0f - f[2];
    w[k] = wx * wy * wz;
  }
}

static void orc_grid_init(orc_grid* g, int n_levels, int log2_t, int base_res, double growth, int hidden, int layers,
                          const float* in_min, const float* in_extent) {
  g->n_levels = n_levels;
  g->hidden = hidden;
  g->layers = layers;
  g->table_size = 1u << log2_t;
  orc_cache_layout(n_levels, log2_t, base_res, growth, g->offset, g->size, g->res, g->dense, &g->total);
  for (int a = 0; a < 3; ++a) {
    g->in_min[a] = in_min[a];
    g->in_scale[a] = (double)(1.0f / in_extent[a]);
  }
}

/* Round to the nearest IEEE binary16 value (ties to even; subnormals kept,
 * overflow to infinity), returned as a double.  Used by the FP16-rounding
 * variant of the cache oracle. */
double orc_round_half(double v) {
  if (v == 0.0 || !isfinite(v)) return v;
  const double a = fabs(v);
  if (a >= 65520.0) return copysign(INFINITY, v);
  int e;
  frexp(a, &e);           /* a in [2^(e-1), 2^e) */
  int ex = e - 1;
  if (ex < -14) ex = -14; /* subnormal range: fixed spacing 2^-24 */
  const double q = ldexp(1.0, ex - 10);
  return copysign(nearbyint(a / q) * q, v);
}

/* forward of one point: x[32], h[4][64], returns y.  With half != 0 the
 * weights, the encoded features and every hidden activation are rounded to
 * binary16 where the product's kernels round them (FP16 MFMA operands); sums
 * stay in double. */
static double orc_mlp_fwd_ex(const orc_grid* g, const float* grid, const float* wt, const float* p, double x[32],
                             double h[4][64], int half) {
  for (int k = 0; k < 32; ++k) x[k] = 0.0;
  if (half) {
    /* FP32 encoding in the product's order, then binary16 */
    float u[3];
    orc_unit_f32(g, p, u);
    for (int l = 0; l < g->n_levels; ++l) {
      uint32_t idx[8];
      float w[8];
      orc_corners_f32(g, l, u, idx, w);
      float f0 = 0.0f, f1 = 0.0f;
      for (int c = 0; c < 8; ++c) {
        f0 += w[c] * grid[2 * (size_t)idx[c]];
        f1 += w[c] * grid[2 * (size_t)idx[c] + 1];
      }
      x[2 * l] = orc_round_half((double)f0);
      x[2 * l + 1] = orc_round_half((double)f1);
    }
  } else {
    double u[3];
    orc_unit(g, p, u);
    for (int l = 0; l < g->n_levels; ++l) {
      uint32_t idx[8];
      double w[8];
      orc_corners(g, l, u, idx, w);
      for (int c = 0; c < 8; ++c) {
        x[2 * l] += w[c] * (double)grid[2 * (size_t)idx[c]];
        x[2 * l + 1] += w[c] * (double)grid[2 * (size_t)idx[c] + 1];
      }
    }
  }
  const double* in = x;
  int nin = 32;
  const int H = g->hidden, L = g->layers;
  for (int layer = 0; layer < L; ++layer) {
    const int off = orc_w_off(g, layer);
    for (int o = 0; o < H; ++o) {
      double z = 0.0;
      for (int k = 0; k < nin; ++k) {
        const double wk = half ? orc_round_half((double)wt[off + o * nin + k]) : (double)wt[off + o * nin + k];
        z += wk * in[k];
      }
      const double r = z > 0.0 ? z : 0.0;
      h[layer][o] = half ? orc_round_half(r) : r;
    }
    in = h[layer];
    nin = H;
  }
  const int out = orc_w_out(g);
  double y = 0.0;
  for (int k = 0; k < H; ++k)
    y += (half ? orc_round_half((double)wt[out + k]) : (double)wt[out + k]) * h[L - 1][k];
  return y;
}

static double orc_mlp_fwd(const orc_grid* g, const float* grid, const float* wt, const float* p, double x[32],
                          double h[4][64]) {
  return orc_mlp_fwd_ex(g, grid, wt, p, x, h, 0);
}

/* The hash-grid encoding alone: out[32 * s + 2 l + f] = feature f of level l
 * at point s (levels >= n_levels are 0).  f32 != 0 uses the FP32 unit-cube
 * mapping and corner weights of the product kernels, else FP64. */
int orc_cache_encode(int n_levels, int log2_t, int base_res, double growth, int hidden, int layers, const float* in_min,
                     const float* in_extent, const float* grid, const float* pts, long n, int f32, double* out) {
  orc_grid g;
  orc_grid_init(&g, n_levels, log2_t, base_res, growth, hidden, layers, in_min, in_extent);
  for (long s = 0; s < n; ++s) {
    double* x = out + 32 * s;
    for (int k = 0; k < 32; ++k) x[k] = 0.0;
    if (f32) {
      float u[3];
      orc_unit_f32(&g, pts + 3 * s, u);
      for (int l = 0; l < g.n_levels; ++l) {
        uint32_t idx[8];
        float w[8];
        orc_corners_f32(&g, l, u, idx, w);
        float f0 = 0.0f, f1 = 0.0f;
        for (int c = 0; c < 8; ++c) {
          f0 += w[c] * grid[2 * (size_t)idx[c]];
          f1 += w[c] * grid[2 * (size_t)idx[c] + 1];
        }
        x[2 * l] = f0;
        x[2 * l + 1] = f1;
      }
    } else {
      double u[3];
      orc_unit(&g, pts + 3 * s, u);
      for (int l = 0; l < g.n_levels; ++l) {
        uint32_t idx[8];
        double w[8];
        orc_corners(&g, l, u, idx, w);
        for (int c = 0; c < 8; ++c) {
          x[2 * l] += w[c] * (double)grid[2 * (size_t)idx[c]];
          x[2 * l + 1] += w[c] * (double)grid[2 * (size_t)idx[c] + 1];
        }
      }
    }
  }
  return 0;
}

int orc_cache_forward(int n_levels, int log2_t, int base_res, double growth, int hidden, int layers, const float* in_min,
                      const float* in_extent, const float* grid, const float* wt, const float* pts, long n,
                      double* out) {
  orc_grid g;
  orc_grid_init(&g, n_levels, log2_t, base_res, growth, hidden, layers, in_min, in_extent);
  double x[32], h[4][64];
  for (long s = 0; s < n; ++s) out[s] = orc_mlp_fwd(&g, grid, wt, pts + 3 * s, x, h);
  return 0;
}

/* Gradients of L = (1/n) sum (y - t)^2 w.r.t. the grid and the weights
 * (accumulated into grid_grad[2 * total] and w_grad[32 H + (L - 1) H^2 + 16 H],
 * which the caller zeroes); returns L. */
static double orc_cache_grads_ex(int n_levels, int log2_t, int base_res, double growth, int hidden, int layers,
                                 const float* in_min,
                                 const float* in_extent, const float* grid, const float* wt, const float* pts,
                                 const float* targets, long n, double* grid_grad, double* w_grad, double loss_scale) {
  /* loss_scale > 0 selects the FP16-rounding variant: dL/dy is scaled by
   * loss_scale, it and every pre-activation gradient dZ are rounded to
   * binary16 (as the product stores them for its MFMA weight-gradient GEMMs),
   * and the sums are unscaled at the end; the input-feature gradient is the
   * unrounded sum (an FP32 accumulator in the product). */
  const int half = loss_scale > 0.0;
  const double scale = half ? loss_scale : 1.0;
  orc_grid g;
  orc_grid_init(&g, n_levels, log2_t, base_res, growth, hidden, layers, in_min, in_extent);
  const int H = hidden, L = layers, out_off = orc_w_out(&g);
  double loss = 0.0;
  for (long s = 0; s < n; ++s) {
    double x[32], h[4][64];
    const double y = orc_mlp_fwd_ex(&g, grid, wt, pts + 3 * s, x, h, half);
    const double diff = y - (double)targets[s];
    loss += diff * diff;
    double dy = 2.0 * diff * (scale / (double)n);
    if (half) dy = orc_round_half(dy);
    double dh[64], dz[64];
    for (int k = 0; k < H; ++k) {
      const double w5 = half ? orc_round_half((double)wt[out_off + k]) : (double)wt[out_off + k];
      w_grad[out_off + k] += dy * h[L - 1][k] / scale;
      dh[k] = dy * w5;
    }
    for (int layer = L - 1; layer >= 0; --layer) {
      const int nin = layer == 0 ? 32 : H;
      const int off = orc_w_off(&g, layer);
      const double* in = layer == 0 ? x : h[layer - 1];
      for (int o = 0; o < H; ++o) {
        const double d = h[layer][o] > 0.0 ? dh[o] : 0.0;
        dz[o] = half ? orc_round_half(d) : d;
      }
      double din[64];
      for (int k = 0; k < nin; ++k) din[k] = 0.0;
      for (int o = 0; o < H; ++o) {
        for (int k = 0; k < nin; ++k) {
          const double wk = half ? orc_round_half((double)wt[off + o * nin + k]) : (double)wt[off + o * nin + k];
          w_grad[off + o * nin + k] += dz[o] * in[k] / scale;
          din[k] += dz[o] * wk;
        }
      }
      for (int k = 0; k < nin; ++k) dh[k] = din[k];
    }
    /* dh now holds dL/dx (times the scale) */
    double u[3];
    float uf[3];
    orc_unit(&g, pts + 3 * s, u);
    orc_unit_f32(&g, pts + 3 * s, uf);
    for (int l = 0; l < g.n_levels; ++l) {
      uint32_t idx[8];
      double w[8];
      if (half) {
        float wf[8];
        orc_corners_f32(&g, l, uf, idx, wf);
        for (int c = 0; c < 8; ++c) w[c] = (double)wf[c];
      } else {
        orc_corners(&g, l, u, idx, w);
      }
      for (int c = 0; c < 8; ++c) {
        grid_grad[2 * (size_t)idx[c]] += w[c] * dh[2 * l] / scale;
        grid_grad[2 * (size_t)idx[c] + 1] += w[c] * dh[2 * l + 1] / scale;
      }
    }
  }
  return loss / (double)n;
}

double orc_cache_grads(int n_levels, int log2_t, int base_res, double growth, int hidden, int layers, const float* in_min,
                       const float* in_extent, const float* grid, const float* wt, const float* pts,
                       const float* targets, long n, double* grid_grad, double* w_grad) {
  return orc_cache_grads_ex(n_levels, log2_t, base_res, growth, hidden, layers, in_min, in_extent, grid, wt, pts, targets, n,
                            grid_grad, w_grad, 0.0);
}

/* The same gradients with the product's FP32 encoding and binary16 rounding
 * points (weights, features, activations, scaled dL/dy and dZ); loss_scale as
 * the product's. */
double orc_cache_grads_fp16(int n_levels, int log2_t, int base_res, double growth, int hidden, int layers, const float* in_min,
                            const float* in_extent, const float* grid, const float* wt, const float* pts,
                            const float* targets, long n, double* grid_grad, double* w_grad, double loss_scale) {
  return orc_cache_grads_ex(n_levels, log2_t, base_res, growth, hidden, layers, in_min, in_extent, grid, wt, pts, targets, n,
                            grid_grad, w_grad, loss_scale);
}

int orc_cache_forward_fp16(int n_levels, int log2_t, int base_res, double growth, int hidden, int layers, const float* in_min,
                           const float* in_extent, const float* grid, const float* wt, const float* pts, long n,
                           double* out) {
  orc_grid g;
  orc_grid_init(&g, n_levels, log2_t, base_res, growth, hidden, layers, in_min, in_extent);
  double x[32], h[4][64];
  for (long s = 0; s < n; ++s) out[s] = orc_mlp_fwd_ex(&g, grid, wt, pts + 3 * s, x, h, 1);
  return 0;
}
