// cache.hip -- learned cache kernels: hash-grid encoding, MLP forward and
// backward on MFMA (FP16 operands, FP32 accumulation), weight gradients as
// MFMA GEMMs over the batch, and Adam.  Layout and formulas: cache.hpp.
//
// MFMA fragment layout of v_mfma_f32_16x16x16_f16 (one wave, lane l, i = l % 16,
// g = l / 16): A operand holds A[i][4g .. 4g+3], B operand B[4g .. 4g+3][i],
// the accumulator D[4g .. 4g+3][i].  The MLP is evaluated transposed
// (features x samples), so one layer's accumulator IS the next layer's B
// operand: lane l holds features 4g..4g+3 of sample i of every 16-feature tile.
// The hash-grid encoding is spread the same way: lane (i, g) encodes levels
// 2g, 2g+1, 8+2g, 8+2g+1 of sample i, exactly the input features its B
// fragments need, so no lane computes a level twice.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "api_internal.hpp"
#include "cache.hpp"
#include "geom.hpp"
#include "wave.hpp"

typedef _Float16 h4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

namespace n2 {

namespace {

__device__ __forceinline__ f4_t mfma16(h4_t a, h4_t b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x32_f16 (gfx950): the same fragments with K = 32, lane
// (i, g) holding A[i][8g .. 8g+7] and B[8g .. 8g+7][i], the accumulator as
// above.  Used where K is summed over as a whole (the weight gradients), so
// only the pairing of A's and B's K elements matters.
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f4_t mfma32(h8_t a, h8_t b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ h4_t relu_h4(f4_t v) {
  h4_t r;
  r[0] = static_cast<_Float16>(fmaxf(v[0], 0.0f));
  r[1] = static_cast<_Float16>(fmaxf(v[1], 0.0f));
  r[2] = static_cast<_Float16>(fmaxf(v[2], 0.0f));
  r[3] = static_cast<_Float16>(fmaxf(v[3], 0.0f));
  return r;
}

__device__ __forceinline__ f4_t zero4() {
  f4_t z = {0.0f, 0.0f, 0.0f, 0.0f};
  return z;
}

// Entry of vertex (x, y, z) within level l: the vertex's corner in its
// 2 x 2 x 2 brick (x & 1 | (y & 1) << 1 | (z & 1) << 2) plus 8 times the
// brick's index, dense (bx + B by + B^2 bz, B = N/2 + 1 bricks per axis) or
// the spatial hash of the brick coordinates mod T / 8.
__host__ __device__ __forceinline__ uint32_t brick_entry(const GridLayout& g, int l, uint32_t x, uint32_t y,
                                                         uint32_t z) {
  const uint32_t bx = x >> 1, by = y >> 1, bz = z >> 1;
  const uint32_t corner = (x & 1u) | ((y & 1u) << 1) | ((z & 1u) << 2);
  uint32_t brick;
  if (g.dense[l]) {
    const uint32_t nb = static_cast<uint32_t>(g.res[l] >> 1) + 1u;
    brick = bx + by * nb + bz * nb * nb;
  } else {
    brick = (bx ^ (by * 2654435761u) ^ (bz * 805459861u)) & ((g.table_size >> 3) - 1u);
  }
  return brick * 8u + corner;
}

// Cell corner indices and trilinear weights of point u (unit cube) on level l.
__host__ __device__ __forceinline__ void level_corners(const GridLayout& g, int l, float ux, float uy, float uz,
                                                       uint32_t idx[8], float w[8]) {
  const int n = g.res[l];
  const float sx = ux * static_cast<float>(n), sy = uy * static_cast<float>(n), sz = uz * static_cast<float>(n);
  int cx = static_cast<int>(floorf(sx)), cy = static_cast<int>(floorf(sy)), cz = static_cast<int>(floorf(sz));
  cx = cx < 0 ? 0 : (cx > n - 1 ? n - 1 : cx);
  cy = cy < 0 ? 0 : (cy > n - 1 ? n - 1 : cy);
  cz = cz < 0 ? 0 : (cz > n - 1 ? n - 1 : cz);
  const float fx = sx - static_cast<float>(cx), fy = sy - static_cast<float>(cy), fz = sz - static_cast<float>(cz);
  for (int c = 0; c < 8; ++c) {
    const uint32_t x = static_cast<uint32_t>(cx + (c & 1));
    const uint32_t y = static_cast<uint32_t>(cy + ((c >> 1) & 1));
    const uint32_t z = static_cast<uint32_t>(cz + ((c >> 2) & 1));
    idx[c] = g.offset[l] + brick_entry(g, l, x, y, z);
    w[c] = ((c & 1) ? fx : 1.0f - fx) * (((c >> 1) & 1) ? fy : 1.0f - fy) * (((c >> 2) & 1) ? fz : 1.0f - fz);
  }
}

__host__ __device__ __forceinline__ void to_unit(const GridLayout& g, float px, float py, float pz, float& ux,
                                                 float& uy, float& uz) {
  ux = fminf(1.0f, fmaxf(0.0f, (px - g.in_min[0]) * g.in_scale[0]));
  uy = fminf(1.0f, fmaxf(0.0f, (py - g.in_min[1]) * g.in_scale[1]));
  uz = fminf(1.0f, fmaxf(0.0f, (pz - g.in_min[2]) * g.in_scale[2]));
}

// The two input fragments of lane (i, g): features of levels {2g, 2g+1} (t=0)
// and {8+2g, 9+2g} (t=1), two features each.
__device__ __forceinline__ void encode_fragments(const GridLayout& gl, const float* __restrict__ grid, float ux,
                                                 float uy, float uz, int g, h4_t xb[2]) {
  for (int t = 0; t < 2; ++t) {
    float f[4];
    for (int h = 0; h < 2; ++h) {
      const int l = 8 * t + 2 * g + h;
      float f0 = 0.0f, f1 = 0.0f;
      if (l < gl.n_levels) {
        uint32_t idx[8];
        float w[8];
        level_corners(gl, l, ux, uy, uz, idx, w);
        for (int c = 0; c < 8; ++c) {
          const float2 v = *reinterpret_cast<const float2*>(grid + 2 * static_cast<size_t>(idx[c]));
          f0 += w[c] * v.x;
          f1 += w[c] * v.y;
        }
      }
      f[2 * h] = f0;
      f[2 * h + 1] = f1;
    }
    xb[t][0] = static_cast<_Float16>(f[0]);
    xb[t][1] = static_cast<_Float16>(f[1]);
    xb[t][2] = static_cast<_Float16>(f[2]);
    xb[t][3] = static_cast<_Float16>(f[3]);
  }
}

__device__ __forceinline__ h4_t load_h4(const _Float16* p) { return *reinterpret_cast<const h4_t*>(p); }
__device__ __forceinline__ h8_t load_h8(const _Float16* p) { return *reinterpret_cast<const h8_t*>(p); }

// A fragment of a row-major [rows x cols] FP16 matrix: row r0 + i, cols k0 + 4g ..
__device__ __forceinline__ h4_t frag_rows(const _Float16* m, int cols, int r, int k) {
  return load_h4(m + r * cols + k);
}

// A fragment of the TRANSPOSE of a row-major [rows x cols] matrix: element
// (r, k) of M^T is M[k][r].
__device__ __forceinline__ h4_t frag_trans(const _Float16* m, int cols, int r, int k) {
  h4_t a;
  a[0] = m[(k + 0) * cols + r];
  a[1] = m[(k + 1) * cols + r];
  a[2] = m[(k + 2) * cols + r];
  a[3] = m[(k + 3) * cols + r];
  return a;
}

// Value of v in lane S of this lane's row of 16 lanes (DPP row_newbcast).
template <int S>
__device__ __forceinline__ float row_lane(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x150 + S, 0xF, 0xF, false));
}

// The four levels one row of the wave scatters (lane (i, g) holds the input
// gradients of levels 2g, 2g+1, 8+2g, 9+2g of sample i): level, resolution,
// dense flag, first grid float (2 x first entry) and bricks per axis.
struct RowLevels {
  int l[4];
  int n[4];
  int dense[4];
  uint32_t base2[4];
  uint32_t nb[4];
};

__device__ __forceinline__ RowLevels row_levels(const GridLayout& gl, int g) {
  RowLevels r;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int l = 8 * (k >> 1) + 2 * g + (k & 1);
    const bool on = l < gl.n_levels;
    r.l[k] = on ? l : -1;
    r.n[k] = on ? gl.res[l] : 1;
    r.dense[k] = on ? gl.dense[l] : 0;
    r.base2[k] = on ? 2u * gl.offset[l] : 0u;
    r.nb[k] = static_cast<uint32_t>(r.n[k] >> 1) + 1u;
  }
  return r;
}

// Grid-gradient scatter of sample S of the tile at one level per row: the
// row adds w_c g_f of the sample's cell into the 2 x 2 x 2 bricks the cell
// touches (1, 2, 4 or 8: two along an axis where the cell starts at an odd
// vertex).  Lane r of the row stands for dword r of a brick's 64 bytes
// (corner r >> 1, feature r & 1).  Along each axis exactly one of the cell's
// bricks holds a cell vertex at this lane's corner: for an even first vertex
// c the brick c/2 (vertex c + o for corner offset o), for an odd one brick
// (c - 1)/2 + 1 - o (vertex c + 1 - o).  So every lane adds exactly one
// value, one of the cell's 8 x 2, and the row's one atomic instruction sends
// one 64-byte request per brick it touches: the memory-side atomic rate is
// set by requests, whatever their size up to 64 bytes (21 G requests/s on
// MI355X for 1 to 16 dwords).  The corner weights are level_corners' own
// products.
template <int S>
__device__ __forceinline__ void scatter_sample(const GridLayout& gl, const RowLevels& rl, float ux, float uy, float uz,
                                               float live, const f4_t (&dx)[2], int r,
                                               float* __restrict__ grid_grad) {
  const float px = row_lane<S>(ux), py = row_lane<S>(uy), pz = row_lane<S>(uz);
  const float pl = row_lane<S>(live);
  const uint32_t rc = static_cast<uint32_t>(r >> 1);
  const uint32_t ox = rc & 1u, oy = (rc >> 1) & 1u, oz = rc >> 2;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float g0 = row_lane<S>(dx[k >> 1][2 * (k & 1)]);
    const float g1 = row_lane<S>(dx[k >> 1][2 * (k & 1) + 1]);
    if (rl.l[k] < 0 || pl == 0.0f || (g0 == 0.0f && g1 == 0.0f)) continue;
    const int n = rl.n[k];
    const float nf = static_cast<float>(n);
    const float sx = px * nf, sy = py * nf, sz = pz * nf;
    int cx = static_cast<int>(floorf(sx)), cy = static_cast<int>(floorf(sy)), cz = static_cast<int>(floorf(sz));
    cx = cx < 0 ? 0 : (cx > n - 1 ? n - 1 : cx);
    cy = cy < 0 ? 0 : (cy > n - 1 ? n - 1 : cy);
    cz = cz < 0 ? 0 : (cz > n - 1 ? n - 1 : cz);
    const float fx = sx - static_cast<float>(cx), fy = sy - static_cast<float>(cy), fz = sz - static_cast<float>(cz);
    // this lane's cell vertex along each axis: offset d (0 or 1) from the
    // first vertex, and the brick that holds it
    const uint32_t ux0 = static_cast<uint32_t>(cx), uy0 = static_cast<uint32_t>(cy), uz0 = static_cast<uint32_t>(cz);
    const uint32_t dxo = (ux0 & 1u) ? 1u - ox : ox, dyo = (uy0 & 1u) ? 1u - oy : oy, dzo = (uz0 & 1u) ? 1u - oz : oz;
    const uint32_t bx = (ux0 + dxo) >> 1, by = (uy0 + dyo) >> 1, bz = (uz0 + dzo) >> 1;
    const float wx = dxo ? fx : 1.0f - fx;
    const float wy = dyo ? fy : 1.0f - fy;
    const float wz = dzo ? fz : 1.0f - fz;
    const float gf = (r & 1) ? g1 : g0;
    const uint32_t brick = rl.dense[k] ? bx + by * rl.nb[k] + bz * rl.nb[k] * rl.nb[k]
                                       : (bx ^ (by * 2654435761u) ^ (bz * 805459861u)) & ((gl.table_size >> 3) - 1u);
    atomicAdd(grid_grad + rl.base2[k] + 16u * static_cast<size_t>(brick) + static_cast<uint32_t>(r),
              wx * wy * wz * gf);
  }
}

template <int... S>
__device__ __forceinline__ void scatter_tile(const GridLayout& gl, const RowLevels& rl, float ux, float uy, float uz,
                                             float live, const f4_t (&dx)[2], int r, float* __restrict__ grid_grad,
                                             std::integer_sequence<int, S...>) {
  (scatter_sample<S>(gl, rl, ux, uy, uz, live, dx, r, grid_grad), ...);
}

// Forward of one 16-sample tile through the MLP with L hidden layers of width
// H (T = H / 16 tiles per layer); returns the output accumulator (row 0 of
// the padded output layer, held in element 0 of lanes with g == 0) and the
// hidden activations h[layer][tile].  Every MFMA here and below runs with the
// whole wave active: the instruction reads all 64 lanes' operands.
template <int H, int L>
__device__ __forceinline__ f4_t mlp_forward(const _Float16* __restrict__ w, const h4_t xb[2], int i, int g,
                                            h4_t (&h)[L][H / 16]) {
  constexpr int T = H / 16;
#pragma unroll
  for (int o = 0; o < T; ++o) {
    f4_t acc = zero4();
#pragma unroll
    for (int t = 0; t < 2; ++t) acc = mfma16(frag_rows(w, kCacheIn, 16 * o + i, 16 * t + 4 * g), xb[t], acc);
    h[0][o] = relu_h4(acc);
  }
#pragma unroll
  for (int k = 1; k < L; ++k) {
    const _Float16* a = w + cache_off_hidden(H, k);
#pragma unroll
    for (int o = 0; o < T; ++o) {
      f4_t acc = zero4();
#pragma unroll
      for (int t = 0; t < T; ++t) acc = mfma16(frag_rows(a, H, 16 * o + i, 16 * t + 4 * g), h[k - 1][t], acc);
      h[k][o] = relu_h4(acc);
    }
  }
  const _Float16* ao = w + cache_off_out(H, L);
  f4_t out = zero4();
#pragma unroll
  for (int t = 0; t < T; ++t) out = mfma16(frag_rows(ao, H, i, 16 * t + 4 * g), h[L - 1][t], out);
  return out;
}

constexpr int kCacheBlock = 256;

__device__ __forceinline__ void stage_weights(_Float16* lds, const _Float16* __restrict__ w, int count) {
  for (int k = threadIdx.x; k < count / 4; k += blockDim.x)
    reinterpret_cast<h4_t*>(lds)[k] = reinterpret_cast<const h4_t*>(w)[k];
  __syncthreads();
}

// Inference: out[s] = C(points[s]) (+ the output offset, when set) for s < n.
template <int H, int L>
__global__ void __launch_bounds__(kCacheBlock) k_cache_infer(GridLayout gl, const float* __restrict__ grid,
                                                            const _Float16* __restrict__ wh,
                                                            const float* __restrict__ points, int n,
                                                            float offset, float* __restrict__ out) {
  __shared__ _Float16 w[cache_weights(H, L)];
  stage_weights(w, wh, cache_weights(H, L));
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int i = lane & 15, g = lane >> 4;
  const int waves = static_cast<int>(gridDim.x * blockDim.x / 64);
  const int wave = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) / 64);
  const int tiles = (n + 15) / 16;
  for (int tile = wave; tile < tiles; tile += waves) {
    const int s = tile * 16 + i;
    float ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (s < n) to_unit(gl, points[3 * s], points[3 * s + 1], points[3 * s + 2], ux, uy, uz);
    h4_t xb[2];
    encode_fragments(gl, grid, ux, uy, uz, g, xb);
    h4_t h[L][H / 16];
    const f4_t y = mlp_forward<H, L>(w, xb, i, g, h);
    if (g == 0 && s < n) out[s] = offset != 0.0f ? y[0] + offset : y[0];
  }
}

// Training pass 1: forward + backward of every 16-sample tile.  Stores the
// layer inputs X, H1..H_L and the pre-activation gradients dZ1..dZ_L, dY
// (FP16, [feature][batch]) for the weight-gradient GEMMs, scatters the
// input-feature gradients into the grid gradient, and adds the loss to
// *loss_sum.
struct TrainStore {
  _Float16* x;    // [32][B]
  _Float16* h;    // [L][H][B]  H1..H_L
  _Float16* dz;   // [L][H][B]  dZ1..dZ_L
  _Float16* dy;   // [B]        dL/dy * loss_scale
  int batch;      // B (multiple of 16)
};

__device__ __forceinline__ void store_frag(_Float16* base, int batch, int feat0, int s, h4_t v) {
  base[static_cast<size_t>(feat0 + 0) * batch + s] = v[0];
  base[static_cast<size_t>(feat0 + 1) * batch + s] = v[1];
  base[static_cast<size_t>(feat0 + 2) * batch + s] = v[2];
  base[static_cast<size_t>(feat0 + 3) * batch + s] = v[3];
}

__device__ __forceinline__ h4_t mask_relu(f4_t d, h4_t h) {
  h4_t r;
  r[0] = static_cast<_Float16>(static_cast<float>(h[0]) > 0.0f ? d[0] : 0.0f);
  r[1] = static_cast<_Float16>(static_cast<float>(h[1]) > 0.0f ? d[1] : 0.0f);
  r[2] = static_cast<_Float16>(static_cast<float>(h[2]) > 0.0f ? d[2] : 0.0f);
  r[3] = static_cast<_Float16>(static_cast<float>(h[3]) > 0.0f ? d[3] : 0.0f);
  return r;
}

// Grid gradients go to copy (xcd_id & copy_mask) of grid_grad (copies
// gstride floats apart), summed by Adam: small grids get one copy per XCD so
// that fewer atomics meet on each address (copy_mask 7), large ones one copy
// (copy_mask 0).  Every copy is updated with device-scope atomics, so the sum
// is right whichever XCD a block runs on.
template <int H, int L>
__global__ void __launch_bounds__(kCacheBlock) k_cache_train_pass1(GridLayout gl, const float* __restrict__ grid,
                                                                  const _Float16* __restrict__ wh,
                                                                  const float* __restrict__ points,
                                                                  const float* __restrict__ targets, int n,
                                                                  float offset, float grad_scale, TrainStore st,
                                                                  float* __restrict__ grid_grad_copies,
                                                                  float* __restrict__ loss_sum, uint32_t copy_mask,
                                                                  size_t gstride, int* __restrict__ bad) {
  constexpr int T = H / 16;
  float* __restrict__ grid_grad = grid_grad_copies + (xcd_id() & copy_mask) * gstride;
  __shared__ _Float16 w[cache_weights(H, L)];
  stage_weights(w, wh, cache_weights(H, L));
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int i = lane & 15, g = lane >> 4;
  const int waves = static_cast<int>(gridDim.x * blockDim.x / 64);
  const int wave = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) / 64);
  const int tiles = st.batch / 16;
  const size_t HB = static_cast<size_t>(H) * st.batch;
  const RowLevels rl = row_levels(gl, g);
  float loss = 0.0f;
  bool finite = true;  // every dX value this lane scattered (and so every grid-gradient term)
  for (int tile = wave; tile < tiles; tile += waves) {
    const int s = tile * 16 + i;
    const bool live = s < n;
    float ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (live) to_unit(gl, points[3 * s], points[3 * s + 1], points[3 * s + 2], ux, uy, uz);
    h4_t xb[2];
    encode_fragments(gl, grid, ux, uy, uz, g, xb);
    h4_t h[L][T];
    const f4_t y = mlp_forward<H, L>(w, xb, i, g, h);
    // dL/dy for L = mean (y - t)^2, scaled; broadcast from lane (i, 0) to (i, g)
    float dy = 0.0f;
    if (g == 0 && live) {
      const float diff = (offset != 0.0f ? y[0] + offset : y[0]) - targets[s];
      loss += diff * diff;
      dy = 2.0f * diff * grad_scale;
    }
    dy = __shfl(dy, i);
    // stores for the weight gradients
    for (int t = 0; t < 2; ++t) store_frag(st.x, st.batch, 16 * t + 4 * g, s, xb[t]);
#pragma unroll
    for (int k = 0; k < L; ++k)
#pragma unroll
      for (int o = 0; o < T; ++o) store_frag(st.h + k * HB, st.batch, 16 * o + 4 * g, s, h[k][o]);
    if (g == 0) st.dy[s] = static_cast<_Float16>(dy);
    // output layer: dH_L = A_out^T dY, dY = [dy; 0; ...] (16 x 16)
    h4_t bdy;
    bdy[0] = static_cast<_Float16>(g == 0 ? dy : 0.0f);
    bdy[1] = bdy[2] = bdy[3] = static_cast<_Float16>(0.0f);
    h4_t dz[T], dzp[T];
    const _Float16* ao = w + cache_off_out(H, L);
#pragma unroll
    for (int o = 0; o < T; ++o) {
      const f4_t acc = mfma16(frag_trans(ao, H, 16 * o + i, 4 * g), bdy, zero4());
      dz[o] = mask_relu(acc, h[L - 1][o]);
      store_frag(st.dz + (L - 1) * HB, st.batch, 16 * o + 4 * g, s, dz[o]);
    }
    // hidden layers, top down: dZ_k = relu'(H_k) (A_{k+1}^T dZ_{k+1})
#pragma unroll
    for (int k = L - 1; k >= 1; --k) {
      const _Float16* a = w + cache_off_hidden(H, k);
#pragma unroll
      for (int o = 0; o < T; ++o) {
        f4_t acc = zero4();
#pragma unroll
        for (int t = 0; t < T; ++t) acc = mfma16(frag_trans(a, H, 16 * o + i, 16 * t + 4 * g), dz[t], acc);
        dzp[o] = mask_relu(acc, h[k - 1][o]);
        store_frag(st.dz + (k - 1) * HB, st.batch, 16 * o + 4 * g, s, dzp[o]);
      }
#pragma unroll
      for (int o = 0; o < T; ++o) dz[o] = dzp[o];
    }
    // input-feature gradients dX = A1^T dZ1 (32 x 16), computed by the whole
    // wave (an MFMA reads every lane's operands), then scattered to the grid
    // brick by brick, one sample and one level per row at a time
    f4_t dx[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      f4_t acc = zero4();
#pragma unroll
      for (int k = 0; k < T; ++k) acc = mfma16(frag_trans(w, kCacheIn, 16 * t + i, 16 * k + 4 * g), dz[k], acc);
      dx[t] = acc;
      finite = finite && isfinite(acc[0]) && isfinite(acc[1]) && isfinite(acc[2]) && isfinite(acc[3]);
    }
    scatter_tile(gl, rl, ux, uy, uz, live ? 1.0f : 0.0f, dx, i, grid_grad, std::make_integer_sequence<int, 16>{});
  }
  // loss: sum over the wave's lanes (only g == 0 lanes carry values)
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
  if (lane == 0) atomicAdd(loss_sum, loss);
  if (!finite) *bad = 1;
}

// Training pass 2: weight gradients dA = dZ H_prev^T over the batch (K =
// samples) on MFMA.  blockIdx.y = layer (0 = input layer, 1 .. L-1 hidden,
// L = output); each wave reduces a K-chunk into ROWS_T x COLS_T accumulator
// tiles, 32 samples per MFMA (16-byte loads; a chunk's last 16 samples, when
// the batch is an odd multiple of 16, with the 16-deep MFMA), and adds them
// to the FP32 gradient with atomics.  ROWS_T = 0 marks the output layer,
// whose dZ is the single row dY (rows 1..15 of the padded tile are zero).
template <int ROWS_T, int COLS_T>
__device__ __forceinline__ void wgrad_layer(const _Float16* __restrict__ dz, const _Float16* __restrict__ hp,
                                            size_t B, int k0, int k1, int i, int g, int cols, float* __restrict__ wg,
                                            int* __restrict__ bad) {
  constexpr int R = ROWS_T == 0 ? 1 : ROWS_T;
  f4_t acc[R * COLS_T];
#pragma unroll
  for (int q = 0; q < R * COLS_T; ++q) acc[q] = zero4();
  int k = k0;
  for (; k + 32 <= k1; k += 32) {
    h8_t bf[COLS_T];
#pragma unroll
    for (int c = 0; c < COLS_T; ++c) bf[c] = load_h8(hp + static_cast<size_t>(16 * c + i) * B + k + 8 * g);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      h8_t af;
      if (ROWS_T == 0) {
        const h8_t v = load_h8(dz + k + 8 * g);
#pragma unroll
        for (int e = 0; e < 8; ++e) af[e] = i == 0 ? v[e] : static_cast<_Float16>(0.0f);
      } else {
        af = load_h8(dz + static_cast<size_t>(16 * r + i) * B + k + 8 * g);
      }
#pragma unroll
      for (int c = 0; c < COLS_T; ++c) acc[r * COLS_T + c] = mfma32(af, bf[c], acc[r * COLS_T + c]);
    }
  }
  if (k < k1) {  // 16 samples left
    h4_t bf[COLS_T];
#pragma unroll
    for (int c = 0; c < COLS_T; ++c) bf[c] = load_h4(hp + static_cast<size_t>(16 * c + i) * B + k + 4 * g);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      h4_t af;
      if (ROWS_T == 0) {
        const h4_t v = load_h4(dz + k + 4 * g);
        const _Float16 z = static_cast<_Float16>(0.0f);
        af[0] = i == 0 ? v[0] : z;
        af[1] = i == 0 ? v[1] : z;
        af[2] = i == 0 ? v[2] : z;
        af[3] = i == 0 ? v[3] : z;
      } else {
        af = load_h4(dz + static_cast<size_t>(16 * r + i) * B + k + 4 * g);
      }
#pragma unroll
      for (int c = 0; c < COLS_T; ++c) acc[r * COLS_T + c] = mfma16(af, bf[c], acc[r * COLS_T + c]);
    }
  }
  // acc[r * COLS_T + c] holds D[4g + e][i] of tile (r, c): row 16r + 4g + e, col 16c + i
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int c = 0; c < COLS_T; ++c)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = 16 * r + 4 * g + e;
        if (ROWS_T == 0 && row != 0) continue;
        if (!isfinite(acc[r * COLS_T + c][e])) *bad = 1;
        atomicAdd(wg + row * cols + 16 * c + i, acc[r * COLS_T + c][e]);
      }
}

template <int H, int L>
__global__ void __launch_bounds__(kCacheBlock) k_cache_wgrad(TrainStore st, int chunk, float* __restrict__ w_grad,
                                                             int* __restrict__ bad) {
  constexpr int T = H / 16;
  const int layer = static_cast<int>(blockIdx.y);
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int i = lane & 15, g = lane >> 4;
  const int wave = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) / 64);
  const int k0 = wave * chunk;
  if (k0 >= st.batch) return;
  const int k1 = min(st.batch, k0 + chunk);
  const size_t B = static_cast<size_t>(st.batch);
  const size_t HB = static_cast<size_t>(H) * B;
  if (layer == 0) {
    wgrad_layer<T, 2>(st.dz, st.x, B, k0, k1, i, g, kCacheIn, w_grad, bad);
  } else if (layer < L) {
    wgrad_layer<T, T>(st.dz + layer * HB, st.h + (layer - 1) * HB, B, k0, k1, i, g, H,
                      w_grad + cache_off_hidden(H, layer), bad);
  } else {
    wgrad_layer<0, T>(st.dy, st.h + (L - 1) * HB, B, k0, k1, i, g, H, w_grad + cache_off_out(H, L), bad);
  }
}

// One Adam step (Kingma & Ba) on FP32 parameters with unscaled gradients:
// the weights in blocks [0, wblocks), which also write their FP16 copy for
// the MFMA kernels, and the grid in the other blocks, whose gradient is the
// sum of its copies in copy order; the gradients are zeroed as they are read.
// Skipped (the gradients still zeroed) when *bad is set.
struct AdamArgs {
  float *w, *w_grad, *w_m, *w_v;
  _Float16* wh;
  int n_w;
  float *grid, *grid_grad, *grid_m, *grid_v;
  size_t n_grid;
  int copies;
  float lr_t, beta1, beta2, eps, inv_scale;
  const int* bad;
  int wblocks;
};

__device__ __forceinline__ float adam_update(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                             size_t k, float gk, const AdamArgs& a) {
  const float mk = a.beta1 * m[k] + (1.0f - a.beta1) * gk;
  const float vk = a.beta2 * v[k] + (1.0f - a.beta2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  const float pk = p[k] - a.lr_t * mk / (sqrtf(vk) + a.eps);
  p[k] = pk;
  return pk;
}

__global__ void k_adam_step(AdamArgs a) {
  const bool skip = *a.bad != 0;
  if (static_cast<int>(blockIdx.x) < a.wblocks) {
    for (int k = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); k < a.n_w; k += a.wblocks * blockDim.x) {
      const float gk = a.w_grad[k] * a.inv_scale;
      a.w_grad[k] = 0.0f;
      if (skip) continue;
      a.wh[k] = static_cast<_Float16>(adam_update(a.w, a.w_m, a.w_v, static_cast<size_t>(k), gk, a));
    }
    return;
  }
  const size_t stride = static_cast<size_t>(gridDim.x - a.wblocks) * blockDim.x;
  for (size_t k = (blockIdx.x - a.wblocks) * static_cast<size_t>(blockDim.x) + threadIdx.x; k < a.n_grid;
       k += stride) {
    float gs = a.grid_grad[k];
    a.grid_grad[k] = 0.0f;
    for (int c = 1; c < a.copies; ++c) {
      gs += a.grid_grad[c * a.n_grid + k];
      a.grid_grad[c * a.n_grid + k] = 0.0f;
    }
    if (skip) continue;
    adam_update(a.grid, a.grid_m, a.grid_v, k, gs * a.inv_scale, a);
  }
}

__global__ void k_to_half(const float* __restrict__ a, _Float16* __restrict__ h, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) h[k] = static_cast<_Float16>(a[k]);
}

__global__ void k_zero_two(int* p, int* q) {
  *p = 0;
  *q = 0;
}

__global__ void k_init_uniform(float* __restrict__ p, size_t n, float lo, float hi, uint32_t seed, uint32_t tag) {
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const U4 r = philox4x32_10(U4{static_cast<uint32_t>(k), static_cast<uint32_t>(k >> 32), tag, 0x63616368u}, seed,
                               0x9e3779b9u);
    p[k] = lo + (hi - lo) * u01(r.x);
  }
}

}  // namespace
}  // namespace n2

// ------------------------------------------------------------------ handle
struct n2wos_cache {
  n2::GridLayout layout{};
  float* grid = nullptr;
  float* grid_grad = nullptr;  // grad_copies copies of the grid's size
  int grad_copies = 1;
  float* grid_m = nullptr;
  float* grid_v = nullptr;
  float* w = nullptr;
  float* w_grad = nullptr;
  float* w_m = nullptr;
  float* w_v = nullptr;
  _Float16* wh = nullptr;
  int* bad = nullptr;
  float* loss = nullptr;
  _Float16* store = nullptr;
  int store_batch = 0;
  int64_t step = 0;
  float lr = 1e-2f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-15f, loss_scale = 128.0f;
  float out_offset = 0.0f;  // added to the network's output (inference and loss)
  int hidden = 64;            // hidden width H (16, 32 or 64)
  int layers = 4;             // hidden layers L (1 .. 4)
  int n_w = n2::cache_weights(64, 4);
};

using namespace n2;

namespace {

size_t grid_floats(const n2wos_cache* c) { return 2 * static_cast<size_t>(c->layout.total_entries); }

// Calls f(std::integral_constant<int, H>, std::integral_constant<int, L>) for
// the cache's MLP shape: the kernels are compiled for H in {16, 32, 64} and
// L in {1, 2, 3, 4}.
template <int H, typename F>
void with_depth(int layers, F&& f) {
  switch (layers) {
    case 1: f(std::integral_constant<int, H>{}, std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, H>{}, std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, H>{}, std::integral_constant<int, 3>{}); break;
    default: f(std::integral_constant<int, H>{}, std::integral_constant<int, 4>{}); break;
  }
}

template <typename F>
void with_shape(const n2wos_cache* c, F&& f) {
  switch (c->hidden) {
    case 16: with_depth<16>(c->layers, f); break;
    case 32: with_depth<32>(c->layers, f); break;
    default: with_depth<64>(c->layers, f); break;
  }
}

// Activation store of a batch of B (multiple of 16) samples: X [32][B], H and
// dZ [L][H][B] each, dY [B], all FP16.
size_t store_halves(int hidden, int layers, size_t batch) {
  return batch * (static_cast<size_t>(kCacheIn) + 2 * static_cast<size_t>(layers) * hidden + 1);
}

TrainStore make_store(const n2wos_cache* c, int batch) {
  TrainStore st;
  st.batch = batch;
  st.x = c->store;
  st.h = st.x + static_cast<size_t>(kCacheIn) * batch;
  st.dz = st.h + static_cast<size_t>(c->layers) * c->hidden * batch;
  st.dy = st.dz + static_cast<size_t>(c->layers) * c->hidden * batch;
  return st;
}

void launch_pass1(const n2wos_cache* c, const float* d_points, const float* d_targets, int n, float grad_scale,
                  const TrainStore& st, hipStream_t s) {
  const int tiles = st.batch / 16;
  const int blocks = std::max(1, std::min(2048, (tiles + 3) / 4));
  with_shape(c, [&](auto HC, auto LC) {
    hipLaunchKernelGGL((k_cache_train_pass1<decltype(HC)::value, decltype(LC)::value>), dim3(blocks),
                       dim3(kCacheBlock), 0, s, c->layout, c->grid, c->wh, d_points, d_targets, n, c->out_offset,
                       grad_scale, st, c->grid_grad, c->loss, static_cast<uint32_t>(c->grad_copies - 1),
                       grid_floats(c), c->bad);
  });
}

void launch_wgrad(const n2wos_cache* c, const TrainStore& st, hipStream_t s) {
  const int chunk = 1024;
  const int waves = (st.batch + chunk - 1) / chunk;
  with_shape(c, [&](auto HC, auto LC) {
    hipLaunchKernelGGL((k_cache_wgrad<decltype(HC)::value, decltype(LC)::value>),
                       dim3((waves + 3) / 4, decltype(LC)::value + 1), dim3(kCacheBlock), 0, s, st, chunk, c->w_grad,
                       c->bad);
  });
}

void free_cache(n2wos_cache* c) {
  for (void* p : {static_cast<void*>(c->grid), static_cast<void*>(c->grid_grad), static_cast<void*>(c->grid_m),
                  static_cast<void*>(c->grid_v), static_cast<void*>(c->w), static_cast<void*>(c->w_grad),
                  static_cast<void*>(c->w_m), static_cast<void*>(c->w_v), static_cast<void*>(c->wh),
                  static_cast<void*>(c->bad), static_cast<void*>(c->loss), static_cast<void*>(c->store)})
    if (p) (void)hipFree(p);
  delete c;
}

}  // namespace

extern "C" int n2wos_cache_create(const n2wos_cache_config* cfg, n2wos_cache** out) {
  return guarded([&]() -> int {
    N2_REQUIRE(cfg && out, N2WOS_ERR_NULL, "NULL argument");
    *out = nullptr;
    N2_REQUIRE(cfg->n_levels >= 1 && cfg->n_levels <= kCacheLevels, N2WOS_ERR_INVALID, "n_levels must be in [1, 16]");
    N2_REQUIRE(cfg->log2_table_size >= 8 && cfg->log2_table_size <= 24, N2WOS_ERR_INVALID,
               "log2_table_size must be in [8, 24]");
    N2_REQUIRE(cfg->base_resolution >= 1 && cfg->growth >= 1.0f, N2WOS_ERR_INVALID, "bad resolution parameters");
    for (int a = 0; a < 3; ++a) N2_REQUIRE(cfg->extent[a] > 0.0f, N2WOS_ERR_INVALID, "extent must be positive");
    const int hidden = cfg->n_neurons > 0 ? cfg->n_neurons : kCacheMaxHidden;
    const int layers = cfg->n_hidden_layers > 0 ? cfg->n_hidden_layers : kCacheMaxLayers;
    N2_REQUIRE(hidden == 16 || hidden == 32 || hidden == 64, N2WOS_ERR_INVALID, "n_neurons must be 16, 32 or 64");
    N2_REQUIRE(layers >= 1 && layers <= kCacheMaxLayers, N2WOS_ERR_INVALID, "n_hidden_layers must be in [1, 4]");
    auto* c = new n2wos_cache();
    c->hidden = hidden;
    c->layers = layers;
    c->n_w = cache_weights(hidden, layers);
    try {
      GridLayout& L = c->layout;
      L.n_levels = cfg->n_levels;
      L.table_size = 1u << cfg->log2_table_size;
      uint64_t off = 0;
      for (int l = 0; l < kCacheLevels; ++l) {
        const double nres = std::floor(static_cast<double>(cfg->base_resolution) * std::pow(cfg->growth, l));
        L.res[l] = static_cast<int>(nres);
        // 8 entries per brick of 2 x 2 x 2 vertices, N/2 + 1 bricks per axis
        const double nb = std::floor(nres / 2.0) + 1.0;
        const double entries = 8.0 * nb * nb * nb;
        L.dense[l] = entries <= static_cast<double>(L.table_size) ? 1 : 0;
        L.size[l] = L.dense[l] ? static_cast<uint32_t>(entries) : L.table_size;
        L.offset[l] = static_cast<uint32_t>(off);
        if (l < L.n_levels) off += L.size[l];
      }
      N2_REQUIRE(off < (uint64_t{1} << 31), N2WOS_ERR_INVALID, "grid too large");
      L.total_entries = static_cast<uint32_t>(off);
      for (int a = 0; a < 3; ++a) {
        L.in_min[a] = cfg->origin[a];
        L.in_scale[a] = 1.0f / cfg->extent[a];
      }
      c->lr = cfg->learning_rate;
      c->beta1 = cfg->beta1;
      c->beta2 = cfg->beta2;
      c->eps = cfg->epsilon;
      c->loss_scale = cfg->loss_scale;
      const int max_batch = cfg->max_batch > 0 ? cfg->max_batch : (1 << 18);
      N2_REQUIRE(max_batch <= (1 << 24), N2WOS_ERR_INVALID, "max_batch must be <= 2^24");
      const size_t gf = grid_floats(c);
      // one gradient copy per XCD for grids up to 2^20 floats (4 MB): measured
      // on a 6-level 2^13 grid, a training step takes 0.24 instead of 0.33 ms;
      // a 16-level 2^19 grid gains nothing and would read 8x the gradients
      c->grad_copies = gf <= (size_t{1} << 20) ? 8 : 1;
      N2_HIP(hipMalloc(&c->grid, gf * 4));
      N2_HIP(hipMalloc(&c->grid_grad, gf * 4 * c->grad_copies));
      N2_HIP(hipMalloc(&c->grid_m, gf * 4));
      N2_HIP(hipMalloc(&c->grid_v, gf * 4));
      const size_t nw = static_cast<size_t>(c->n_w);
      N2_HIP(hipMalloc(&c->w, nw * 4));
      N2_HIP(hipMalloc(&c->w_grad, nw * 4));
      N2_HIP(hipMalloc(&c->w_m, nw * 4));
      N2_HIP(hipMalloc(&c->w_v, nw * 4));
      N2_HIP(hipMalloc(&c->wh, nw * 2));
      N2_HIP(hipMalloc(&c->bad, 64));
      N2_HIP(hipMalloc(&c->loss, 64));
      // the activation store of the largest batch: training allocates nothing
      c->store_batch = (max_batch + 15) / 16 * 16;
      N2_HIP(hipMalloc(&c->store, store_halves(hidden, layers, static_cast<size_t>(c->store_batch)) * 2));
      N2_HIP(hipMemset(c->grid_grad, 0, gf * 4 * c->grad_copies));
      N2_HIP(hipMemset(c->grid_m, 0, gf * 4));
      N2_HIP(hipMemset(c->grid_v, 0, gf * 4));
      N2_HIP(hipMemset(c->w_grad, 0, nw * 4));
      N2_HIP(hipMemset(c->w_m, 0, nw * 4));
      N2_HIP(hipMemset(c->w_v, 0, nw * 4));
      N2_HIP(hipMemset(c->bad, 0, 64));
      // grid: uniform +-1e-4; MLP: Xavier-uniform per layer (Philox, seeded)
      hipLaunchKernelGGL(k_init_uniform, dim3(1024), dim3(256), 0, nullptr, c->grid, gf, -1e-4f, 1e-4f,
                         static_cast<uint32_t>(cfg->seed), 0u);
      // layer k (0 = input, 1 .. L-1 hidden, L = output): fan in, fan out,
      // offset; Philox stream tag k + 1
      for (int l = 0; l <= layers; ++l) {
        const int fin = l == 0 ? kCacheIn : hidden;
        const int fout = l == layers ? 1 : hidden;
        const int off = l == 0 ? 0 : (l == layers ? cache_off_out(hidden, layers) : cache_off_hidden(hidden, l));
        const float a = std::sqrt(6.0f / static_cast<float>(fin + fout));
        const size_t cnt = static_cast<size_t>(fin) * fout;
        hipLaunchKernelGGL(k_init_uniform, dim3(8), dim3(256), 0, nullptr, c->w + off, cnt, -a, a,
                           static_cast<uint32_t>(cfg->seed), static_cast<uint32_t>(l + 1));
      }
      // rows 1..15 of the padded output layer stay zero
      N2_HIP(hipMemset(c->w + cache_off_out(hidden, layers) + hidden, 0, static_cast<size_t>(15 * hidden) * 4));
      hipLaunchKernelGGL(k_to_half, dim3((c->n_w + 255) / 256), dim3(256), 0, nullptr, c->w, c->wh, c->n_w);
      N2_HIP(hipGetLastError());
      N2_HIP(hipDeviceSynchronize());
    } catch (...) {
      free_cache(c);
      throw;
    }
    *out = c;
    return N2WOS_OK;
  });
}

extern "C" int n2wos_cache_destroy(n2wos_cache* c) {
  return guarded([&]() -> int {
    if (c) free_cache(c);
    return N2WOS_OK;
  });
}

extern "C" int n2wos_cache_infer(const n2wos_cache* c, const float* d_points, int n, float* d_out, void* stream) {
  return guarded([&]() -> int {
    N2_REQUIRE(c != nullptr, N2WOS_ERR_NULL, "NULL cache");
    N2_REQUIRE(n >= 0, N2WOS_ERR_INVALID, "n must be non-negative");
    if (n == 0) return N2WOS_OK;
    N2_REQUIRE(d_points && d_out, N2WOS_ERR_NULL, "NULL device pointer");
    const int tiles = (n + 15) / 16;
    const int blocks = std::max(1, std::min(4096, (tiles + 3) / 4));
    with_shape(c, [&](auto HC, auto LC) {
      hipLaunchKernelGGL((k_cache_infer<decltype(HC)::value, decltype(LC)::value>), dim3(blocks), dim3(kCacheBlock),
                         0, static_cast<hipStream_t>(stream), c->layout, c->grid, c->wh, d_points, n, c->out_offset,
                         d_out);
    });
    N2_HIP(hipGetLastError());
    return N2WOS_OK;
  });
}

extern "C" int n2wos_cache_train_step(n2wos_cache* c, const float* d_points, const float* d_targets, int n,
                                      float* d_loss_out, void* stream) {
  return guarded([&]() -> int {
    N2_REQUIRE(c != nullptr, N2WOS_ERR_NULL, "NULL cache");
    N2_REQUIRE(n >= 1, N2WOS_ERR_INVALID, "batch must be >= 1");
    N2_REQUIRE(d_points && d_targets, N2WOS_ERR_NULL, "NULL device pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int batch = (n + 15) / 16 * 16;
    N2_REQUIRE(batch <= c->store_batch, N2WOS_ERR_INVALID,
               "batch of " + std::to_string(n) + " exceeds the cache's max_batch (" +
                   std::to_string(c->store_batch) + ")");
    const TrainStore st = make_store(c, batch);
    // the step is skipped when a gradient term is not finite: pass 1 and the
    // weight-gradient kernel flag it where they compute the terms (a grid
    // gradient is a sum of the scattered terms, far too few and too small to
    // overflow FP32 when each is finite)
    hipLaunchKernelGGL(k_zero_two, dim3(1), dim3(1), 0, s, c->bad, reinterpret_cast<int*>(c->loss));
    launch_pass1(c, d_points, d_targets, n, c->loss_scale / static_cast<float>(n), st, s);
    launch_wgrad(c, st, s);
    c->step += 1;
    const double bc1 = 1.0 - std::pow(static_cast<double>(c->beta1), static_cast<double>(c->step));
    const double bc2 = 1.0 - std::pow(static_cast<double>(c->beta2), static_cast<double>(c->step));
    AdamArgs a;
    a.w = c->w; a.w_grad = c->w_grad; a.w_m = c->w_m; a.w_v = c->w_v; a.wh = c->wh; a.n_w = c->n_w;
    a.grid = c->grid; a.grid_grad = c->grid_grad; a.grid_m = c->grid_m; a.grid_v = c->grid_v;
    a.n_grid = grid_floats(c); a.copies = c->grad_copies;
    a.lr_t = static_cast<float>(c->lr * std::sqrt(bc2) / bc1);
    a.beta1 = c->beta1; a.beta2 = c->beta2;
    a.eps = static_cast<float>(c->eps * std::sqrt(bc2));
    a.inv_scale = 1.0f / c->loss_scale;
    a.bad = c->bad;
    a.wblocks = (c->n_w + 255) / 256;
    hipLaunchKernelGGL(k_adam_step, dim3(a.wblocks + 2048), dim3(256), 0, s, a);
    N2_HIP(hipGetLastError());
    if (d_loss_out) N2_HIP(hipMemcpyAsync(d_loss_out, c->loss, 4, hipMemcpyDeviceToDevice, s));
    return N2WOS_OK;
  });
}

extern "C" int n2wos_cache_set_output_offset(n2wos_cache* c, float offset) {
  return guarded([&]() -> int {
    N2_REQUIRE(c != nullptr, N2WOS_ERR_NULL, "NULL cache");
    N2_REQUIRE(std::isfinite(offset), N2WOS_ERR_INVALID, "offset must be finite");
    c->out_offset = offset;
    return N2WOS_OK;
  });
}

// Device-to-device copy of one cache's parameters (grid, FP32 and FP16
// weights, output offset) into another of the same layout, queued on a
// stream: a snapshot for inference while the source keeps training.
extern "C" int n2wos_cache_copy_params(n2wos_cache* dst, const n2wos_cache* src, void* stream) {
  return guarded([&]() -> int {
    N2_REQUIRE(dst && src, N2WOS_ERR_NULL, "NULL cache");
    const GridLayout& a = dst->layout;
    const GridLayout& b = src->layout;
    bool same = a.n_levels == b.n_levels && a.table_size == b.table_size && a.total_entries == b.total_entries;
    for (int l = 0; same && l < kCacheLevels; ++l) same = a.offset[l] == b.offset[l] && a.res[l] == b.res[l];
    same = same && dst->hidden == src->hidden && dst->layers == src->layers;
    N2_REQUIRE(same, N2WOS_ERR_INVALID, "caches have different layouts");
    if (dst == src) return N2WOS_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    N2_HIP(hipMemcpyAsync(dst->grid, src->grid, grid_floats(src) * 4, hipMemcpyDeviceToDevice, s));
    N2_HIP(hipMemcpyAsync(dst->w, src->w, static_cast<size_t>(src->n_w) * 4, hipMemcpyDeviceToDevice, s));
    N2_HIP(hipMemcpyAsync(dst->wh, src->wh, static_cast<size_t>(src->n_w) * 2, hipMemcpyDeviceToDevice, s));
    dst->out_offset = src->out_offset;
    return N2WOS_OK;
  });
}

extern "C" int n2wos_cache_layout(const n2wos_cache* c, int* n_levels, uint32_t* table_size, uint32_t* offsets,
                                  uint32_t* sizes, int* res, int* dense, uint32_t* total_entries) {
  return guarded([&]() -> int {
    N2_REQUIRE(c != nullptr, N2WOS_ERR_NULL, "NULL cache");
    const GridLayout& L = c->layout;
    if (n_levels) *n_levels = L.n_levels;
    if (table_size) *table_size = L.table_size;
    for (int l = 0; l < kCacheLevels; ++l) {
      if (offsets) offsets[l] = L.offset[l];
      if (sizes) sizes[l] = L.size[l];
      if (res) res[l] = L.res[l];
      if (dense) dense[l] = L.dense[l];
    }
    if (total_entries) *total_entries = L.total_entries;
    return N2WOS_OK;
  });
}

// Host copies of (or host uploads to) the FP32 parameters: grid [2 * entries],
// MLP weights [32H + (L-1)H^2 + 16H] (layout in the header); either pointer
// may be NULL.
extern "C" int n2wos_cache_get_params(const n2wos_cache* c, float* grid, float* weights) {
  return guarded([&]() -> int {
    N2_REQUIRE(c != nullptr, N2WOS_ERR_NULL, "NULL cache");
    N2_HIP(hipDeviceSynchronize());
    if (grid) N2_HIP(hipMemcpy(grid, c->grid, grid_floats(c) * 4, hipMemcpyDeviceToHost));
    if (weights) N2_HIP(hipMemcpy(weights, c->w, static_cast<size_t>(c->n_w) * 4, hipMemcpyDeviceToHost));
    return N2WOS_OK;
  });
}

extern "C" int n2wos_cache_set_params(n2wos_cache* c, const float* grid, const float* weights) {
  return guarded([&]() -> int {
    N2_REQUIRE(c != nullptr, N2WOS_ERR_NULL, "NULL cache");
    N2_HIP(hipDeviceSynchronize());
    if (grid) N2_HIP(hipMemcpy(c->grid, grid, grid_floats(c) * 4, hipMemcpyHostToDevice));
    if (weights) {
      N2_HIP(hipMemcpy(c->w, weights, static_cast<size_t>(c->n_w) * 4, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_to_half, dim3((c->n_w + 255) / 256), dim3(256), 0, nullptr, c->w, c->wh, c->n_w);
      N2_HIP(hipGetLastError());
    }
    N2_HIP(hipDeviceSynchronize());
    return N2WOS_OK;
  });
}

// Gradients of the last training step are consumed by Adam; this entry point
// computes them WITHOUT updating (for gradient checks): pass 1 and pass 2 into
// the gradient buffers, copied to the host (unscaled), then zeroed.
extern "C" int n2wos_cache_gradients(n2wos_cache* c, const float* d_points, const float* d_targets, int n,
                                     float* grid_grad, float* weight_grad, float* loss) {
  return guarded([&]() -> int {
    N2_REQUIRE(c != nullptr, N2WOS_ERR_NULL, "NULL cache");
    N2_REQUIRE(n >= 1 && d_points && d_targets, N2WOS_ERR_INVALID, "bad batch");
    const int batch = (n + 15) / 16 * 16;
    N2_REQUIRE(batch <= c->store_batch, N2WOS_ERR_INVALID, "batch exceeds the cache's max_batch");
    const TrainStore st = make_store(c, batch);
    N2_HIP(hipMemset(c->loss, 0, 4));
    launch_pass1(c, d_points, d_targets, n, c->loss_scale / static_cast<float>(n), st, nullptr);
    launch_wgrad(c, st, nullptr);
    N2_HIP(hipGetLastError());
    N2_HIP(hipDeviceSynchronize());
    const size_t gf = grid_floats(c);
    std::vector<float> tmp;
    if (grid_grad) {
      // the sum of the per-XCD copies, in copy order (as k_adam_step sums them)
      tmp.resize(gf * static_cast<size_t>(c->grad_copies));
      N2_HIP(hipMemcpy(tmp.data(), c->grid_grad, tmp.size() * 4, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < gf; ++k) {
        float s = tmp[k];
        for (int x = 1; x < c->grad_copies; ++x) s += tmp[x * gf + k];
        grid_grad[k] = s;
      }
    }
    if (weight_grad) N2_HIP(hipMemcpy(weight_grad, c->w_grad, static_cast<size_t>(c->n_w) * 4, hipMemcpyDeviceToHost));
    const float inv = 1.0f / c->loss_scale;
    if (grid_grad)
      for (size_t k = 0; k < gf; ++k) grid_grad[k] *= inv;
    if (weight_grad)
      for (int k = 0; k < c->n_w; ++k) weight_grad[k] *= inv;
    if (loss) {
      N2_HIP(hipMemcpy(loss, c->loss, 4, hipMemcpyDeviceToHost));
      *loss /= static_cast<float>(n);
    }
    N2_HIP(hipMemset(c->grid_grad, 0, gf * 4 * static_cast<size_t>(c->grad_copies)));
    N2_HIP(hipMemset(c->w_grad, 0, static_cast<size_t>(c->n_w) * 4));
    return N2WOS_OK;
  });
}
