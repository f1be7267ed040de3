// Device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fs_internal.hpp"

namespace fs {

// --------------------------------------------------------------------------
// reductions
// --------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;
}

// Sum NV values over the block; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double s[NV][kBlock / kWave];
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = wave_sum(v[k]);
    if (lane == 0) s[k][wv] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kBlock / kWave; ++w) t += s[k][w];
      v[k] = t;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void store_partial(double* p, double v) { *p = v; }

// out[r] = sum over b of partials[r*nb + b] for r < nr (NR >= nr), each in a
// fixed order, all quantities in one pass.  Results valid in thread 0.
template <int NR>
__device__ __forceinline__ void final_sums(const double* __restrict__ partials, int nb, int nr,
                                           double (&out)[NR]) {
#pragma unroll
  for (int r = 0; r < NR; ++r) out[r] = 0.0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r >= nr) continue;
    const double* pr = partials + (int64_t)r * nb;
    // four independent chains keep four loads in flight per thread
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int b = threadIdx.x;
    for (; b + 3 * kBlock < nb; b += 4 * kBlock) {
      a0 += pr[b];
      a1 += pr[b + kBlock];
      a2 += pr[b + 2 * kBlock];
      a3 += pr[b + 3 * kBlock];
    }
    for (; b < nb; b += kBlock) a0 += pr[b];
    out[r] = (a0 + a1) + (a2 + a3);
  }
  block_sum<NR>(out);
}

// Sum of one quantity's partials (result in thread 0).
__device__ __forceinline__ double sum_partials(const double* __restrict__ pr, int nb) {
  double v[1];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int b = threadIdx.x;
  for (; b + 3 * kBlock < nb; b += 4 * kBlock) {
    a0 += pr[b];
    a1 += pr[b + kBlock];
    a2 += pr[b + 2 * kBlock];
    a3 += pr[b + 3 * kBlock];
  }
  for (; b < nb; b += kBlock) a0 += pr[b];
  v[0] = (a0 + a1) + (a2 + a3);
  block_sum<1>(v);
  return v[0];
}

// ----------------------------------------------------------------------------
// raw operator row: (A x)_n = sum_e w_e (x_n - x_m) + extra_n x_n   (FP64)
// ----------------------------------------------------------------------------
// The six faces of node n = (i, j, k), in the order +x, -x, +y, -y, +z, -z:
// m[f] the neighbour across face f (-1: none), e[f] the edge slot holding its
// weight (the slot of the face's lower node).  Along an axis in d.per the
// last node's + face and the first node's - face are the wrap-around edge,
// whose weight sits in the last node's slot.
struct Faces {
  int64_t m[6];
  int64_t e[6];
};

__device__ __forceinline__ Faces faces_of(const Dims& d, int i, int j, int k, int64_t n) {
  Faces f;
  const int pos[3] = {i, j, k};
  const int len[3] = {d.nx, d.ny, d.nz};
  const int64_t st[3] = {1, (int64_t)d.nx, d.sxy};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const bool wrap = ((d.per >> a) & 1) && len[a] > 1;
    const int64_t span = (int64_t)(len[a] - 1) * st[a];
    f.e[2 * a] = n;
    f.m[2 * a] = pos[a] + 1 < len[a] ? n + st[a] : (wrap ? n - span : -1);
    if (pos[a] > 0) {
      f.m[2 * a + 1] = n - st[a];
      f.e[2 * a + 1] = n - st[a];
    } else {
      f.m[2 * a + 1] = wrap ? n + span : -1;
      f.e[2 * a + 1] = wrap ? n + span : n;
    }
  }
  return f;
}

__device__ __forceinline__ double raw_row(const double* __restrict__ x,
                                          const double* __restrict__ wx,
                                          const double* __restrict__ wy,
                                          const double* __restrict__ wz,
                                          const double* __restrict__ extra, const Dims& d,
                                          int i, int j, int k, int64_t n) {
  const double xc = x[n];
  const double* w[3] = {wx, wy, wz};
  const Faces f = faces_of(d, i, j, k, n);
  double y = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q)
    if (f.m[q] >= 0) y += w[q >> 1][f.e[q]] * (xc - x[f.m[q]]);
  if (extra) y += extra[n] * xc;
  return y;
}


// (A (x - y))_n without forming x - y
__device__ __forceinline__ double raw_row_diff(const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ wx,
                                               const double* __restrict__ wy,
                                               const double* __restrict__ wz,
                                               const double* __restrict__ extra, const Dims& d,
                                               int i, int j, int k, int64_t n) {
  const double c = x[n] - y[n];
  double v = 0.0;
  if (i + 1 < d.nx) v += wx[n] * (c - (x[n + 1] - y[n + 1]));
  if (i > 0) v += wx[n - 1] * (c - (x[n - 1] - y[n - 1]));
  if (j + 1 < d.ny) v += wy[n] * (c - (x[n + d.nx] - y[n + d.nx]));
  if (j > 0) v += wy[n - d.nx] * (c - (x[n - d.nx] - y[n - d.nx]));
  if (k + 1 < d.nz) v += wz[n] * (c - (x[n + d.sxy] - y[n + d.sxy]));
  if (k > 0) v += wz[n - d.sxy] * (c - (x[n - d.sxy] - y[n - d.sxy]));
  if (extra) v += extra[n] * c;
  return v;
}

// --------------------------------------------------------------------------
// Krylov vectors: NR = compile-time right-hand-side capacity, nr <= NR live;
// the vectors of right-hand side r start at r*vs.
// --------------------------------------------------------------------------
struct VecArgs {
  int64_t n;
  int64_t vs;
  int nr;
};

__device__ __forceinline__ bool rhs_live(const CgScalars* sc, int r, int mode) {
  return mode == 0 ? sc[r].done == 0 : (sc[r].done == 0 && sc[r].cycle_done == 0);
}

__device__ __forceinline__ bool all_idle(const CgScalars* sc, int nr, int mode) {
  for (int r = 0; r < nr; ++r)
    if (rhs_live(sc, r, mode)) return false;
  return true;
}

// (i, j, k) of node n = (k*ny + j)*nx + i: 32-bit divisions when the node
// count fits (64-bit integer division is a long emulated sequence)
__device__ __forceinline__ void node_coords(const Dims& d, int64_t n, int& i, int& j, int& k) {
  if (d.n <= 0xffffffffLL) {
    const uint32_t u = (uint32_t)n, nx = (uint32_t)d.nx, ny = (uint32_t)d.ny;
    const uint32_t row = u / nx;
    i = (int)(u - row * nx);
    k = (int)(row / ny);
    j = (int)(row - (uint32_t)k * ny);
  } else {
    i = (int)(n % d.nx);
    j = (int)((n / d.nx) % d.ny);
    k = (int)(n / d.sxy);
  }
}

// Bind a tile of the XCD-remapped launch; returns false for padding blocks.
__device__ __forceinline__ bool tile_of_block(const Tiling& tl, int& tx, int& ty, int& tz) {
  const int bid = blockIdx.x;
  const int per = tl.nblocks >> 3;
  const int t = (bid & 7) * per + (bid >> 3);
  if (t >= tl.ntiles) return false;
  tx = t % tl.tiles_x;
  ty = (t / tl.tiles_x) % tl.tiles_y;
  tz = t / (tl.tiles_x * tl.tiles_y);
  return true;
}

}  // namespace fs
