// Aggregation hierarchy of the learned preconditioner (FP32).
//
// Level 0 is the prepared system's equilibrated operator A = I - C (unit
// diagonal; couplings that touch Dirichlet or pinned nodes are zero).  Level
// l+1 aggregates boxes of level-l nodes: the tensor product of one grouping
// of consecutive indices per axis.  With the piecewise-constant prolongation
// P of such boxes the Galerkin product P^T A P is again a 7-point operator:
//   d_I  = sum_{n in I} d_n - 2 sum_{edges inside I} c_e
//   c_IJ = sum of the couplings of the level-l edges between boxes I and J.
// One cycle (applied to a residual r on level l):
//   z1 = omega D^-1 r
//   r_{l+1} = P^T (r - A z1)                      (restrict_)
//   e_{l+1} = cycle on level l+1 (the network, or one Jacobi step, at the bottom)
//   z2 = z1 + P e_{l+1};  e_l = z2 + omega D^-1 (r - A z2)   (prolong)

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fs_device.hpp"
#include "fs_internal.hpp"
#include "fs_launch.hpp"
#include "fs_nn.hpp"

namespace fs {
namespace mg {

__device__ __forceinline__ void node_ijk(const Level& L, int64_t n, int& i, int& j, int& k) {
  if (L.n <= 0xffffffffLL) {  // 32-bit divisions (64-bit ones are long emulated sequences)
    const uint32_t u = (uint32_t)n, nx = (uint32_t)L.nx, ny = (uint32_t)L.ny;
    const uint32_t row = u / nx;
    i = (int)(u - row * nx);
    k = (int)(row / ny);
    j = (int)(row - (uint32_t)k * ny);
    return;
  }
  i = (int)(n % L.nx);
  j = (int)((n / L.nx) % L.ny);
  k = (int)(n / ((int64_t)L.nx * L.ny));
}

// diagonal, inverse diagonal and activity of node n (level 0: unit, mask)
template <bool UNIT>
__device__ __forceinline__ float dinv_of(const Level& L, int64_t n) {
  if constexpr (UNIT) return L.mask[n] == 0 ? 1.f : 0.f;
  else return L.dinv[n];
}
template <bool UNIT>
__device__ __forceinline__ float diag_of(const Level& L, int64_t n) {
  if constexpr (UNIT) return 1.f;
  else return L.dg[n];
}

// -------------------------------------------------------------------------
// setup: Galerkin coarse operator of one aggregation step (double sums)
// -------------------------------------------------------------------------
template <bool UNIT>
__global__ __launch_bounds__(kBlock) void k_coarsen(Level f, Map m, float* __restrict__ dg,
                                                    float* __restrict__ dinv,
                                                    float* __restrict__ c0,
                                                    float* __restrict__ c1,
                                                    float* __restrict__ c2) {
  const int64_t nc = (int64_t)m.cn[0] * m.cn[1] * m.cn[2];
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < nc;
       t += (int64_t)gridDim.x * kBlock) {
    const int I = (int)(t % m.cn[0]), J = (int)((t / m.cn[0]) % m.cn[1]),
              K = (int)(t / ((int64_t)m.cn[0] * m.cn[1]));
    const int i0 = m.start[0][I], i1 = m.start[0][I + 1];
    const int j0 = m.start[1][J], j1 = m.start[1][J + 1];
    const int k0 = m.start[2][K], k1 = m.start[2][K + 1];
    double d = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    int act = 0;
    for (int k = k0; k < k1; ++k)
      for (int j = j0; j < j1; ++j)
        for (int i = i0; i < i1; ++i) {
          const int64_t n = ((int64_t)k * f.ny + j) * f.nx + i;
          if (dinv_of<UNIT>(f, n) > 0.f) {
            d += (double)diag_of<UNIT>(f, n);
            ++act;
          }
          // edges to the + neighbour: internal ones count twice against the
          // diagonal, the ones leaving the box are the coarse couplings
          if (i + 1 < i1) d -= 2.0 * (double)f.c[0][n];
          else if (I + 1 < m.cn[0]) cx += (double)f.c[0][n];
          if (j + 1 < j1) d -= 2.0 * (double)f.c[1][n];
          else if (J + 1 < m.cn[1]) cy += (double)f.c[1][n];
          if (k + 1 < k1) d -= 2.0 * (double)f.c[2][n];
          else if (K + 1 < m.cn[2]) cz += (double)f.c[2][n];
        }
    const bool active = act > 0 && d > 0.0;
    dg[t] = active ? (float)d : 1.f;
    dinv[t] = active ? (float)(1.0 / d) : 0.f;
    c0[t] = active ? (float)cx : 0.f;
    c1[t] = active ? (float)cy : 0.f;
    c2[t] = active ? (float)cz : 0.f;
  }
}

// network inputs that depend on the operator only: couplings normalised by
// sqrt(d_I d_J) along +x, +y, +z, and activity; dh = 1/sqrt(d) on active nodes
__global__ __launch_bounds__(kBlock) void k_bottom_features(Level L, float* __restrict__ feat4,
                                                            float* __restrict__ dh) {
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < L.n;
       n += (int64_t)gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(L, n, i, j, k);
    const float di = L.dinv[n];
    const float hi = di > 0.f ? sqrtf(di) : 0.f;
    const int64_t nb[3] = {i + 1 < L.nx ? n + 1 : -1, j + 1 < L.ny ? n + L.nx : -1,
                           k + 1 < L.nz ? n + (int64_t)L.nx * L.ny : -1};
    for (int a = 0; a < 3; ++a) {
      float v = 0.f;
      if (nb[a] >= 0) {
        const float dj = L.dinv[nb[a]];
        if (hi > 0.f && dj > 0.f) v = L.c[a][n] * hi * sqrtf(dj);
      }
      feat4[n * 4 + a] = v;
    }
    feat4[n * 4 + 3] = di > 0.f ? 1.f : 0.f;
    dh[n] = hi;
  }
}

// -------------------------------------------------------------------------
// cycle kernels
// -------------------------------------------------------------------------
// rc_I = sum over the fine nodes n of aggregate I of t_n,
// t_n = (r - A omega D^-1 r)_n = (1 - omega) r_n + omega sum_m c_nm dinv_m r_m
// (on inactive nodes r = 0 and every coupling is 0, so t = 0 there)
template <bool UNIT>
__global__ __launch_bounds__(kBlock) void k_restrict(Level f, Map m, const float* __restrict__ r,
                                                     float* __restrict__ rc, int64_t vsc, int nr,
                                                     float omega,
    const int* __restrict__ gate) {
  if (gate && *gate) return;  // every right-hand side of the batch is idle
  // one thread per (coarse node, right-hand side), the right-hand side
  // slowest: a small level has few coarse nodes, and each sums a group of
  // fine nodes in sequence
  const int64_t nc = (int64_t)m.cn[0] * m.cn[1] * m.cn[2];
  const Level cl{m.cn[0], m.cn[1], m.cn[2], nc, 0, {}, nullptr, nullptr, nullptr};
  const int64_t sxy = (int64_t)f.nx * f.ny;
  const int64_t nt = nc * nr;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int b = nt <= 0xffffffffLL ? (int)((uint32_t)t / (uint32_t)nc) : (int)(t / nc);
    const int64_t c = t - (int64_t)b * nc;
    int I, J, K;
    node_ijk(cl, c, I, J, K);
    const int i0 = m.start[0][I], i1 = m.start[0][I + 1];
    const int j0 = m.start[1][J], j1 = m.start[1][J + 1];
    const int k0 = m.start[2][K], k1 = m.start[2][K + 1];
    const float* __restrict__ rb = r + b * f.vs;
    float s = 0.f;
    for (int k = k0; k < k1; ++k)
      for (int j = j0; j < j1; ++j)
        for (int i = i0; i < i1; ++i) {
          // t_n = (1 - omega) r_n + omega sum_m c_nm dinv_m r_m
          const int64_t n = ((int64_t)k * f.ny + j) * f.nx + i;
          float a = 0.f;
          if (i + 1 < f.nx) a += f.c[0][n] * dinv_of<UNIT>(f, n + 1) * rb[n + 1];
          if (i > 0) a += f.c[0][n - 1] * dinv_of<UNIT>(f, n - 1) * rb[n - 1];
          if (j + 1 < f.ny) a += f.c[1][n] * dinv_of<UNIT>(f, n + f.nx) * rb[n + f.nx];
          if (j > 0) a += f.c[1][n - f.nx] * dinv_of<UNIT>(f, n - f.nx) * rb[n - f.nx];
          if (k + 1 < f.nz) a += f.c[2][n] * dinv_of<UNIT>(f, n + sxy) * rb[n + sxy];
          if (k > 0) a += f.c[2][n - sxy] * dinv_of<UNIT>(f, n - sxy) * rb[n - sxy];
          s += (1.f - omega) * rb[n] + omega * a;
        }
    rc[b * vsc + c] = s;
  }
}

// e_n = z2_n + omega dinv_n (r - A z2)_n with z2_n = omega dinv_n r_n +
// (active ? ec[P(n)] : 0); optional FCG partials <e, r>, <e, q>
// SPLIT (a small level without partials): one thread per (node, right-hand
// side), the right-hand side slowest, for enough threads to hide the
// gathers' latency; otherwise one thread per node for every right-hand side.
template <bool UNIT, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_prolong(Level f, Map m, const float* __restrict__ r,
                                                    const float* __restrict__ ec, int64_t vsc,
                                                    float* __restrict__ e, int nr, float omega,
                                                    const float* __restrict__ q,
                                                    double* __restrict__ partials,
    const int* __restrict__ gate) {
  if (gate && *gate) return;  // every right-hand side of the batch is idle
  double acc[2 * kMaxRhs];
#pragma unroll
  for (int a = 0; a < 2 * kMaxRhs; ++a) acc[a] = 0.0;
  const int64_t sxy = (int64_t)f.nx * f.ny;
  const int64_t csxy = (int64_t)m.cn[0] * m.cn[1];
  const int64_t nt = SPLIT ? f.n * nr : f.n;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < nt;
       t += (int64_t)gridDim.x * kBlock) {
    int b0 = 0;
    int64_t n = t;
    if constexpr (SPLIT) {
      b0 = nt <= 0xffffffffLL ? (int)((uint32_t)t / (uint32_t)f.n) : (int)(t / f.n);
      n = t - (int64_t)b0 * f.n;
    }
    int i, j, k;
    node_ijk(f, n, i, j, k);
    const float di = dinv_of<UNIT>(f, n);
    const float dd = diag_of<UNIT>(f, n);
    // the node's 6 neighbours (+x, -x, +y, -y, +z, -z): index, coarse index,
    // inverse diagonal and coupling, read once for every right-hand side
    const bool has[6] = {i + 1 < f.nx, i > 0, j + 1 < f.ny, j > 0, k + 1 < f.nz, k > 0};
    const int64_t nb[6] = {n + 1, n - 1, n + f.nx, n - f.nx, n + sxy, n - sxy};
    const int ci = m.f2c[0][i], cj = m.f2c[1][j], ck = m.f2c[2][k];
    const int64_t cc = ((int64_t)ck * m.cn[1] + cj) * m.cn[0] + ci;
    int64_t cb[6];
    float dn[6], cw[6];
    cb[0] = cc + ((has[0] ? m.f2c[0][i + 1] : ci) - ci);
    cb[1] = cc + ((has[1] ? m.f2c[0][i - 1] : ci) - ci);
    cb[2] = cc + (int64_t)((has[2] ? m.f2c[1][j + 1] : cj) - cj) * m.cn[0];
    cb[3] = cc + (int64_t)((has[3] ? m.f2c[1][j - 1] : cj) - cj) * m.cn[0];
    cb[4] = cc + (int64_t)((has[4] ? m.f2c[2][k + 1] : ck) - ck) * csxy;
    cb[5] = cc + (int64_t)((has[5] ? m.f2c[2][k - 1] : ck) - ck) * csxy;
    cw[0] = has[0] ? f.c[0][n] : 0.f;
    cw[1] = has[1] ? f.c[0][n - 1] : 0.f;
    cw[2] = has[2] ? f.c[1][n] : 0.f;
    cw[3] = has[3] ? f.c[1][n - f.nx] : 0.f;
    cw[4] = has[4] ? f.c[2][n] : 0.f;
    cw[5] = has[5] ? f.c[2][n - sxy] : 0.f;
#pragma unroll
    for (int q = 0; q < 6; ++q) dn[q] = has[q] ? dinv_of<UNIT>(f, nb[q]) : 0.f;
#pragma unroll
    for (int bb = 0; bb < (SPLIT ? 1 : kMaxRhs); ++bb) {
      const int b = SPLIT ? b0 : bb;
      if (!SPLIT && b >= nr) break;
      const float* rb = r + b * f.vs;
      const float* eb = ec + b * vsc;
      float out = 0.f;
      if (di > 0.f) {
        // z2_m = omega dinv_m r_m + ec[P(m)] on active nodes, 0 elsewhere
        const float zc = omega * di * rb[n] + eb[cc];
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < 6; ++q)
          if (has[q]) s += cw[q] * (dn[q] > 0.f ? omega * dn[q] * rb[nb[q]] + eb[cb[q]] : 0.f);
        out = zc + omega * di * (rb[n] - dd * zc + s);
      }
      e[b * f.vs + n] = out;
      if (!SPLIT && partials) {
        acc[bb] += (double)out * (double)rb[n];
        acc[kMaxRhs + bb] += (double)out * (double)q[b * f.vs + n];
      }
    }
  }
  if (SPLIT || !partials) return;  // uniform over the launch
  block_sum<2 * kMaxRhs>(acc);
  if (threadIdx.x == 0) {
    for (int b = 0; b < nr; ++b) {
      store_partial(partials + (int64_t)b * gridDim.x + blockIdx.x, acc[b]);
      store_partial(partials + (int64_t)(nr + b) * gridDim.x + blockIdx.x, acc[kMaxRhs + b]);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_bottom_jacobi(Level L, const float* __restrict__ r,
                                                          float* __restrict__ e, int nr,
                                                          float omega,
    const int* __restrict__ gate) {
  if (gate && *gate) return;  // every right-hand side of the batch is idle
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < L.n;
       n += (int64_t)gridDim.x * kBlock)
    for (int b = 0; b < nr; ++b) e[b * L.vs + n] = omega * L.dinv[n] * r[b * L.vs + n];
}

// level-0 training loss: w = A z, g = 2 (w - r) gscale_b, partials z.w, z.r
__global__ __launch_bounds__(kBlock) void k_loss_grad(Level L, const float* __restrict__ z,
                                                      const float* __restrict__ r,
                                                      const float* __restrict__ gscale,
                                                      float* __restrict__ g, int nr,
                                                      double* __restrict__ partials) {
  double acc[2 * kMaxRhs];
#pragma unroll
  for (int a = 0; a < 2 * kMaxRhs; ++a) acc[a] = 0.0;
  const int64_t sxy = (int64_t)L.nx * L.ny;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < L.n;
       n += (int64_t)gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(L, n, i, j, k);
#pragma unroll
    for (int b = 0; b < kMaxRhs; ++b) {
      if (b >= nr) break;
      const float* zb = z + b * L.vs;
      float s = 0.f;
      if (i + 1 < L.nx) s += L.c[0][n] * zb[n + 1];
      if (i > 0) s += L.c[0][n - 1] * zb[n - 1];
      if (j + 1 < L.ny) s += L.c[1][n] * zb[n + L.nx];
      if (j > 0) s += L.c[1][n - L.nx] * zb[n - L.nx];
      if (k + 1 < L.nz) s += L.c[2][n] * zb[n + sxy];
      if (k > 0) s += L.c[2][n - sxy] * zb[n - sxy];
      const float w = zb[n] - s;
      const float rv = r[b * L.vs + n];
      g[b * L.vs + n] = 2.f * (w - rv) * gscale[b];
      acc[b] += (double)zb[n] * (double)w;
      acc[kMaxRhs + b] += (double)zb[n] * (double)rv;
    }
  }
  block_sum<2 * kMaxRhs>(acc);
  if (threadIdx.x == 0) {
    for (int b = 0; b < nr; ++b) {
      store_partial(partials + (int64_t)b * gridDim.x + blockIdx.x, acc[b]);
      store_partial(partials + (int64_t)(nr + b) * gridDim.x + blockIdx.x, acc[kMaxRhs + b]);
    }
  }
}

// one training pair from an exact FP64 error e (equilibrated, level 0):
// r = A e with the FP64 couplings; e32, r32 stored; partial of e.r (the
// energy ||e||_A^2 that normalises the loss)
__global__ __launch_bounds__(kBlock) void k_pair_make(Level L, const double* __restrict__ cx,
                                                      const double* __restrict__ cy,
                                                      const double* __restrict__ cz,
                                                      const double* __restrict__ e,
                                                      float* __restrict__ e32,
                                                      float* __restrict__ r32,
                                                      double* __restrict__ partials) {
  double acc[1] = {0.0};
  const int64_t sxy = (int64_t)L.nx * L.ny;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < L.n;
       n += (int64_t)gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(L, n, i, j, k);
    double rv = 0.0, ev = 0.0;
    if (L.mask[n] == 0) {
      double s = 0.0;
      if (i + 1 < L.nx) s += cx[n] * e[n + 1];
      if (i > 0) s += cx[n - 1] * e[n - 1];
      if (j + 1 < L.ny) s += cy[n] * e[n + L.nx];
      if (j > 0) s += cy[n - L.nx] * e[n - L.nx];
      if (k + 1 < L.nz) s += cz[n] * e[n + sxy];
      if (k > 0) s += cz[n - sxy] * e[n - sxy];
      ev = e[n];
      rv = ev - s;
    }
    e32[n] = (float)ev;
    r32[n] = (float)rv;
    acc[0] += ev * rv;
  }
  block_sum<1>(acc);
  if (threadIdx.x == 0) store_partial(partials + blockIdx.x, acc[0]);
}

// bottom training loss: per vector b, c_b = r^T D^-1 r (normalisation),
// partials [nr][nblocks]
__global__ __launch_bounds__(kBlock) void k_jacobi_norm(Level L, const float* __restrict__ r,
                                                        int nr, double* __restrict__ partials) {
  double acc[kMaxRhs];
#pragma unroll
  for (int a = 0; a < kMaxRhs; ++a) acc[a] = 0.0;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < L.n;
       n += (int64_t)gridDim.x * kBlock) {
    const double di = L.dinv[n];
#pragma unroll
    for (int b = 0; b < kMaxRhs; ++b) {
      if (b >= nr) break;
      const double v = r[b * L.vs + n];
      acc[b] += v * v * di;
    }
  }
  block_sum<kMaxRhs>(acc);
  if (threadIdx.x == 0)
    for (int b = 0; b < nr; ++b) store_partial(partials + (int64_t)b * gridDim.x + blockIdx.x, acc[b]);
}

// bottom energy loss L_b = (y^T A y - 2 y^T r) / c_b: gradient
// g = 2 (A y - r) / (c_b nr) and partials [2*nr][nblocks] of y.Ay, y.r
__global__ __launch_bounds__(kBlock) void k_energy_grad(Level L, const float* __restrict__ y,
                                                        const float* __restrict__ r,
                                                        const double* __restrict__ c,
                                                        float* __restrict__ g, int nr,
                                                        double* __restrict__ partials) {
  double acc[2 * kMaxRhs];
#pragma unroll
  for (int a = 0; a < 2 * kMaxRhs; ++a) acc[a] = 0.0;
  const int64_t sxy = (int64_t)L.nx * L.ny;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < L.n;
       n += (int64_t)gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(L, n, i, j, k);
#pragma unroll
    for (int b = 0; b < kMaxRhs; ++b) {
      if (b >= nr) break;
      const float* yb = y + b * L.vs;
      float s = 0.f;
      if (i + 1 < L.nx) s += L.c[0][n] * yb[n + 1];
      if (i > 0) s += L.c[0][n - 1] * yb[n - 1];
      if (j + 1 < L.ny) s += L.c[1][n] * yb[n + L.nx];
      if (j > 0) s += L.c[1][n - L.nx] * yb[n - L.nx];
      if (k + 1 < L.nz) s += L.c[2][n] * yb[n + sxy];
      if (k > 0) s += L.c[2][n - sxy] * yb[n - sxy];
      // inactive nodes: y = r = 0 and every coupling is 0
      const float w = L.dinv[n] > 0.f ? L.dg[n] * yb[n] - s : 0.f;
      const float rv = r[b * L.vs + n];
      const double cb = c[b] > 0.0 ? c[b] : 1.0;
      g[b * L.vs + n] = (float)(2.0 * ((double)w - (double)rv) / (cb * nr));
      acc[b] += (double)yb[n] * (double)w;
      acc[kMaxRhs + b] += (double)yb[n] * (double)rv;
    }
  }
  block_sum<2 * kMaxRhs>(acc);
  if (threadIdx.x == 0) {
    for (int b = 0; b < nr; ++b) {
      store_partial(partials + (int64_t)b * gridDim.x + blockIdx.x, acc[b]);
      store_partial(partials + (int64_t)(nr + b) * gridDim.x + blockIdx.x, acc[kMaxRhs + b]);
    }
  }
}

// -------------------------------------------------------------------------
// level 0 (the prepared system, unit diagonal) on the stencil's tiles: one
// thread per (i, j) column of a 64x4 tile marching over z, couplings read
// once per node, no index division.  A coupling that touches a Dirichlet or
// pinned node is zero, so neighbour terms need no mask test.
// -------------------------------------------------------------------------
__device__ __forceinline__ bool live_of(const CgScalars* sc, int r, int mode) {
  return sc == nullptr || rhs_live(sc, r, mode);
}

// Profiling counters of a level-0 launch carrying right-hand sides b0 ..
// b0+nr-1 of nrt: [0] iterations that found a live right-hand side (counted
// by the launch that carries b0 = 0), [1] live right-hand sides summed.
__device__ __forceinline__ void prof_level0(unsigned long long* prof, const CgScalars* sc,
                                            int mode, int b0, int nr, int nrt) {
  if (!prof || blockIdx.x != 0 || threadIdx.x != 0) return;
  unsigned long long mine = 0, any = 0;
  for (int b = 0; b < nrt; ++b) {
    const bool l = live_of(sc, b, mode);
    any += l ? 1ull : 0ull;
    if (b >= b0 && b < b0 + nr) mine += l ? 1ull : 0ull;
  }
  if (b0 == 0 && any) prof[0] += 1ull;
  prof[1] += mine;
}

// e = z2 + omega (r - z2 + sum_m c_nm z2_m), z2 = omega r + ec[P(n)] on free
// nodes; optional partials <e, r>, <e, q> ([2*nr][gridDim.x]).
template <int NR>
__global__ __launch_bounds__(kBlock) void k_prolong_l0(
    const float* __restrict__ r, const float* __restrict__ ec, int64_t vsc,
    float* __restrict__ e, const float* __restrict__ cx, const float* __restrict__ cy,
    const float* __restrict__ cz, const uint8_t* __restrict__ mask, Dims d, Tiling tl, Map m,
    int64_t vs, int nr, float omega, const float* __restrict__ q,
    double* __restrict__ partials, const CgScalars* sc, int mode,
    unsigned long long* prof, const int* __restrict__ gate) {
  if (gate && *gate) return;  // a batch's device control skipped this cycle
  prof_level0(prof, sc, mode, 0, nr, nr);
  double acc[2 * NR];
#pragma unroll
  for (int a = 0; a < 2 * NR; ++a) acc[a] = 0.0;
  int tx, ty, tz;
  if (tile_of_block(tl, tx, ty, tz)) {
    const int i = tx * kTileX + (int)(threadIdx.x % kTileX);
    const int j = ty * kTileY + (int)(threadIdx.x / kTileX);
    const int k0 = tz * tl.kz;
    const int k1 = min(d.nz, k0 + tl.kz);
    if (i < d.nx && j < d.ny) {
      const int64_t base = (int64_t)j * d.nx + i;
      const bool has_e = i + 1 < d.nx, has_w = i > 0, has_n = j + 1 < d.ny, has_s = j > 0;
      const int cnx = m.cn[0], cny = m.cn[1];
      const int ci = m.f2c[0][i], cj = m.f2c[1][j];
      const int cie = has_e ? m.f2c[0][i + 1] : ci, ciw = has_w ? m.f2c[0][i - 1] : ci;
      const int cjn = has_n ? m.f2c[1][j + 1] : cj, cjs = has_s ? m.f2c[1][j - 1] : cj;
#pragma unroll
      for (int b = 0; b < NR; ++b) {
        if (b >= nr || !live_of(sc, b, mode)) continue;
        const float* __restrict__ rb = r + b * vs;
        const float* __restrict__ eb = ec + b * vsc;
        float* __restrict__ ob = e + b * vs;
        int64_t nn = (int64_t)k0 * d.sxy + base;
        // z2 and r along the column: below, current
        float czm = 0.f, zm = 0.f;
        if (k0 > 0) {
          czm = cz[nn - d.sxy];
          zm = omega * rb[nn - d.sxy] + eb[((int64_t)m.f2c[2][k0 - 1] * cny + cj) * cnx + ci];
        }
        float rc = rb[nn];
        int ck = m.f2c[2][k0];
        float zc = omega * rc + eb[((int64_t)ck * cny + cj) * cnx + ci];
        double a0 = 0.0, a1 = 0.0;
        for (int k = k0; k < k1; ++k, nn += d.sxy) {
          const bool up = k + 1 < d.nz;
          const int cku = up ? m.f2c[2][k + 1] : ck;
          const float ru = up ? rb[nn + d.sxy] : 0.f;
          const float zu = up ? omega * ru + eb[((int64_t)cku * cny + cj) * cnx + ci] : 0.f;
          const int64_t plane = (int64_t)ck * cny;
          const float czc = cz[nn];
          float out = 0.f;
          if (mask[nn] == 0) {
            float s = czc * zu + czm * zm;
            if (has_e) s += cx[nn] * (omega * rb[nn + 1] + eb[(plane + cj) * cnx + cie]);
            if (has_w) s += cx[nn - 1] * (omega * rb[nn - 1] + eb[(plane + cj) * cnx + ciw]);
            if (has_n) s += cy[nn] * (omega * rb[nn + d.nx] + eb[(plane + cjn) * cnx + ci]);
            if (has_s) s += cy[nn - d.nx] * (omega * rb[nn - d.nx] + eb[(plane + cjs) * cnx + ci]);
            out = zc + omega * (rc - zc + s);
          }
          ob[nn] = out;
          if (partials) {
            a0 += (double)out * (double)rc;
            a1 += (double)out * (double)q[b * vs + nn];
          }
          zm = zc;
          zc = zu;
          rc = ru;
          czm = czc;
          ck = cku;
        }
        acc[b] = a0;
        acc[NR + b] = a1;
      }
    }
  } else if (!partials) {
    return;  // padding block (uniform)
  }
  if (!partials) return;  // uniform over the launch
  block_sum<2 * NR>(acc);
  if (threadIdx.x == 0) {
    for (int b = 0; b < nr; ++b) {
      store_partial(partials + (int64_t)b * gridDim.x + blockIdx.x, acc[b]);
      store_partial(partials + (int64_t)(nr + b) * gridDim.x + blockIdx.x, acc[NR + b]);
    }
  }
}

// rc = P^T (r - omega A r) when level 0 pairs nodes along x and y (groups
// {2I, 2I+1}) and the z-tiles [kstart[t], kstart[t+1]) hold whole z-groups:
// each thread sums its column over a z-group, x pairs meet by a lane shuffle,
// y pairs through LDS, in a fixed order.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_restrict_l0(
    const float* __restrict__ r, float* __restrict__ rc, int64_t vsc,
    const float* __restrict__ cx, const float* __restrict__ cy, const float* __restrict__ cz,
    Dims d, Tiling tl, const int* __restrict__ kstart, Map m, int64_t vs, int nr, float omega,
    const CgScalars* sc, int mode, unsigned long long* prof, const int* __restrict__ gate) {
  if (gate && *gate) return;  // a batch's device control skipped this cycle
  prof_level0(prof, sc, mode, 0, nr, nr);
  __shared__ float lds[kTileY][kTileX];
  int tx, ty, tz;
  if (!tile_of_block(tl, tx, ty, tz)) return;  // uniform over the block
  const int lx = (int)(threadIdx.x % kTileX), ly = (int)(threadIdx.x / kTileX);
  const int i = tx * kTileX + lx;
  const int j = ty * kTileY + ly;
  const int k0 = kstart[tz], k1 = kstart[tz + 1];
  const bool valid = i < d.nx && j < d.ny;
  const int64_t base = valid ? (int64_t)j * d.nx + i : 0;
  const bool has_e = i + 1 < d.nx, has_w = i > 0, has_n = j + 1 < d.ny, has_s = j > 0;
  const int cnx = m.cn[0], cny = m.cn[1];
  for (int b = 0; b < nr; ++b) {
    if (!live_of(sc, b, mode)) continue;  // uniform over the block
    const float* __restrict__ rb = r + b * vs;
    int64_t nn = (int64_t)k0 * d.sxy + base;
    float rm = 0.f, czm = 0.f, rcur = 0.f;
    if (valid) {
      if (k0 > 0) {
        rm = rb[nn - d.sxy];
        czm = cz[nn - d.sxy];
      }
      rcur = rb[nn];
    }
    float col = 0.f;
    for (int k = k0; k < k1; ++k, nn += d.sxy) {
      const int ck = m.f2c[2][k];
      if (valid) {
        const bool up = k + 1 < d.nz;
        const float ru = up ? rb[nn + d.sxy] : 0.f;
        const float czc = cz[nn];
        float s = czc * ru + czm * rm;
        if (has_e) s += cx[nn] * rb[nn + 1];
        if (has_w) s += cx[nn - 1] * rb[nn - 1];
        if (has_n) s += cy[nn] * rb[nn + d.nx];
        if (has_s) s += cy[nn - d.nx] * rb[nn - d.nx];
        col += (1.f - omega) * rcur + omega * s;
        rm = rcur;
        rcur = ru;
        czm = czc;
      }
      const bool group_end = k + 1 == k1 || m.f2c[2][k + 1] != ck;  // uniform
      if (group_end) {
        // x pairs: lanes 2I and 2I+1 of the wave
        const float v = col + __shfl_down(col, 1, kWave);
        lds[ly][lx] = v;
        __syncthreads();
        if ((lx & 1) == 0 && (ly & 1) == 0 && valid) {
          const float w = v + (ly + 1 < kTileY ? lds[ly + 1][lx] : 0.f);
          rc[b * vsc + ((int64_t)ck * cny + (j >> 1)) * cnx + (i >> 1)] = w;
        }
        __syncthreads();
        col = 0.f;
      }
    }
  }
}

// -------------------------------------------------------------------------
// level 0 with 4 consecutive x nodes per lane (rows a multiple of 4): one
// 16-byte load per array and plane, every right-hand side advanced in the
// same z-march so couplings and masks are read once per node.
//
// Both kernels issue every global load of a plane before the first use of
// any of them and keep loads out of lane-divergent branches: lanes and rows
// outside the grid read the grid's last node (in bounds) and store nothing,
// and the two x-halo lanes select their neighbour from one load that every
// lane makes.  A load under a divergent branch, merged into a default value,
// makes the compiler wait for it at the merge, which serialised five or more
// memory round trips per plane (256^3, 5 RHS: prolongation 418 -> 280 us,
// restriction 185 -> 144 us).  Node offsets are 32-bit: the set-up selects
// these kernels only when n + nx*ny < 2^31.
// -------------------------------------------------------------------------
__device__ __forceinline__ float4 f4(float a, float b, float c, float d) {
  return make_float4(a, b, c, d);
}
__device__ __forceinline__ float4 ldf4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ float2 ldf2(const float* p) { return *reinterpret_cast<const float2*>(p); }
// component-wise select of two float4 values (the conditional operator on the
// vector struct selects an address and keeps both operands in scratch memory)
__device__ __forceinline__ float4 pick(bool c, float4 a, float4 b) {
  return make_float4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w);
}

// e = z2 + omega D^-1 (r - D z2 + C z2), z2 = omega D^-1 r + ec[P(n)], on
// active nodes; optional partials <e, r>, <e, q>.  DIAG: a coarse level
// (diagonal dgv, inverse diagonal dinvv, activity dinvv > 0) instead of level
// 0 (unit diagonal, Dirichlet mask).  The smoothed iterate of the rows above
// and below comes from the neighbouring waves of the block through LDS; the
// block's two edge rows read the one row outside it.
// HZ (level 0 inside flexible CG): e is stored as FP16 e / |r|_2 per
// right-hand side, |r|_2 = sqrt(sc.aux0), the FP32 cycle's current residual
// norm (saturated at the FP16 range), and the partials use the stored value;
// the direction update reads it back with the same scale.
template <int NR, bool DIAG, bool HZ = false>
__global__ __launch_bounds__(kBlock) void k_prolong_l0x4(
    const float* __restrict__ r, const float* __restrict__ ec, int64_t vsc,
    float* __restrict__ e, const float* __restrict__ cx, const float* __restrict__ cy,
    const float* __restrict__ cz, const uint8_t* __restrict__ mask,
    const float* __restrict__ dgv, const float* __restrict__ dinvv, Dims d, Tiling tl, Map m,
    int64_t vs, int nr, float omega, const float* __restrict__ q,
    double* __restrict__ partials, const CgScalars* sc, int mode, int b0, int nrt,
    unsigned long long* prof, const int* __restrict__ gate) {
  if (gate && *gate) return;  // a batch's device control skipped this cycle
  prof_level0(prof, sc, mode, b0, nr, nrt);
  double acc[2 * NR];
#pragma unroll
  for (int a = 0; a < 2 * NR; ++a) acc[a] = 0.0;
  bool live[NR];
#pragma unroll
  for (int b = 0; b < NR; ++b) live[b] = b < nr && live_of(sc, b0 + b, mode);
  __shared__ float4 zsh[2][NR][kTileY][kWave];
  int tx, ty, tz;
  if (tile_of_block(tl, tx, ty, tz)) {  // uniform over the block
    const int lane = (int)(threadIdx.x % kWave);
    const int ly = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int i = (tx * kWave + lane) * 4;
    const int j = ty * kTileY + ly;  // uniform over the wave
    const bool valid = i < d.nx && j < d.ny;
    // lanes and rows outside the grid read the last node's data and store nothing
    const int ic = min(i, d.nx - 4), jc = min(j, d.ny - 1);
    const bool has_w = ic > 0, has_e = ic + 4 < d.nx, has_n = jc + 1 < d.ny, has_s = jc > 0;
    // the x halo: lane 0 reads its west node, the last lane its east node
    const int xo = (lane == 0 && has_w) ? -1 : ((lane == kWave - 1 && has_e) ? 4 : 0);
    // the one row outside the block that an edge wave reads: south for the
    // first wave, north for the last (the other rows come through LDS)
    const bool own_n = ly + 1 == kTileY, own_s = ly == 0;
    const bool edge = (own_s && has_s) || (own_n && has_n);
    const int sxy = (int)d.sxy, nx = d.nx;
    const int oh = own_s ? -nx : nx;
    const int os = has_s ? -nx : 0;
    const int cnx = m.cn[0], cny = m.cn[1];
    const int ci0 = ic >> 1;
    const int cih = xo < 0 ? (ic - 1) >> 1 : (xo > 0 ? (ic + 4) >> 1 : ci0);
    const int cj = jc >> 1;
    const int cjh = own_s ? (jc - 1) >> 1 : (jc + 1) >> 1;  // used on edge waves only
    auto z2 = [&](float4 rr, float4 dd, float2 ee) {
      if constexpr (DIAG)
        return f4(omega * dd.x * rr.x + ee.x, omega * dd.y * rr.y + ee.x,
                  omega * dd.z * rr.z + ee.y, omega * dd.w * rr.w + ee.y);
      else
        return f4(omega * rr.x + ee.x, omega * rr.y + ee.x, omega * rr.z + ee.y,
                  omega * rr.w + ee.y);
    };
    const float4 z4 = f4(0.f, 0.f, 0.f, 0.f);
    const float4 one4 = f4(1.f, 1.f, 1.f, 1.f);
    const int k0 = tz * tl.kz;
    const int k1 = min(d.nz, k0 + tl.kz);
    int nn = k0 * sxy + jc * nx + ic;
    int ck = __builtin_amdgcn_readfirstlane(m.f2c[2][k0]);
    // state of each live right-hand side: z2 at planes k-1 and k, r at plane k
    float4 zm[NR], zc[NR], rc[NR];
    float4 czm = z4, dcur = one4;
    {
      const int kd = k0 > 0 ? k0 - 1 : k0;
      const int nd = nn - (k0 > 0 ? sxy : 0);
      const int ckd = __builtin_amdgcn_readfirstlane(m.f2c[2][kd]);
      const float4 czd = ldf4(cz + nd);
      float4 dprev = one4;
      if constexpr (DIAG) {
        dcur = ldf4(dinvv + nn);
        dprev = ldf4(dinvv + nd);
      }
      float4 rd[NR];
      float2 ed[NR], e0[NR];
#pragma unroll
      for (int b = 0; b < NR; ++b) {
        if (!live[b]) continue;  // uniform
        const float* rb = r + b * vs;
        const float* eb = ec + b * vsc;
        rd[b] = ldf4(rb + nd);
        rc[b] = ldf4(rb + nn);
        ed[b] = ldf2(eb + (ckd * cny + cj) * cnx + ci0);
        e0[b] = ldf2(eb + (ck * cny + cj) * cnx + ci0);
      }
      if (k0 > 0) czm = czd;
#pragma unroll
      for (int b = 0; b < NR; ++b) {
        if (!live[b]) continue;
        zm[b] = k0 > 0 ? z2(rd[b], dprev, ed[b]) : z4;
        zc[b] = z2(rc[b], dcur, e0[b]);
      }
    }
    for (int k = k0; k < k1; ++k, nn += sxy) {
      const bool up = k + 1 < d.nz;
      const int nu = up ? nn + sxy : nn;
      const int cku = __builtin_amdgcn_readfirstlane(up ? m.f2c[2][k + 1] : ck);
      const int rowc = (ck * cny + cj) * cnx, rowu = (cku * cny + cj) * cnx;
      const int rowh = (ck * cny + cjh) * cnx;
      // ---- every load of this plane ----
      const float4 cxc = ldf4(cx + nn), cyc = ldf4(cy + nn), czc = ldf4(cz + nn);
      const float4 cyl = ldf4(cy + nn + os);
      const float cxh = cx[nn + min(xo, 0)];
      uchar4 mk = make_uchar4(0, 0, 0, 0);
      float4 dgc = one4, dup = one4, dhr = one4;
      float dh1 = 1.f;
      if constexpr (DIAG) {
        dgc = ldf4(dgv + nn);
        dup = ldf4(dinvv + nu);
        dh1 = dinvv[nn + xo];
        if (edge) dhr = ldf4(dinvv + nn + oh);
      } else {
        mk = *reinterpret_cast<const uchar4*>(mask + nn);
      }
      float4 ru[NR], qq[NR], rh4[NR];
      float2 eu[NR], eh2[NR];
      float rh[NR], eh[NR];
#pragma unroll
      for (int b = 0; b < NR; ++b) {
        if (!live[b]) continue;  // uniform over the launch
        const float* __restrict__ rb = r + b * vs;
        const float* __restrict__ eb = ec + b * vsc;
        ru[b] = ldf4(rb + nu);
        eu[b] = ldf2(eb + rowu + ci0);
        rh[b] = rb[nn + xo];
        eh[b] = eb[rowc + cih];
        if (partials) qq[b] = ldf4(q + b * vs + nn);
        if (edge) {
          rh4[b] = ldf4(rb + nn + oh);
          eh2[b] = ldf2(eb + rowh + ci0);
        }
      }
      // ---- z2 of the rows above and below from the neighbouring waves ----
      const int buf = (k - k0) & 1;
#pragma unroll
      for (int b = 0; b < NR; ++b)
        if (live[b]) zsh[buf][b][ly][lane] = zc[b];
      __syncthreads();
      // ---- compute ----
      const float4 cym = has_s ? cyl : z4;
      const float cxw_s = __shfl_up(cxc.w, 1, kWave);
      const float cxw = has_w ? (lane == 0 ? cxh : cxw_s) : 0.f;
#pragma unroll
      for (int b = 0; b < NR; ++b) {
        if (!live[b]) continue;
        const float4 c = zc[b];
        const float4 zu = up ? z2(ru[b], dup, eu[b]) : z4;
        float4 zn = z4, zs = z4;
        if (has_n) zn = own_n ? z2(rh4[b], dhr, eh2[b]) : zsh[buf][b][min(ly + 1, kTileY - 1)][lane];
        if (has_s) zs = own_s ? z2(rh4[b], dhr, eh2[b]) : zsh[buf][b][max(ly - 1, 0)][lane];
        const float zh = omega * (DIAG ? dh1 : 1.f) * rh[b] + eh[b];
        const float sw = __shfl_up(c.w, 1, kWave), se = __shfl_down(c.x, 1, kWave);
        const float zw = has_w ? (lane == 0 ? zh : sw) : 0.f;
        const float ze = has_e ? (lane == kWave - 1 ? zh : se) : 0.f;
        const float4 mm = zm[b], rr = rc[b];
        const float s0 = cxc.x * c.y + cxw * zw + cyc.x * zn.x + cym.x * zs.x + czc.x * zu.x + czm.x * mm.x;
        const float s1 = cxc.y * c.z + cxc.x * c.x + cyc.y * zn.y + cym.y * zs.y + czc.y * zu.y + czm.y * mm.y;
        const float s2 = cxc.z * c.w + cxc.y * c.y + cyc.z * zn.z + cym.z * zs.z + czc.z * zu.z + czm.z * mm.z;
        const float s3 = cxc.w * ze + cxc.z * c.z + cyc.w * zn.w + cym.w * zs.w + czc.w * zu.w + czm.w * mm.w;
        float4 out;
        if constexpr (DIAG) {  // e = z2 + omega D^-1 (r - D z2 + C z2) on active nodes
          out.x = dcur.x > 0.f ? c.x + omega * dcur.x * (rr.x - dgc.x * c.x + s0) : 0.f;
          out.y = dcur.y > 0.f ? c.y + omega * dcur.y * (rr.y - dgc.y * c.y + s1) : 0.f;
          out.z = dcur.z > 0.f ? c.z + omega * dcur.z * (rr.z - dgc.z * c.z + s2) : 0.f;
          out.w = dcur.w > 0.f ? c.w + omega * dcur.w * (rr.w - dgc.w * c.w + s3) : 0.f;
        } else {
          out.x = mk.x == 0 ? c.x + omega * (rr.x - c.x + s0) : 0.f;
          out.y = mk.y == 0 ? c.y + omega * (rr.y - c.y + s1) : 0.f;
          out.z = mk.z == 0 ? c.z + omega * (rr.z - c.z + s2) : 0.f;
          out.w = mk.w == 0 ? c.w + omega * (rr.w - c.w + s3) : 0.f;
        }
        if constexpr (HZ) {  // the value the direction update will read
          const float hs = zscale(sc[b0 + b]);
          const float zi = 1.f / hs;
          const h4 hv = {(_Float16)fminf(fmaxf(out.x * zi, -65504.f), 65504.f),
                         (_Float16)fminf(fmaxf(out.y * zi, -65504.f), 65504.f),
                         (_Float16)fminf(fmaxf(out.z * zi, -65504.f), 65504.f),
                         (_Float16)fminf(fmaxf(out.w * zi, -65504.f), 65504.f)};
          if (valid) *reinterpret_cast<h4*>(reinterpret_cast<_Float16*>(e) + b * vs + nn) = hv;
          out = f4((float)hv[0] * hs, (float)hv[1] * hs, (float)hv[2] * hs, (float)hv[3] * hs);
        }
        if (partials) {
          // the lane's 4 products in FP32, their sums accumulated in FP64
          const float4 qv = qq[b];
          const double a0 = (double)(out.x * rr.x + out.y * rr.y + out.z * rr.z + out.w * rr.w);
          const double a1 = (double)(out.x * qv.x + out.y * qv.y + out.z * qv.z + out.w * qv.w);
          acc[b] += valid ? a0 : 0.0;
          acc[NR + b] += valid ? a1 : 0.0;
        }
        if constexpr (!HZ)
          if (valid) *reinterpret_cast<float4*>(e + b * vs + nn) = out;
        zm[b] = c;
        zc[b] = zu;
        rc[b] = ru[b];
      }
      czm = czc;
      ck = cku;
      dcur = dup;
    }
  }
  if (!partials) return;  // uniform over the launch
  block_sum<2 * NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int b = 0; b < NR; ++b) {
      if (b >= nr) continue;
      partials[(int64_t)(b0 + b) * gridDim.x + blockIdx.x] = acc[b];
      partials[(int64_t)(nrt + b0 + b) * gridDim.x + blockIdx.x] = acc[NR + b];
    }
  }
}

// rc = P^T (r - omega A r) with 4 nodes per lane: x pairs {2I, 2I+1} sum
// inside the lane, y pairs through LDS at the end of each z-group.
// DIAG: a coarse level, t_n = (1 - omega) r_n + omega sum_m c_nm dinv_m r_m
// (level 0: dinv = 1 on free nodes, and r = 0 on the others)
template <int NR, bool DIAG>
__global__ __launch_bounds__(kBlock) void k_restrict_l0x4(
    const float* __restrict__ r, float* __restrict__ rc, int64_t vsc,
    const float* __restrict__ cx, const float* __restrict__ cy, const float* __restrict__ cz,
    const float* __restrict__ dinvv, Dims d, Tiling tl, const int* __restrict__ kstart, Map m,
    int64_t vs, int nr, float omega, const CgScalars* sc, int mode, unsigned long long* prof,
    const int* __restrict__ gate) {
  if (gate && *gate) return;  // a batch's device control skipped this cycle
  prof_level0(prof, sc, mode, 0, nr, nr);
  __shared__ float2 lds[NR][kTileY][kWave];
  int tx, ty, tz;
  if (!tile_of_block(tl, tx, ty, tz)) return;  // uniform over the block
  bool live[NR];
#pragma unroll
  for (int b = 0; b < NR; ++b) live[b] = b < nr && live_of(sc, b, mode);
  const int lane = (int)(threadIdx.x % kWave);
  const int ly = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  const int i = (tx * kWave + lane) * 4;
  const int j = ty * kTileY + ly;
  const bool valid = i < d.nx && j < d.ny;
  const int ic = min(i, d.nx - 4), jc = min(j, d.ny - 1);
  const bool has_w = ic > 0, has_e = ic + 4 < d.nx, has_n = jc + 1 < d.ny, has_s = jc > 0;
  const int xo = (lane == 0 && has_w) ? -1 : ((lane == kWave - 1 && has_e) ? 4 : 0);
  const int sxy = (int)d.sxy, nx = d.nx;
  const int on = has_n ? nx : 0, os = has_s ? -nx : 0;
  const int cnx = m.cn[0], cny = m.cn[1];
  const int k0 = __builtin_amdgcn_readfirstlane(kstart[tz]);
  const int k1 = __builtin_amdgcn_readfirstlane(kstart[tz + 1]);
  const float4 z4 = f4(0.f, 0.f, 0.f, 0.f);
  const float4 one4 = f4(1.f, 1.f, 1.f, 1.f);
  int nn = k0 * sxy + jc * nx + ic;
  float4 rm[NR], rcur[NR];
  float2 col[NR];
  float4 czm = z4, dcur = one4, dprev = one4;
  {
    const int nd = nn - (k0 > 0 ? sxy : 0);
    const float4 czd = ldf4(cz + nd);
    if constexpr (DIAG) {
      dcur = ldf4(dinvv + nn);
      const float4 dd = ldf4(dinvv + nd);
      if (k0 > 0) dprev = dd;
    }
    float4 rd[NR];
#pragma unroll
    for (int b = 0; b < NR; ++b) {
      col[b] = make_float2(0.f, 0.f);
      if (!live[b]) continue;
      rd[b] = ldf4(r + b * vs + nd);
      rcur[b] = ldf4(r + b * vs + nn);
    }
    if (k0 > 0) czm = czd;
#pragma unroll
    for (int b = 0; b < NR; ++b)
      if (live[b]) rm[b] = pick(k0 > 0, rd[b], z4);
  }
  auto mul = [](float4 a, float4 b) { return f4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); };
  for (int k = k0; k < k1; ++k, nn += sxy) {
    const int ck = __builtin_amdgcn_readfirstlane(m.f2c[2][k]);
    const bool up = k + 1 < d.nz;
    const int nu = up ? nn + sxy : nn;
    // ---- every load of this plane ----
    const float4 cxc = ldf4(cx + nn), cyc = ldf4(cy + nn), czc = ldf4(cz + nn);
    const float4 cyl = ldf4(cy + nn + os);
    const float cxh = cx[nn + min(xo, 0)];
    float4 dup = one4, dnr = one4, dsr = one4;
    float dh = 1.f;
    if constexpr (DIAG) {
      dup = ldf4(dinvv + nu);
      dnr = ldf4(dinvv + nn + on);
      dsr = ldf4(dinvv + nn + os);
      dh = dinvv[nn + xo];
    }
    float4 ru[NR], rn[NR], rs[NR];
    float rh[NR];
#pragma unroll
    for (int b = 0; b < NR; ++b) {
      if (live[b]) {  // uniform over the block
        const float* __restrict__ rb = r + b * vs;
        ru[b] = ldf4(rb + nu);
        rn[b] = ldf4(rb + nn + on);
        rs[b] = ldf4(rb + nn + os);
        rh[b] = rb[nn + xo];
      }
    }
    // ---- compute ----
    const float4 cym = pick(has_s, cyl, z4);
    const float cxw_s = __shfl_up(cxc.w, 1, kWave);
    const float cxw = has_w ? (lane == 0 ? cxh : cxw_s) : 0.f;
    if constexpr (DIAG) {
      if (!up) dup = one4;
    }
#pragma unroll
    for (int b = 0; b < NR; ++b) {
      if (!live[b]) continue;
      const float4 cr = rcur[b];
      // neighbour values enter as D^-1 r (level 0: r)
      float4 c = cr;
      if constexpr (DIAG) c = mul(cr, dcur);
      const float hv = (DIAG ? dh : 1.f) * rh[b];
      const float sw = __shfl_up(c.w, 1, kWave), se = __shfl_down(c.x, 1, kWave);
      const float rw = has_w ? (lane == 0 ? hv : sw) : 0.f;
      const float re = has_e ? (lane == kWave - 1 ? hv : se) : 0.f;
      const float4 ruv = pick(up, ru[b], z4);
      float4 ruS = ruv;
      if constexpr (DIAG) ruS = mul(ruv, dup);
      float4 rnS = rn[b], rsS = rs[b], mm = rm[b];
      if constexpr (DIAG) {
        rnS = mul(rnS, dnr);
        rsS = mul(rsS, dsr);
        mm = mul(mm, dprev);
      }
      rnS = pick(has_n, rnS, z4);
      rsS = pick(has_s, rsS, z4);
      const float s0 = cxc.x * c.y + cxw * rw + cyc.x * rnS.x + cym.x * rsS.x + czc.x * ruS.x + czm.x * mm.x;
      const float s1 = cxc.y * c.z + cxc.x * c.x + cyc.y * rnS.y + cym.y * rsS.y + czc.y * ruS.y + czm.y * mm.y;
      const float s2 = cxc.z * c.w + cxc.y * c.y + cyc.z * rnS.z + cym.z * rsS.z + czc.z * ruS.z + czm.z * mm.z;
      const float s3 = cxc.w * re + cxc.z * c.z + cyc.w * rnS.w + cym.w * rsS.w + czc.w * ruS.w + czm.w * mm.w;
      const float w1 = 1.f - omega;
      const float cxs = (w1 * cr.x + omega * s0) + (w1 * cr.y + omega * s1);
      const float cys = (w1 * cr.z + omega * s2) + (w1 * cr.w + omega * s3);
      if (valid) {
        col[b].x += cxs;
        col[b].y += cys;
      }
      rm[b] = cr;
      rcur[b] = ruv;
    }
    czm = czc;
    dprev = dcur;
    dcur = dup;
    const bool group_end =
        k + 1 == k1 || __builtin_amdgcn_readfirstlane(m.f2c[2][k + 1 < d.nz ? k + 1 : k]) != ck;
    if (group_end) {  // uniform over the block
#pragma unroll
      for (int b = 0; b < NR; ++b) lds[b][ly][lane] = col[b];
      __syncthreads();
      if ((ly & 1) == 0 && valid) {
#pragma unroll
        for (int b = 0; b < NR; ++b) {
          if (!live[b]) continue;
          const float2 w = col[b], v = lds[b][ly + 1][lane];
          float* out = rc + b * vsc + ((int64_t)ck * cny + (j >> 1)) * cnx + (i >> 1);
          out[0] = w.x + v.x;
          if ((i >> 1) + 1 < cnx) out[1] = w.y + v.y;
        }
      }
      __syncthreads();
#pragma unroll
      for (int b = 0; b < NR; ++b) col[b] = make_float2(0.f, 0.f);
    }
  }
}

// -------------------------------------------------------------------------
// launchers
// -------------------------------------------------------------------------
void jacobi_norm(hipStream_t s, const Level& L, const float* r, int nr, double* partials,
                 int nblocks) {
  hipLaunchKernelGGL(k_jacobi_norm, dim3(nblocks), dim3(kBlock), 0, s, L, r, nr, partials);
}

void energy_grad(hipStream_t s, const Level& L, const float* y, const float* r, const double* c,
                 float* g, int nr, double* partials, int nblocks) {
  hipLaunchKernelGGL(k_energy_grad, dim3(nblocks), dim3(kBlock), 0, s, L, y, r, c, g, nr,
                     partials);
}

int prolong_l0(hipStream_t s, const L0Args& a, const float* r, const float* ec, int64_t vsc,
               float* e, int nr, float omega, const float* q, double* partials,
               const CgScalars* sc, int mode, hipEvent_t ea, hipEvent_t eb,
               unsigned long long* prof, const int* gate, int zhalf) {
  if (zhalf && (!a.x4 || a.dinv || !sc)) return -1;  // FP16 output: level 0, 4-wide, in a solve
  if (a.x4) {
    // ~16 registers per RHS and lane: launched for the exact RHS count, and
    // in two halves above 5 RHS (8 in one launch would run 1 wave per SIMD);
    // the start event goes with the first launch, the stop event with the last
    hipEvent_t e0 = ea, e1 = nullptr;
#define FS_L0P4(NR_, DG, B0, NB)                                                                 \
  launch_timed(k_prolong_l0x4<NR_, DG>, dim3(a.tl4.nblocks), s, e0, e1, r + (B0) * a.vs,          \
               ec + (B0) * vsc, vsc, e + (B0) * a.vs, a.c[0], a.c[1], a.c[2], a.mask, a.dg,      \
               a.dinv, a.d, a.tl4, a.m, a.vs, NB, omega, q ? q + (B0) * a.vs : nullptr,        \
               partials, sc, mode, B0, nr, prof, gate)
#define FS_L0P4H(NR_, B0, NB)                                                                    \
  launch_timed(k_prolong_l0x4<NR_, false, true>, dim3(a.tl4.nblocks), s, e0, e1,                 \
               r + (B0) * a.vs, ec + (B0) * vsc, vsc,                                          \
               reinterpret_cast<float*>(reinterpret_cast<_Float16*>(e) + (B0) * a.vs), a.c[0],  \
               a.c[1], a.c[2], a.mask, a.dg, a.dinv, a.d, a.tl4, a.m, a.vs, NB, omega,          \
               q ? q + (B0) * a.vs : nullptr, partials, sc, mode, B0, nr, prof, gate)
    auto part = [&](int b0, int nb) {
      if (zhalf) {
        switch (nb) {
          case 1: FS_L0P4H(1, b0, nb); break;
          case 2: FS_L0P4H(2, b0, nb); break;
          case 3: FS_L0P4H(3, b0, nb); break;
          case 4: FS_L0P4H(4, b0, nb); break;
          default: FS_L0P4H(5, b0, nb); break;
        }
      } else if (a.dinv) {
        switch (nb) {
          case 1: FS_L0P4(1, true, b0, nb); break;
          case 2: FS_L0P4(2, true, b0, nb); break;
          case 3: FS_L0P4(3, true, b0, nb); break;
          case 4: FS_L0P4(4, true, b0, nb); break;
          default: FS_L0P4(5, true, b0, nb); break;
        }
      } else {
        switch (nb) {
          case 1: FS_L0P4(1, false, b0, nb); break;
          case 2: FS_L0P4(2, false, b0, nb); break;
          case 3: FS_L0P4(3, false, b0, nb); break;
          case 4: FS_L0P4(4, false, b0, nb); break;
          default: FS_L0P4(5, false, b0, nb); break;
        }
      }
    };
#undef FS_L0P4
#undef FS_L0P4H
    // a coarse level's kernel (with its diagonal) holds more registers: 5
    // right-hand sides there need 256 VGPRs and some AGPRs (1 wave per SIMD),
    // 4 fit in 230 (2 waves)
    if (nr <= (a.dinv ? 4 : 5)) {
      e1 = eb;
      part(0, nr);
    } else {
      const int h = (nr + 1) / 2;
      part(0, h);
      e0 = nullptr;
      e1 = eb;
      part(h, nr - h);
    }
    return a.tl4.nblocks;
  }
#define FS_L0P(NR_)                                                                             \
  launch_timed(k_prolong_l0<NR_>, dim3(a.tl.nblocks), s, ea, eb, r, ec, vsc, e, a.c[0], a.c[1], \
               a.c[2], a.mask, a.d, a.tl, a.m, a.vs, nr, omega, q, partials, sc, mode, prof, gate)
  if (nr <= 1) FS_L0P(1);
  else if (nr <= 2) FS_L0P(2);
  else if (nr <= 4) FS_L0P(4);
  else FS_L0P(8);
#undef FS_L0P
  return a.tl.nblocks;
}

void restrict_l0(hipStream_t s, const L0Args& a, const float* r, float* rc, int64_t vsc, int nr,
                 float omega, const CgScalars* sc, int mode, hipEvent_t ea, hipEvent_t eb,
                 unsigned long long* prof, const int* gate) {
  if (a.x4) {
#define FS_L0R4(NR_, DG)                                                                        \
  launch_timed(k_restrict_l0x4<NR_, DG>, dim3(a.tlr4.nblocks), s, ea, eb, r, rc, vsc, a.c[0],     \
               a.c[1], a.c[2], a.dinv, a.d, a.tlr4, a.kstart, a.m, a.vs, nr, omega, sc, mode, prof, \
               gate)
    // exact RHS count: these kernels hold ~16 registers per RHS and lane
    if (a.dinv) {
      if (nr == 8) {
        // a coarse level's 8 right-hand sides need 256 VGPRs and some AGPRs
        // (1 wave per SIMD): two launches of 4 (2 waves); no profiling here
        launch_timed(k_restrict_l0x4<4, true>, dim3(a.tlr4.nblocks), s, nullptr, nullptr, r, rc,
                     vsc, a.c[0], a.c[1], a.c[2], a.dinv, a.d, a.tlr4, a.kstart, a.m, a.vs, 4,
                     omega, sc, mode, (unsigned long long*)nullptr, gate);
        launch_timed(k_restrict_l0x4<4, true>, dim3(a.tlr4.nblocks), s, nullptr, nullptr,
                     r + 4 * a.vs, rc + 4 * vsc, vsc, a.c[0], a.c[1], a.c[2], a.dinv, a.d, a.tlr4,
                     a.kstart, a.m, a.vs, 4, omega, sc ? sc + 4 : sc, mode,
                     (unsigned long long*)nullptr, gate);
        return;
      }
      switch (nr) {
        case 1: FS_L0R4(1, true); break;
        case 2: FS_L0R4(2, true); break;
        case 3: FS_L0R4(3, true); break;
        case 4: FS_L0R4(4, true); break;
        case 5: FS_L0R4(5, true); break;
        case 6: FS_L0R4(6, true); break;
        default: FS_L0R4(7, true); break;
      }
    } else {
      switch (nr) {
        case 1: FS_L0R4(1, false); break;
        case 2: FS_L0R4(2, false); break;
        case 3: FS_L0R4(3, false); break;
        case 4: FS_L0R4(4, false); break;
        case 5: FS_L0R4(5, false); break;
        case 6: FS_L0R4(6, false); break;
        case 7: FS_L0R4(7, false); break;
        default: FS_L0R4(8, false); break;
      }
    }
#undef FS_L0R4
    return;
  }
  launch_timed(k_restrict_l0<8>, dim3(a.tlr.nblocks), s, ea, eb, r, rc, vsc, a.c[0], a.c[1],
               a.c[2], a.d, a.tlr, a.kstart, a.m, a.vs, nr, omega, sc, mode, prof, gate);
}

int vec_blocks(int64_t n) {
  const int64_t b = (n + kBlock - 1) / kBlock;
  return (int)(b < 2048 ? (b > 0 ? b : 1) : 2048);
}

void coarsen(hipStream_t s, const Level& f, const Map& m, float* dg, float* dinv, float* c0,
             float* c1, float* c2) {
  const int64_t nc = (int64_t)m.cn[0] * m.cn[1] * m.cn[2];
  if (f.dg == nullptr)
    hipLaunchKernelGGL(k_coarsen<true>, dim3(vec_blocks(nc)), dim3(kBlock), 0, s, f, m, dg, dinv,
                       c0, c1, c2);
  else
    hipLaunchKernelGGL(k_coarsen<false>, dim3(vec_blocks(nc)), dim3(kBlock), 0, s, f, m, dg, dinv,
                       c0, c1, c2);
}

void bottom_features(hipStream_t s, const Level& L, float* feat4, float* dh) {
  hipLaunchKernelGGL(k_bottom_features, dim3(vec_blocks(L.n)), dim3(kBlock), 0, s, L, feat4, dh);
}

void restrict_(hipStream_t s, const Level& f, const Map& m, const float* r, float* rc,
               int64_t vsc, int nr, float omega, const int* gate) {
  const int64_t nt = (int64_t)m.cn[0] * m.cn[1] * m.cn[2] * nr;
  const int nb = (int)std::min<int64_t>(std::max<int64_t>((nt + kBlock - 1) / kBlock, 1), 16384);
  if (f.dg == nullptr)
    hipLaunchKernelGGL(k_restrict<true>, dim3(nb), dim3(kBlock), 0, s, f, m, r, rc, vsc, nr,
                       omega, gate);
  else
    hipLaunchKernelGGL(k_restrict<false>, dim3(nb), dim3(kBlock), 0, s, f, m, r, rc, vsc, nr,
                       omega, gate);
}

void prolong(hipStream_t s, const Level& f, const Map& m, const float* r, const float* ec,
             int64_t vsc, float* e, int nr, float omega, const float* q, double* partials,
             int nblocks, const int* gate) {
  // levels below 2^20 nodes without partials: a thread per (node, RHS)
  if (!partials && f.n < (int64_t)1 << 20) {
    const int64_t nt = f.n * nr;
    const int nbs = (int)std::min<int64_t>(std::max<int64_t>((nt + kBlock - 1) / kBlock, 1), 16384);
    if (f.dg == nullptr)
      hipLaunchKernelGGL((k_prolong<true, true>), dim3(nbs), dim3(kBlock), 0, s, f, m, r, ec, vsc,
                         e, nr, omega, q, partials, gate);
    else
      hipLaunchKernelGGL((k_prolong<false, true>), dim3(nbs), dim3(kBlock), 0, s, f, m, r, ec,
                         vsc, e, nr, omega, q, partials, gate);
    return;
  }
  const int nb = nblocks > 0 ? nblocks : vec_blocks(f.n);
  if (f.dg == nullptr)
    hipLaunchKernelGGL((k_prolong<true, false>), dim3(nb), dim3(kBlock), 0, s, f, m, r, ec, vsc, e,
                       nr, omega, q, partials, gate);
  else
    hipLaunchKernelGGL((k_prolong<false, false>), dim3(nb), dim3(kBlock), 0, s, f, m, r, ec, vsc,
                       e, nr, omega, q, partials, gate);
}

void bottom_jacobi(hipStream_t s, const Level& L, const float* r, float* e, int nr,
                   float omega, const int* gate) {
  hipLaunchKernelGGL(k_bottom_jacobi, dim3(vec_blocks(L.n)), dim3(kBlock), 0, s, L, r, e, nr,
                     omega, gate);
}

// out = 1 (skip the cycle) when the external gate is set or every
// right-hand side is idle; count, when given, counts the cycles that run
__global__ void k_gate(const CgScalars* sc, int nr, int mode, const int* ext, int* out,
                       int* count) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int skip = (ext && *ext) || all_idle(sc, nr, mode) ? 1 : 0;
  *out = skip;
  if (count && !skip) *count += 1;
}

void gate(hipStream_t s, const CgScalars* sc, int nr, int mode, int* out, const int* ext,
          int* count) {
  hipLaunchKernelGGL(k_gate, dim3(1), dim3(64), 0, s, sc, nr, mode, ext, out, count);
}

void loss_grad(hipStream_t s, const Level& L, const float* z, const float* r,
               const float* gscale, float* g, int nr, double* partials, int nblocks) {
  hipLaunchKernelGGL(k_loss_grad, dim3(nblocks), dim3(kBlock), 0, s, L, z, r, gscale, g, nr,
                     partials);
}

void pair_make(hipStream_t s, const Level& L, const double* const c64[3], const double* e,
               float* e32, float* r32, double* partials, int nblocks) {
  hipLaunchKernelGGL(k_pair_make, dim3(nblocks), dim3(kBlock), 0, s, L, c64[0], c64[1], c64[2],
                     e, e32, r32, partials);
}

}  // namespace mg
}  // namespace fs
