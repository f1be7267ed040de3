// HIP kernels of the fieldspice MI355X field solve (gfx950, wave64).
//
// Operator: A = G^T diag(w) G + diag(extra) on a tensor-product node grid,
// stored matrix-free as three node-padded edge-weight arrays in an
// x-fastest layout (see fs_internal.hpp).  A prepared system keeps the
// Jacobi-equilibrated, Dirichlet-masked couplings
//     c_e = w_e / sqrt(D_n D_m)   (0 when either end is fixed or pinned)
// so the Krylov operator is  y = x - sum_e c_e x_other  with an implicit unit
// diagonal: the equilibrated free block S A_ff S of
// fieldspice/solvers/poisson.py:364-381, with no mask or diagonal stream.
//
// Reductions are deterministic: each block writes one partial per quantity
// and a single-block finalisation kernel sums them in index order and updates
// the per-RHS scalars on the device, so the host never has to synchronise
// inside the Krylov loop.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fs_internal.hpp"
#include "fs_device.hpp"
#include "fs_launch.hpp"

namespace fs {

// --------------------------------------------------------------------------
// layout permutation between fieldspice's z-fastest arrays and x-fast
// --------------------------------------------------------------------------
// in: C order (A, B, C), c fastest.  out[(c*OB + b)*OA + a] (a fastest).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_permute_in(const T* __restrict__ in,
                                                       T* __restrict__ out, int A, int B,
                                                       int C, int OA, int OB) {
  __shared__ T tile[64][65];
  const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
  const int c0 = blockIdx.x * 64, a0 = blockIdx.y * 64, b = blockIdx.z;
  for (int ra = ty; ra < 64; ra += 4) {
    const int a = a0 + ra, c = c0 + tx;
    if (a < A && c < C) tile[ra][tx] = in[((int64_t)a * B + b) * C + c];
  }
  __syncthreads();
  for (int rc = ty; rc < 64; rc += 4) {
    const int c = c0 + rc, a = a0 + tx;
    if (a < A && c < C) out[((int64_t)c * OB + b) * OA + a] = tile[tx][rc];
  }
}

// in: x-fast [(c*IB + b)*IA + a];  out: C order (A, B, C).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_permute_out(const T* __restrict__ in,
                                                        T* __restrict__ out, int A, int B,
                                                        int C, int IA, int IB) {
  __shared__ T tile[64][65];
  const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
  const int a0 = blockIdx.x * 64, c0 = blockIdx.y * 64, b = blockIdx.z;
  for (int rc = ty; rc < 64; rc += 4) {
    const int c = c0 + rc, a = a0 + tx;
    if (a < A && c < C) tile[rc][tx] = in[((int64_t)c * IB + b) * IA + a];
  }
  __syncthreads();
  for (int ra = ty; ra < 64; ra += 4) {
    const int a = a0 + ra, c = c0 + tx;
    if (a < A && c < C) out[((int64_t)a * B + b) * C + c] = tile[tx][ra];
  }
}

template <typename T>
void permute_in(hipStream_t s, const T* in, T* out, int A, int B, int C, int OA, int OB) {
  dim3 grid((C + 63) / 64, (A + 63) / 64, B);
  hipLaunchKernelGGL(k_permute_in<T>, grid, dim3(kBlock), 0, s, in, out, A, B, C, OA, OB);
}
template <typename T>
void permute_out(hipStream_t s, const T* in, T* out, int A, int B, int C, int IA, int IB) {
  dim3 grid((A + 63) / 64, (C + 63) / 64, B);
  hipLaunchKernelGGL(k_permute_out<T>, grid, dim3(kBlock), 0, s, in, out, A, B, C, IA, IB);
}
template void permute_in<double>(hipStream_t, const double*, double*, int, int, int, int, int);
template void permute_in<uint8_t>(hipStream_t, const uint8_t*, uint8_t*, int, int, int, int, int);
template void permute_in<int32_t>(hipStream_t, const int32_t*, int32_t*, int, int, int, int, int);
template void permute_out<double>(hipStream_t, const double*, double*, int, int, int, int, int);

// --------------------------------------------------------------------------
// edge weights from cell properties: cell_to_edge + edge_mass
// (fieldspice/operators.py:144-237, 302-315) fused in one pass
// --------------------------------------------------------------------------
struct GridMetric {
  const double *h[3];   // primal widths, per axis (N_d entries)
  const double *hd[3];  // dual widths, per axis (N_d + 1 entries)
  int N[3];             // cells per axis
};

__global__ __launch_bounds__(kBlock) void k_cell_to_edge(const double* __restrict__ prop,
                                                         GridMetric g, Dims d, int dir,
                                                         int mode, double* __restrict__ w) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n >= d.n) return;
  int idx[3];
  idx[0] = (int)(n % d.nx);
  idx[1] = (int)((n / d.nx) % d.ny);
  idx[2] = (int)(n / d.sxy);
  if (idx[dir] >= g.N[dir]) {  // past the last cell along dir: no edge
    w[n] = 0.0;
    return;
  }
  const int t0 = dir == 0 ? 1 : 0;
  const int t1 = dir == 2 ? 1 : 2;
  double q[4], wt[4];
  int m = 0;
  for (int s0 = 0; s0 < 2; ++s0) {
    for (int s1 = 0; s1 < 2; ++s1, ++m) {
      const int i0 = idx[t0], i1 = idx[t1];
      // half width of the cell below (s=0) / above (s=1) the node plane
      const double w0 = s0 == 0 ? (i0 > 0 ? 0.5 * g.h[t0][i0 - 1] : 0.0)
                                : (i0 < g.N[t0] ? 0.5 * g.h[t0][i0] : 0.0);
      const double w1 = s1 == 0 ? (i1 > 0 ? 0.5 * g.h[t1][i1 - 1] : 0.0)
                                : (i1 < g.N[t1] ? 0.5 * g.h[t1][i1] : 0.0);
      wt[m] = w0 * w1;
      // edge-replicated padding: missing cells repeat the boundary cell
      int c[3];
      c[dir] = idx[dir];
      c[t0] = min(max(i0 - 1 + s0, 0), g.N[t0] - 1);
      c[t1] = min(max(i1 - 1 + s1, 0), g.N[t1] - 1);
      q[m] = prop[((int64_t)c[2] * g.N[1] + c[1]) * g.N[0] + c[0]];
    }
  }
  const double wsum = ((wt[0] + wt[1]) + wt[2]) + wt[3];
  const double safe = wsum > 0.0 ? wsum : 1.0;
  double e;
  if (mode == FS_AVG_PARALLEL) {
    e = (((q[0] * wt[0] + q[1] * wt[1]) + q[2] * wt[2]) + q[3] * wt[3]) / safe;
  } else if (mode == FS_AVG_HARMONIC) {
    double acc = 0.0;
    for (int k = 0; k < 4; ++k) acc = acc + wt[k] / (q[k] > 0.0 ? q[k] : 2.2250738585072014e-308);
    e = safe / acc;
  } else if (mode == FS_AVG_MIN) {
    e = fmin(fmin(q[0], q[1]), fmin(q[2], q[3]));
  } else {
    e = fmax(fmax(q[0], q[1]), fmax(q[2], q[3]));
  }
  const double area = g.hd[t0][idx[t0]] * g.hd[t1][idx[t1]];
  w[n] = (e * area) / g.h[dir][idx[dir]];
}

void cell_to_edge(hipStream_t s, const double* prop, const double* const h[3],
                  const double* const hd[3], const int N[3], const Dims& d, int dir,
                  int mode, double* w) {
  GridMetric g;
  for (int a = 0; a < 3; ++a) {
    g.h[a] = h[a];
    g.hd[a] = hd[a];
    g.N[a] = N[a];
  }
  const int nb = (int)((d.n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_cell_to_edge, dim3(nb), dim3(kBlock), 0, s, prop, g, d, dir, mode, w);
}

// out = a*x + b*y elementwise (y may be null)
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, double a, const double* x,
                                                  double b, const double* y, double* out) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    double v = x ? a * x[i] : 0.0;
    if (y) v += b * y[i];
    out[i] = v;
  }
}

void axpby(hipStream_t s, int64_t n, double a, const double* x, double b, const double* y,
           double* out) {
  const int nb = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, 4096);
  hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(kBlock), 0, s, n, a, x, b, y, out);
}

// --------------------------------------------------------------------------
// raw operator apply: y = sum_e w_e (x_n - x_m) + extra_n x_n   (FP64)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_apply_raw(const double* __restrict__ x,
                                                      double* __restrict__ y,
                                                      const double* __restrict__ wx,
                                                      const double* __restrict__ wy,
                                                      const double* __restrict__ wz,
                                                      const double* __restrict__ extra,
                                                      Dims d) {
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n >= d.n) return;
  int i, j, k;
  node_coords(d, n, i, j, k);
  y[n] = raw_row(x, wx, wy, wz, extra, d, i, j, k, n);
}

void apply_raw(hipStream_t s, const fs_op* op, const double* x, double* y) {
  const int nb = (int)((op->d.n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_apply_raw, dim3(nb), dim3(kBlock), 0, s, x, y, op->w[0], op->w[1],
                     op->w[2], op->extra, op->d);
}

// --------------------------------------------------------------------------
// system setup: diagonal, masks, gauge, equilibration, couplings
// --------------------------------------------------------------------------
struct SetupCounts {
  unsigned long long n_fixed;
  unsigned long long n_empty;
  unsigned long long n_nonzero_rowsum;  // free rows whose sum is not ~0
  unsigned long long dmax_bits;         // max D over non-empty rows (bit pattern)
  long long gauge_ref;                  // min reference index among argmax rows
};

// D_n = sum of the <= 6 incident weights + extra_n; classify rows.  Counts
// and the maximum are combined per wavefront before one atomic per wave.
__global__ __launch_bounds__(kBlock) void k_setup_diag(const fs_op op, uint8_t* __restrict__ mask,
                                                       double* __restrict__ dsum,
                                                       SetupCounts* cnt) {
#pragma clang fp contract(off)
  const Dims d = op.d;
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  bool is_fixed = false, is_empty = false, nonzero_sum = false;
  unsigned long long dbits = 0ull;
  if (n < d.n) {
    int i, j, k;
    node_coords(d, n, i, j, k);
    // summed in fieldspice's edge order (x, y, z families; the lower edge of
    // each first), as the diagonal of its G^T W G product accumulates, so
    // the diagonal, its maximum (the gauge node) and the equilibration
    // match the reference to the last bit
    const Faces f = faces_of(d, i, j, k, n);
    double ws = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (f.m[2 * a + 1] >= 0) ws = ws + op.w[a][f.e[2 * a + 1]];
      if (f.m[2 * a] >= 0) ws = ws + op.w[a][f.e[2 * a]];
    }
    const double ex = op.extra ? op.extra[n] : 0.0;
    const double D = ws + ex;
    dsum[n] = D;
    if (mask[n] == FS_NODE_DIRICHLET) {
      is_fixed = true;
    } else {
      const double row_abs = 2.0 * ws + fabs(ex);
      if (row_abs == 0.0) {  // coupled to nothing: pinned to 0
        mask[n] = 2;
        is_empty = true;
      } else {
        nonzero_sum = fabs(ex) > 1e-9 * row_abs;
        dbits = (unsigned long long)__double_as_longlong(fabs(D));
      }
    }
  }
  const unsigned long long bf = __ballot(is_fixed), be = __ballot(is_empty),
                           bn = __ballot(nonzero_sum);
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_down(dbits, off, kWave);
    dbits = o > dbits ? o : dbits;
  }
  if (threadIdx.x % kWave == 0) {
    if (bf) atomicAdd(&cnt->n_fixed, (unsigned long long)__popcll(bf));
    if (be) atomicAdd(&cnt->n_empty, (unsigned long long)__popcll(be));
    if (bn) atomicAdd(&cnt->n_nonzero_rowsum, (unsigned long long)__popcll(bn));
    if (dbits) atomicMax(&cnt->dmax_bits, dbits);
  }
}

// Among free non-empty rows with |D| == dmax, the smallest reference index
// (fieldspice picks np.argmax(|diag|), i.e. the first in its C order).
__global__ __launch_bounds__(kBlock) void k_setup_gauge(const uint8_t* __restrict__ mask,
                                                        const double* __restrict__ dsum,
                                                        Dims d, SetupCounts* cnt) {
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n >= d.n) return;
  if (mask[n] != 0) return;
  if ((unsigned long long)__double_as_longlong(fabs(dsum[n])) != cnt->dmax_bits) return;
  const int64_t i = n % d.nx, j = (n / d.nx) % d.ny, k = n / d.sxy;
  const long long ref = (long long)((i * d.ny + j) * d.nz + k);
  atomicMin(&cnt->gauge_ref, ref);
}

__global__ __launch_bounds__(kBlock) void k_setup_pin(uint8_t* mask, Dims d, long long ref) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int64_t i = ref / ((int64_t)d.ny * d.nz);
  const int64_t j = (ref / d.nz) % d.ny;
  const int64_t k = ref % d.nz;
  mask[(k * d.ny + j) * d.nx + i] = 2;
}

__global__ __launch_bounds__(kBlock) void k_setup_equil(const uint8_t* __restrict__ mask,
                                                        const double* __restrict__ dsum,
                                                        double* __restrict__ dhalf,
                                                        double* __restrict__ dinv, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (mask[i] == 0) {
    const double h = sqrt(dsum[i]);
    dhalf[i] = h;
    dinv[i] = 1.0 / h;
  } else {
    dhalf[i] = 0.0;
    dinv[i] = 0.0;
  }
}

// c_e = dinv_n * (w_e * dinv_m); zero whenever either end is not free.
__global__ __launch_bounds__(kBlock) void k_setup_couplings(const fs_op op,
                                                            const double* __restrict__ dinv,
                                                            double* cx, double* cy, double* cz,
                                                            float* fx, float* fy, float* fz) {
  const Dims d = op.d;
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n >= d.n) return;
  int i, j, k;
  node_coords(d, n, i, j, k);
  const double s = dinv[n];
  const Faces f = faces_of(d, i, j, k, n);  // + faces: the edge in slot n
  const double vx = f.m[0] >= 0 ? s * (op.w[0][n] * dinv[f.m[0]]) : 0.0;
  const double vy = f.m[2] >= 0 ? s * (op.w[1][n] * dinv[f.m[2]]) : 0.0;
  const double vz = f.m[4] >= 0 ? s * (op.w[2][n] * dinv[f.m[4]]) : 0.0;
  cx[n] = vx;
  cy[n] = vy;
  cz[n] = vz;
  fx[n] = (float)vx;
  fy[n] = (float)vy;
  fz[n] = (float)vz;
}

int setup_system(fs_sys* sys, double* dsum, void* counts_dev, void* counts_host, int pin) {
  hipStream_t s = sys->stream;
  const Dims d = sys->d;
  const int nb = (int)((d.n + kBlock - 1) / kBlock);
  SetupCounts* cd = (SetupCounts*)counts_dev;
  SetupCounts* ch = (SetupCounts*)counts_host;  // pinned
  *ch = SetupCounts{0, 0, 0, 0, (long long)INT64_MAX};
  if (hipMemcpyAsync(cd, ch, sizeof(SetupCounts), hipMemcpyHostToDevice, s) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(k_setup_diag, dim3(nb), dim3(kBlock), 0, s, *sys->op, sys->mask, dsum, cd);
  if (hipMemcpyAsync(ch, cd, sizeof(SetupCounts), hipMemcpyDeviceToHost, s) != hipSuccess)
    return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  sys->info.n_nodes = d.n;
  sys->info.n_fixed = (int64_t)ch->n_fixed;
  sys->info.n_empty = (int64_t)ch->n_empty;
  const int64_t n_keep = d.n - sys->info.n_fixed - sys->info.n_empty;
  sys->info.singular = (sys->info.n_fixed == 0 && n_keep > 0 && ch->n_nonzero_rowsum == 0) ? 1 : 0;
  sys->info.gauge = -1;
  if (sys->info.singular) {
    if (!pin) return 1;
    hipLaunchKernelGGL(k_setup_gauge, dim3(nb), dim3(kBlock), 0, s, sys->mask, dsum, d, cd);
    if (hipMemcpyAsync(ch, cd, sizeof(SetupCounts), hipMemcpyDeviceToHost, s) != hipSuccess)
      return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (ch->gauge_ref < 0 || ch->gauge_ref >= d.n) return -1;  // no candidate row
    sys->info.gauge = ch->gauge_ref;
    hipLaunchKernelGGL(k_setup_pin, dim3(1), dim3(kBlock), 0, s, sys->mask, d, ch->gauge_ref);
  }
  sys->info.n_free = d.n - sys->info.n_fixed - sys->info.n_empty - (sys->info.gauge >= 0 ? 1 : 0);
  hipLaunchKernelGGL(k_setup_equil, dim3(nb), dim3(kBlock), 0, s, sys->mask, dsum, sys->dhalf,
                     sys->dinv, d.n);
  hipLaunchKernelGGL(k_setup_couplings, dim3(nb), dim3(kBlock), 0, s, *sys->op, sys->dinv,
                     sys->c64[0], sys->c64[1], sys->c64[2], sys->c32[0], sys->c32[1],
                     sys->c32[2]);
  return 0;
}

// --------------------------------------------------------------------------
// Krylov kernels.  NR = compile-time right-hand-side capacity, nr <= NR live.
// Vectors of right-hand side r start at r*vs.
// --------------------------------------------------------------------------

// q = A_s p for each live RHS (z-marching, couplings read once for all RHS),
// fused with the block partials of p.q (alpha is set by k_finalize).
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_stencil_dot(const T* __restrict__ p, T* __restrict__ q,
                                                        const T* __restrict__ cx,
                                                        const T* __restrict__ cy,
                                                        const T* __restrict__ cz, Dims d,
                                                        Tiling tl, VecArgs va, int mode,
                                                        CgScalars* sc, double* partials,
                                                        unsigned long long* prof) {
  if (all_idle(sc, va.nr, mode)) return;
  bool live[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) live[r] = r < va.nr && rhs_live(sc, r, mode);
  if (prof && blockIdx.x == 0 && threadIdx.x == 0) {  // launches with work, live RHS summed
    unsigned long long nl = 0;
    for (int r = 0; r < NR; ++r) nl += live[r] ? 1ull : 0ull;
    prof[0] += 1ull;
    prof[1] += nl;
  }
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;

  int tx, ty, tz;
  if (tile_of_block(tl, tx, ty, tz)) {
    const int i = tx * kTileX + (int)(threadIdx.x % kTileX);
    const int j = ty * kTileY + (int)(threadIdx.x / kTileX);
    const int k0 = tz * tl.kz;
    const int k1 = min(d.nz, k0 + tl.kz);
    if (i < d.nx && j < d.ny) {
      const int64_t base = (int64_t)j * d.nx + i;
      const bool has_e = i + 1 < d.nx, has_w = i > 0, has_n = j + 1 < d.ny, has_s = j > 0;
      // wrap-around neighbours along periodic axes (offsets from the node)
      const bool px = (d.per & 1) && d.nx > 1, py = (d.per & 2) && d.ny > 1,
                 pz = (d.per & 4) && d.nz > 1;
      const int64_t ox = d.nx - 1, oy = (int64_t)(d.ny - 1) * d.nx, oz = (d.nz - 1) * d.sxy;
      // One right-hand side at a time through the tile: the tile's couplings
      // come from HBM for the first live RHS and from L1/L2 for the others,
      // and each pass is the register-light single-RHS z-march.
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (!live[r]) continue;
        const T* __restrict__ pr = p + r * va.vs;
        T* __restrict__ qr = q + r * va.vs;
        int64_t nn = (int64_t)k0 * d.sxy + base;
        T czm = k0 > 0 ? cz[nn - d.sxy] : (pz ? cz[nn + oz] : T(0));
        T pm = k0 > 0 ? pr[nn - d.sxy] : (pz ? pr[nn + oz] : T(0));
        T pc = pr[nn];
        double a = 0.0;
        for (int k = k0; k < k1; ++k, nn += d.sxy) {
          const T cxc = cx[nn], cyc = cy[nn], czc = cz[nn];
          const T cxm = has_w ? cx[nn - 1] : (px ? cx[nn + ox] : T(0));
          const T cym = has_s ? cy[nn - d.nx] : (py ? cy[nn + oy] : T(0));
          const T pu = k + 1 < d.nz ? pr[nn + d.sxy] : (pz ? pr[nn - oz] : T(0));
          const T pe = has_e ? pr[nn + 1] : (px ? pr[nn - ox] : T(0));
          const T pw = has_w ? pr[nn - 1] : (px ? pr[nn + ox] : T(0));
          const T pn = has_n ? pr[nn + d.nx] : (py ? pr[nn - oy] : T(0));
          const T ps = has_s ? pr[nn - d.nx] : (py ? pr[nn + oy] : T(0));
          const T y = pc - (cxc * pe + cxm * pw + cyc * pn + cym * ps + czc * pu + czm * pm);
          qr[nn] = y;
          a += (double)pc * (double)y;
          pm = pc;
          pc = pu;
          czm = czc;
        }
        acc[r] = a;
      }
    }
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr) store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
  }
}

// FP32 variant of k_stencil_dot with W consecutive x nodes per lane (W = 2,
// 3 or 4; rows a multiple of W): one 8-, 12- or 16-byte load per array, a
// wave covers 64*W nodes of a row, and all right-hand sides advance together
// in the z-march so the couplings are read once per node.  The x neighbours
// at the ends of a lane's W nodes come from the adjacent lanes (shuffles); the
// rows above and below come from the neighbouring waves of the block through
// LDS, and the block's two edge rows read the one row outside it.  Every
// global load of a plane is issued before the first use of any of them, with
// no lane-divergent branch around a load: lanes and rows outside the grid
// read the grid's last node and store nothing, and the two x-halo lanes
// select their neighbour from one load every lane makes (256^3, 5 RHS, W = 4:
// 195 -> 176 us, 79% of a kernel streaming the same arrays once).  Node
// offsets are 32-bit (the set-up selects this kernel only when
// n + nx*ny < 2^31).  Stores bypass the caches (q is not read again here).
typedef float f4n __attribute__((ext_vector_type(4)));

// W consecutive floats of one lane
template <int W>
struct Lw {
  float v[W];
};
template <int W>
__device__ __forceinline__ Lw<W> ldw(const float* p) {
  Lw<W> r;
  if constexpr (W == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
  } else if constexpr (W == 3) {
    const float3 t = *reinterpret_cast<const float3*>(p);
    r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z;
  } else {
    const float2 t = *reinterpret_cast<const float2*>(p);
    r.v[0] = t.x; r.v[1] = t.y;
  }
  return r;
}
template <int W>
__device__ __forceinline__ void st_nt(float* p, const Lw<W>& y) {
  if constexpr (W == 4) {
    f4n t;
    t.x = y.v[0]; t.y = y.v[1]; t.z = y.v[2]; t.w = y.v[3];
    __builtin_nontemporal_store(t, reinterpret_cast<f4n*>(p));
  } else {
#pragma unroll
    for (int w = 0; w < W; ++w) __builtin_nontemporal_store(y.v[w], p + w);
  }
}
// component-wise select (the conditional operator on a vector struct selects
// an address and keeps both operands in scratch memory)
template <int W>
__device__ __forceinline__ Lw<W> pickw(bool c, const Lw<W>& a, const Lw<W>& b) {
  Lw<W> r;
#pragma unroll
  for (int w = 0; w < W; ++w) r.v[w] = c ? a.v[w] : b.v[w];
  return r;
}
template <int W>
__device__ __forceinline__ Lw<W> zerow() {
  Lw<W> r;
#pragma unroll
  for (int w = 0; w < W; ++w) r.v[w] = 0.f;
  return r;
}

template <int NR, int W>
__global__ __launch_bounds__(kBlock) void k_stencil_dotw(const float* __restrict__ p,
                                                         float* __restrict__ q,
                                                         const float* __restrict__ cx,
                                                         const float* __restrict__ cy,
                                                         const float* __restrict__ cz, Dims d,
                                                         Tiling tl, VecArgs va, int mode,
                                                         CgScalars* sc, double* partials,
                                                         unsigned long long* prof) {
  if (all_idle(sc, va.nr, mode)) return;
  bool live[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) live[r] = r < va.nr && rhs_live(sc, r, mode);
  if (prof && blockIdx.x == 0 && threadIdx.x == 0) {
    unsigned long long nl = 0;
    for (int r = 0; r < NR; ++r) nl += live[r] ? 1ull : 0ull;
    prof[0] += 1ull;
    prof[1] += nl;
  }
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  __shared__ Lw<W> psh[2][NR][kTileY][kWave];
  int tx, ty, tz;
  if (tile_of_block(tl, tx, ty, tz)) {  // uniform over the block
    const int lane = (int)(threadIdx.x % kWave);
    const int ly = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int i = (tx * kWave + lane) * W;
    const int j = ty * kTileY + ly;  // uniform over the wave
    const bool valid = i < d.nx && j < d.ny;
    const int ic = min(i, d.nx - W), jc = min(j, d.ny - 1);
    const bool has_w = ic > 0, has_e = ic + W < d.nx, has_n = jc + 1 < d.ny, has_s = jc > 0;
    // the x halo: lane 0 reads its west node, the last lane its east node
    const int xo = (lane == 0 && has_w) ? -1 : ((lane == kWave - 1 && has_e) ? W : 0);
    // the one row outside the block an edge wave reads: south for the first
    // wave, north for the last
    const bool own_n = ly + 1 == kTileY, own_s = ly == 0;
    const bool edge = (own_s && has_s) || (own_n && has_n);
    const int k0 = tz * tl.kz;
    const int k1 = min(d.nz, k0 + tl.kz);
    const int sxy = (int)d.sxy, nx = d.nx;
    const int os = has_s ? -nx : 0, oh = own_s ? -nx : nx;
    int nn = k0 * sxy + jc * nx + ic;
    const Lw<W> zw = zerow<W>();
    Lw<W> pm[NR], pc[NR];
    Lw<W> czm = zw;
    {
      const int nd = k0 > 0 ? nn - sxy : nn;
      const Lw<W> czd = ldw<W>(cz + nd);
      Lw<W> pd[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (!live[r]) continue;  // uniform over the launch
        pd[r] = ldw<W>(p + r * va.vs + nd);
        pc[r] = ldw<W>(p + r * va.vs + nn);
      }
      czm = pickw<W>(k0 > 0, czd, zw);
#pragma unroll
      for (int r = 0; r < NR; ++r)
        if (live[r]) pm[r] = pickw<W>(k0 > 0, pd[r], zw);
    }
    for (int k = k0; k < k1; ++k, nn += sxy) {
      const bool up = k + 1 < d.nz;
      const int nu = up ? nn + sxy : nn;
      // ---- every load of this plane ----
      const Lw<W> cxc = ldw<W>(cx + nn), cyc = ldw<W>(cy + nn), czc = ldw<W>(cz + nn);
      const Lw<W> cyl = ldw<W>(cy + nn + os);
      const float cxh = cx[nn + min(xo, 0)];
      Lw<W> pu[NR], ph4[NR];
      float ph[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (!live[r]) continue;
        const float* __restrict__ pr = p + r * va.vs;
        pu[r] = ldw<W>(pr + nu);
        ph[r] = pr[nn + xo];
        if (edge) ph4[r] = ldw<W>(pr + nn + oh);
      }
      // ---- the rows above and below from the neighbouring waves ----
      const int buf = (k - k0) & 1;
#pragma unroll
      for (int r = 0; r < NR; ++r)
        if (live[r]) psh[buf][r][ly][lane] = pc[r];
      __syncthreads();
      // ---- compute ----
      const Lw<W> cym = pickw<W>(has_s, cyl, zw);
      const float cxw_s = __shfl_up(cxc.v[W - 1], 1, kWave);
      const float cxw = has_w ? (lane == 0 ? cxh : cxw_s) : 0.f;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (!live[r]) continue;
        const Lw<W> c = pc[r];
        const Lw<W> u = pickw<W>(up, pu[r], zw);
        const Lw<W> ln = psh[buf][r][min(ly + 1, kTileY - 1)][lane];
        const Lw<W> ls = psh[buf][r][max(ly - 1, 0)][lane];
        const Lw<W> nb = pickw<W>(has_n, pickw<W>(own_n, ph4[r], ln), zw);
        const Lw<W> sb = pickw<W>(has_s, pickw<W>(own_s, ph4[r], ls), zw);
        const float sw = __shfl_up(c.v[W - 1], 1, kWave), se = __shfl_down(c.v[0], 1, kWave);
        const float pw = has_w ? (lane == 0 ? ph[r] : sw) : 0.f;
        const float pe = has_e ? (lane == kWave - 1 ? ph[r] : se) : 0.f;
        const Lw<W> m = pm[r];
        Lw<W> y;
        double dp = 0.0;
        float dps = 0.f;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const float east = w + 1 < W ? c.v[w + 1] : pe;   // x neighbour above
          const float cwest = w > 0 ? cxc.v[w - 1] : cxw;   // coupling to the one below
          const float west = w > 0 ? c.v[w - 1] : pw;
          y.v[w] = c.v[w] - (cxc.v[w] * east + cwest * west + cyc.v[w] * nb.v[w] +
                             cym.v[w] * sb.v[w] + czc.v[w] * u.v[w] + czm.v[w] * m.v[w]);
          // the lane's W products in FP32, their sum accumulated in FP64
          dps = w == 0 ? c.v[0] * y.v[0] : dps + c.v[w] * y.v[w];
        }
        dp = (double)dps;
        acc[r] += valid ? dp : 0.0;
        if (valid) st_nt<W>(q + r * va.vs + nn, y);
        pm[r] = c;
        pc[r] = u;
      }
      czm = czc;
    }
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr) store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
  }
}

// FP64 q = A_s p with 2 consecutive x nodes per lane (16-byte loads; rows a
// multiple of 2): the layout of k_stencil_dot4 for doubles, 128 x nodes per
// wavefront, every right-hand side advanced in the same z-march so the
// couplings are read once per node.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_stencil_dot2(const double* __restrict__ p,
                                                         double* __restrict__ q,
                                                         const double* __restrict__ cx,
                                                         const double* __restrict__ cy,
                                                         const double* __restrict__ cz, Dims d,
                                                         Tiling tl, VecArgs va, int mode,
                                                         CgScalars* sc, double* partials,
                                                         unsigned long long* prof) {
  if (all_idle(sc, va.nr, mode)) return;
  bool live[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) live[r] = r < va.nr && rhs_live(sc, r, mode);
  if (prof && blockIdx.x == 0 && threadIdx.x == 0) {
    unsigned long long nl = 0;
    for (int r = 0; r < NR; ++r) nl += live[r] ? 1ull : 0ull;
    prof[0] += 1ull;
    prof[1] += nl;
  }
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  int tx, ty, tz;
  if (tile_of_block(tl, tx, ty, tz)) {
    const int lane = (int)(threadIdx.x % kWave);
    const int i = (tx * kWave + lane) * 2;
    const int j = ty * kTileY + (int)(threadIdx.x / kWave);
    const int k0 = tz * tl.kz;
    const int k1 = min(d.nz, k0 + tl.kz);
    const bool valid = i < d.nx && j < d.ny;  // i, i+1 inside: nx % 2 == 0
    const int64_t base = valid ? (int64_t)j * d.nx + i : 0;
    const bool has_w = i > 0, has_e = i + 2 < d.nx, has_n = j + 1 < d.ny, has_s = j > 0;
    int64_t nn = (int64_t)k0 * d.sxy + base;
    const double2 z2 = make_double2(0.0, 0.0);
    auto ld2 = [](const double* a) { return *reinterpret_cast<const double2*>(a); };
    double2 pm[NR], pc[NR];
    double2 czm = z2;
    if (valid && k0 > 0) czm = ld2(cz + nn - d.sxy);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      pm[r] = z2;
      pc[r] = z2;
      if (valid && live[r]) {
        if (k0 > 0) pm[r] = ld2(p + r * va.vs + nn - d.sxy);
        pc[r] = ld2(p + r * va.vs + nn);
      }
    }
    for (int k = k0; k < k1; ++k, nn += d.sxy) {
      double2 cxc = z2, cyc = z2, czc = z2, cym = z2;
      if (valid) {
        cxc = ld2(cx + nn);
        cyc = ld2(cy + nn);
        czc = ld2(cz + nn);
        if (has_s) cym = ld2(cy + nn - d.nx);
      }
      // coupling to the west of this lane's first node
      const double cxw_l = __shfl_up(cxc.y, 1, kWave);
      double cxw = 0.0;
      if (valid && has_w) cxw = lane > 0 ? cxw_l : cx[nn - 1];
      const bool up = k + 1 < d.nz;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (!live[r]) continue;  // uniform over the launch
        const double* __restrict__ pr = p + r * va.vs;
        double2 pu = z2, pn = z2, ps = z2;
        if (valid) {
          if (up) pu = ld2(pr + nn + d.sxy);
          if (has_n) pn = ld2(pr + nn + d.nx);
          if (has_s) ps = ld2(pr + nn - d.nx);
        }
        const double2 c = pc[r];
        double pw = __shfl_up(c.y, 1, kWave), pe = __shfl_down(c.x, 1, kWave);
        if (valid) {
          if (!has_w) pw = 0.0;
          else if (lane == 0) pw = pr[nn - 1];
          if (!has_e) pe = 0.0;
          else if (lane == kWave - 1) pe = pr[nn + 2];
        }
        const double2 m = pm[r];
        double2 y;
        y.x = c.x - (cxc.x * c.y + cxw * pw + cyc.x * pn.x + cym.x * ps.x + czc.x * pu.x +
                     czm.x * m.x);
        y.y = c.y - (cxc.y * pe + cxc.x * c.x + cyc.y * pn.y + cym.y * ps.y + czc.y * pu.y +
                     czm.y * m.y);
        if (valid) {
          *reinterpret_cast<double2*>(q + r * va.vs + nn) = y;
          acc[r] += c.x * y.x + c.y * y.y;
        }
        pm[r] = c;
        pc[r] = pu;
      }
      czm = czc;
    }
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr) store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
  }
}

// x += alpha p; r -= alpha q; rho_new = r.r; beta = rho_new/rho; stop tests.
// mode 0: FP64 main iteration (threshold thresh2, sets done).
// mode 1: FP32 cycle of mixed precision (threshold aux1 on the cycle's own
//         rho kept in aux0, sets cycle_done; the iteration cap still ends it).
// DX: x += alpha p is deferred to the next direction update (k_pupdate DX),
// which reads p anyway; the refresh adds the pending term of the last step.
// Launches with work and their live right-hand sides, counted by block 0 when
// the profiling counters are given.
template <int NR>
__device__ __forceinline__ void prof_count(unsigned long long* prof, const bool (&live)[NR]) {
  if (prof && blockIdx.x == 0 && threadIdx.x == 0) {
    unsigned long long nl = 0;
    for (int r = 0; r < NR; ++r) nl += live[r] ? 1ull : 0ull;
    prof[0] += 1ull;
    prof[1] += nl;
  }
}

template <typename T, int NR, bool DX = false>
__global__ __launch_bounds__(kBlock) void k_update(T* __restrict__ x, T* __restrict__ r_,
                                                   const T* __restrict__ p,
                                                   const T* __restrict__ q, VecArgs va,
                                                   int mode, CgScalars* sc, double* partials,
                                                   unsigned long long* prof = nullptr) {
  if (all_idle(sc, va.nr, mode)) return;
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  bool live[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) live[r] = r < va.nr && rhs_live(sc, r, mode);
  prof_count<NR>(prof, live);
  const int64_t stride = (int64_t)gridDim.x * kBlock;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!live[r]) continue;
    const T a = (T)sc[r].alpha;
    T* xr = x + r * va.vs;
    T* rr = r_ + r * va.vs;
    const T* pr = p + r * va.vs;
    const T* qr = q + r * va.vs;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride) {
      if constexpr (!DX) xr[i] += a * pr[i];
      const T rn = rr[i] - a * qr[i];
      rr[i] = rn;
      s += (double)rn * (double)rn;
    }
    acc[r] = s;
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr) store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
  }
}

// The pending step of a DX loop: x += alpha p for every right-hand side,
// with the alpha of its last iteration (0 at the start and after a breakdown).
__global__ __launch_bounds__(kBlock) void k_flush_x(double* __restrict__ x,
                                                    const double* __restrict__ p, VecArgs va,
                                                    const CgScalars* sc) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int r = 0; r < va.nr; ++r) {
    const double a = sc[r].alpha;
    if (a == 0.0 || !isfinite(a)) continue;  // uniform over the launch
    double* xr = x + r * va.vs;
    const double* pr = p + r * va.vs;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride)
      xr[i] += a * pr[i];
  }
}

// p = r + beta p; DX: first x += alpha p with the previous step's alpha (the
// deferred update of k_update DX; alpha is 0 at the start of a cycle)
template <typename T, int NR, bool DX = false>
__global__ __launch_bounds__(kBlock) void k_pupdate(T* __restrict__ p, const T* __restrict__ r_,
                                                    VecArgs va, int mode, const CgScalars* sc,
                                                    T* __restrict__ x = nullptr,
                                                    unsigned long long* prof = nullptr) {
  if (all_idle(sc, va.nr, mode)) return;
  if (prof) {
    bool live[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) live[r] = r < va.nr && rhs_live(sc, r, mode);
    prof_count<NR>(prof, live);
  }
  const int64_t stride = (int64_t)gridDim.x * kBlock;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!(r < va.nr && rhs_live(sc, r, mode))) continue;
    const T b = (T)sc[r].beta;
    T* pr = p + r * va.vs;
    const T* rr = r_ + r * va.vs;
    if constexpr (DX) {
      const T a = (T)sc[r].alpha;
      T* xr = x + r * va.vs;
      for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride) {
        const T pv = pr[i];
        xr[i] += a * pv;
        pr[i] = rr[i] + b * pv;
      }
    } else {
      for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride)
        pr[i] = rr[i] + b * pr[i];
    }
  }
}

// flexible CG with an FP16 preconditioned residual: p = zscale * zh + beta p,
// after the deferred x += alpha p
template <int NR>
__global__ __launch_bounds__(kBlock) void k_pupdate_h(float* __restrict__ p,
                                                      const _Float16* __restrict__ zh,
                                                      VecArgs va, int mode, const CgScalars* sc,
                                                      float* __restrict__ x,
                                                      unsigned long long* prof = nullptr) {
  if (all_idle(sc, va.nr, mode)) return;
  if (prof) {
    bool live[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) live[r] = r < va.nr && rhs_live(sc, r, mode);
    prof_count<NR>(prof, live);
  }
  const int64_t stride = (int64_t)gridDim.x * kBlock;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!(r < va.nr && rhs_live(sc, r, mode))) continue;
    const float b = (float)sc[r].beta, a = (float)sc[r].alpha, hs = zscale(sc[r]);
    float* pr = p + r * va.vs;
    float* xr = x + r * va.vs;
    const _Float16* zr = zh + r * va.vs;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride) {
      const float pv = pr[i];
      xr[i] += a * pv;
      pr[i] = (float)zr[i] * hs + b * pv;
    }
  }
}

// --------------------------------------------------------------------------
// Finalisation of the block partials: one single-block kernel launched right
// after each reducing kernel.  Kernel boundaries make the partials of every
// XCD visible, so no device-scope fence is needed inside the reducing kernels
// (a per-block release fence writes back the whole L2 and quadruples the
// cost of a streaming kernel on this part).
// --------------------------------------------------------------------------
struct FinArgs {
  int op;
  int nr;
  int nb;          // partials per quantity
  int mode;        // 0 FP64 CG, 1 FP32 cycle
  double tol;
  int maxiter;
  double inner_tol2;
};

// Launched with one block per right-hand side: block r owns sc[r].
__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ partials,
                                                     FinArgs fa, CgScalars* sc,
                                                     unsigned long long* prof,
                                                     const int* __restrict__ gate) {
  if (gate && *gate) return;  // a batch's device control skipped this step
  const int r = blockIdx.x;
  if (fa.op == FIN_ALPHA || fa.op == FIN_ALPHA_ZR || fa.op == FIN_BETA || fa.op == FIN_FCG) {
    if (!rhs_live(sc, r, fa.mode)) return;  // uniform over the block
  }
  if (fa.op == FIN_FCG0 && sc[r].done != 0) return;
  const double s0 = sum_partials(partials + (int64_t)r * fa.nb, fa.nb);
  double s1 = 0.0;
  if (fa.op == FIN_RESID || fa.op == FIN_FCG)
    s1 = sum_partials(partials + (int64_t)(fa.nr + r) * fa.nb, fa.nb);
  if (threadIdx.x != 0) return;
  CgScalars& c = sc[r];
  if (fa.op == FIN_ALPHA || fa.op == FIN_ALPHA_ZR) {
    const double pq = s0;
    const double rho = fa.op == FIN_ALPHA_ZR ? c.zr
                                              : (fa.mode == 0 ? c.rho : (fa.mode == 1 ? c.aux0 : c.zr));
    c.pq = pq;
    if (!(pq > 0.0) || !isfinite(pq)) {
      c.done = 3;  // breakdown: A_s not positive definite on p
      c.alpha = 0.0;
    } else {
      c.alpha = rho / pq;
    }
  } else if (fa.op == FIN_BETA) {
    const double rho_new = s0;
    c.iter += 1;
    c.cycle_iter += 1;
    if (fa.mode == 0) {
      c.beta = rho_new / c.rho;
      c.rho = rho_new;
      if (rho_new < c.thresh2) c.done = 1;
      else if (c.iter >= c.maxiter) c.done = 2;
      else if (!isfinite(rho_new)) c.done = 3;
    } else {
      // FP32 cycle; flexible CG (mode 2) takes beta from FIN_FCG instead
      if (fa.mode == 1) c.beta = rho_new / c.aux0;
      c.aux0 = rho_new;
      if (rho_new < c.aux1) c.cycle_done = 1;
      else if (c.iter >= c.maxiter) c.cycle_done = 1;
      else if (!isfinite(rho_new)) c.done = 3;
    }
  } else if (fa.op == FIN_INIT) {
    c.rho = s0;
    c.bnorm2 = s0;
    c.thresh2 = fa.tol * fa.tol * s0;
    c.beta = 0.0;
    c.alpha = 0.0;
    c.iter = 0;
    c.cycle_iter = 0;
    c.maxiter = fa.maxiter;
    c.done = s0 == 0.0 ? 1 : 0;
    c.cycle_done = 0;
    c.refreshes = 0;
  } else if (fa.op == FIN_BNORM) {
    c.aux2 = s0;
  } else if (fa.op == FIN_RESID) {
    c.aux3 = s0;
    c.aux1 = s1;
  } else if (fa.op == FIN_FCG || fa.op == FIN_FCG0) {
    // beta = -<z_new, A p>/<p, A p> keeps the new direction A-orthogonal to
    // the previous one whatever the preconditioner did (flexible CG)
    c.zr = s0;
    c.beta = fa.op == FIN_FCG ? -s1 / c.pq : 0.0;
    if (!(s0 > 0.0) || !isfinite(s0) || !isfinite(c.beta)) c.done = 3;
  } else if (fa.op == FIN_CYCLE) {
    if (c.done != 0) return;
    c.scale = 1.0 / sqrt(c.rho);
    c.aux0 = s0;
    c.aux1 = fmax(fa.inner_tol2, 0.25 * c.thresh2 / c.rho) * s0;
    c.beta = 0.0;
    c.alpha = 0.0;  // no deferred x update pending at a cycle start
    c.cycle_iter = 0;
    c.cycle_done = 0;
  }
}

template <int NR>
static void finalize(hipStream_t st, const double* partials, int op, int nr, int nb, int mode,
                     CgScalars* sc, unsigned long long* prof = nullptr, double tol = 0.0,
                     int maxiter = 0, double inner_tol2 = 0.0, const int* gate = nullptr) {
  FinArgs fa{op, nr, nb, mode, tol, maxiter, inner_tol2};
  hipLaunchKernelGGL(k_finalize, dim3(nr), dim3(kBlock), 0, st, partials, fa, sc, prof, gate);
}

void launch_finalize(hipStream_t st, const double* partials, int op, int nr, int nb, int mode,
                     CgScalars* sc, double inner_tol2) {
  FinArgs fa{op, nr, nb, mode, 0.0, 0, inner_tol2};
  hipLaunchKernelGGL(k_finalize, dim3(nr), dim3(kBlock), 0, st, partials, fa, sc, nullptr,
                     (const int*)nullptr);
}

// --------------------------------------------------------------------------
// right-hand sides, initial state, expansion
// --------------------------------------------------------------------------
// General LinearSystem.solve RHS: b (x-fast, full) -> b_s = b*dinv on free
// rows; ||b||^2 with pinned rows zeroed goes to sc[r].aux2.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_lift_general(const double* __restrict__ b,
                                                         const uint8_t* __restrict__ mask,
                                                         const double* __restrict__ dinv,
                                                         double* __restrict__ bs, VecArgs va,
                                                         double* partials) {
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride) {
    const uint8_t m = mask[i];
    const double di = dinv[i];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r >= va.nr) continue;
      const double v = m == 2 ? 0.0 : b[r * va.vs + i];
      acc[r] += v * v;
      bs[r * va.vs + i] = m == 0 ? v * di : 0.0;
    }
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr) store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
  }
}

// Capacitance RHS for excitation r (terminal id r at 1 V, other fixed at 0):
// b_n = sum over neighbours m with tid[m]==r of w(n,m) on free rows, so
// b_s = dinv * b; ||b_full||^2 = sum b^2 + |terminal r|  -> aux2.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_lift_cap(const fs_op op, const int32_t* __restrict__ tid,
                                                     const uint8_t* __restrict__ mask,
                                                     const double* __restrict__ dinv,
                                                     double* __restrict__ bs, VecArgs va,
                                                     int j0, double* partials) {
  const Dims d = op.d;
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < va.n; n += stride) {
    int i, j, k;
    node_coords(d, n, i, j, k);
    const int t_self = tid[n];
    if (mask[n] != 0) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (r >= va.nr) continue;
        bs[r * va.vs + n] = 0.0;
        if (t_self == j0 + r) acc[r] += 1.0;
      }
      continue;
    }
    int tn[6];
    double wn[6];
    const Faces f = faces_of(d, i, j, k, n);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      tn[q] = f.m[q] >= 0 ? tid[f.m[q]] : -1;
      wn[q] = f.m[q] >= 0 ? op.w[q >> 1][f.e[q]] : 0.0;
    }
    const double di = dinv[n];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r >= va.nr) continue;
      double v = 0.0;
#pragma unroll
      for (int e = 0; e < 6; ++e)
        if (tn[e] == j0 + r) v += wn[e];
      acc[r] += v * v;
      bs[r * va.vs + n] = v * di;
    }
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr) store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
  }
}

// x = 0, p = 0, r = b_s, rho = ||b_s||^2; thresholds.  (FP64 state)
template <int NR>
__global__ __launch_bounds__(kBlock) void k_cg_init(const double* __restrict__ bs,
                                                    double* __restrict__ x,
                                                    double* __restrict__ r_,
                                                    double* __restrict__ p, VecArgs va,
                                                    double* partials) {
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r >= va.nr) continue;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride) {
      const double b = bs[r * va.vs + i];
      x[r * va.vs + i] = 0.0;
      p[r * va.vs + i] = 0.0;
      r_[r * va.vs + i] = b;
      s += b * b;
    }
    acc[r] = s;
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr) store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
  }
}

// phi = x_s*dinv on free rows, the Dirichlet value on fixed rows, 0 pinned.
// Dirichlet values come from b (general solve) or from the terminal map
// (capacitance: 1 on terminal r, 0 elsewhere).
template <int NR>
__global__ __launch_bounds__(kBlock) void k_expand(const double* __restrict__ xs,
                                                   const uint8_t* __restrict__ mask,
                                                   const double* __restrict__ dinv,
                                                   const double* __restrict__ b,
                                                   const int32_t* __restrict__ tid,
                                                   double* __restrict__ phi, VecArgs va,
                                                   int j0) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride) {
    const uint8_t m = mask[i];
    const double di = dinv[i];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r >= va.nr) continue;
      double v;
      if (m == 0) v = xs[r * va.vs + i] * di;
      else if (m == 2) v = 0.0;
      else v = b ? b[r * va.vs + i] : (tid[i] == j0 + r ? 1.0 : 0.0);
      phi[r * va.vs + i] = v;
    }
  }
}

// True equilibrated residual r = b_s - A_s x (FP64 couplings), written to
// `res`, with ||r||^2 -> aux3 and ||r * dhalf||^2 -> aux1 per RHS.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_residual(const double* __restrict__ bs,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ res,
                                                     const double* __restrict__ cx,
                                                     const double* __restrict__ cy,
                                                     const double* __restrict__ cz,
                                                     const double* __restrict__ dhalf, Dims d,
                                                     VecArgs va, double* partials,
                                                     const int* __restrict__ gate) {
  if (gate && *gate) return;  // uniform over the launch
  double acc[2 * NR];
#pragma unroll
  for (int r = 0; r < 2 * NR; ++r) acc[r] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < va.n; n += stride) {
    int i, j, k;
    node_coords(d, n, i, j, k);
    const Faces f = faces_of(d, i, j, k, n);
    const double* c[3] = {cx, cy, cz};
    double cf[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) cf[q] = f.m[q] >= 0 ? c[q >> 1][f.e[q]] : 0.0;
    const double dh = dhalf[n];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r >= va.nr) continue;
      const double* xr = x + r * va.vs;
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q)
        if (f.m[q] >= 0) s += cf[q] * xr[f.m[q]];
      // Dirichlet and pinned rows are identity rows with x_s = b_s = 0 there
      const double rv = bs[r * va.vs + n] - (xr[n] - s);
      res[r * va.vs + n] = rv;
      acc[r] += rv * rv;
      acc[NR + r] += (rv * dh) * (rv * dh);
    }
  }
  block_sum<2 * NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r >= va.nr) continue;
      store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
      store_partial(partials + (int64_t)(va.nr + r) * gridDim.x + blockIdx.x, acc[NR + r]);
    }
  }
}

// FP64 residual r = b_s - A_s x with 2 consecutive x nodes per lane (the
// tiles and z-march of k_stencil_dot2; rows a multiple of 2) and the two
// partial sums of k_residual: ||r||^2 and ||r D^1/2||^2 per right-hand side.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_residual2(const double* __restrict__ bs,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ res,
                                                      const double* __restrict__ cx,
                                                      const double* __restrict__ cy,
                                                      const double* __restrict__ cz,
                                                      const double* __restrict__ dhalf, Dims d,
                                                      Tiling tl, VecArgs va, double* partials,
                                                      const int* __restrict__ gate) {
  if (gate && *gate) return;  // uniform over the launch
  double acc[2 * NR];
#pragma unroll
  for (int a = 0; a < 2 * NR; ++a) acc[a] = 0.0;
  int tx, ty, tz;
  if (tile_of_block(tl, tx, ty, tz)) {
    const int lane = (int)(threadIdx.x % kWave);
    const int i = (tx * kWave + lane) * 2;
    const int j = ty * kTileY + (int)(threadIdx.x / kWave);
    const int k0 = tz * tl.kz;
    const int k1 = min(d.nz, k0 + tl.kz);
    const bool valid = i < d.nx && j < d.ny;  // i, i+1 inside: nx % 2 == 0
    const int64_t base = valid ? (int64_t)j * d.nx + i : 0;
    const bool has_w = i > 0, has_e = i + 2 < d.nx, has_n = j + 1 < d.ny, has_s = j > 0;
    int64_t nn = (int64_t)k0 * d.sxy + base;
    const double2 z2 = make_double2(0.0, 0.0);
    auto ld2 = [](const double* a) { return *reinterpret_cast<const double2*>(a); };
    double2 xm[NR], xc[NR];
    double2 czm = z2;
    if (valid && k0 > 0) czm = ld2(cz + nn - d.sxy);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      xm[r] = z2;
      xc[r] = z2;
      if (valid && r < va.nr) {
        if (k0 > 0) xm[r] = ld2(x + r * va.vs + nn - d.sxy);
        xc[r] = ld2(x + r * va.vs + nn);
      }
    }
    for (int k = k0; k < k1; ++k, nn += d.sxy) {
      double2 cxc = z2, cyc = z2, czc = z2, cym = z2, dh = z2;
      if (valid) {
        cxc = ld2(cx + nn);
        cyc = ld2(cy + nn);
        czc = ld2(cz + nn);
        dh = ld2(dhalf + nn);
        if (has_s) cym = ld2(cy + nn - d.nx);
      }
      const double cxw_l = __shfl_up(cxc.y, 1, kWave);
      double cxw = 0.0;
      if (valid && has_w) cxw = lane > 0 ? cxw_l : cx[nn - 1];
      const bool up = k + 1 < d.nz;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (r >= va.nr) continue;  // uniform over the launch
        const double* __restrict__ xr = x + r * va.vs;
        double2 xu = z2, xn = z2, xs = z2, b = z2;
        if (valid) {
          if (up) xu = ld2(xr + nn + d.sxy);
          if (has_n) xn = ld2(xr + nn + d.nx);
          if (has_s) xs = ld2(xr + nn - d.nx);
          b = ld2(bs + r * va.vs + nn);
        }
        const double2 c = xc[r];
        double xw = __shfl_up(c.y, 1, kWave), xe = __shfl_down(c.x, 1, kWave);
        if (valid) {
          if (!has_w) xw = 0.0;
          else if (lane == 0) xw = xr[nn - 1];
          if (!has_e) xe = 0.0;
          else if (lane == kWave - 1) xe = xr[nn + 2];
        }
        const double2 m = xm[r];
        // Dirichlet and pinned rows are identity rows with x_s = b_s = 0 there
        double2 rv;
        rv.x = b.x - (c.x - (cxc.x * c.y + cxw * xw + cyc.x * xn.x + cym.x * xs.x +
                             czc.x * xu.x + czm.x * m.x));
        rv.y = b.y - (c.y - (cxc.y * xe + cxc.x * c.x + cyc.y * xn.y + cym.y * xs.y +
                             czc.y * xu.y + czm.y * m.y));
        if (valid) {
          *reinterpret_cast<double2*>(res + r * va.vs + nn) = rv;
          acc[r] += rv.x * rv.x + rv.y * rv.y;
          acc[NR + r] += (rv.x * dh.x) * (rv.x * dh.x) + (rv.y * dh.y) * (rv.y * dh.y);
        }
        xm[r] = c;
        xc[r] = xu;
      }
      czm = czc;
    }
  }
  block_sum<2 * NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r >= va.nr) continue;
      store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
      store_partial(partials + (int64_t)(va.nr + r) * gridDim.x + blockIdx.x, acc[NR + r]);
    }
  }
}

// --------------------------------------------------------------------------
// mixed precision: FP32 cycles with FP64 residual refresh
// --------------------------------------------------------------------------
// Start (or restart) an FP32 cycle from the FP64 residual r64:
// scale = 1/||r64||, r32 = r64*scale, d32 = p32 = 0, rho32 = ||r32||^2 (aux0),
// cycle threshold aux1 = max(inner_tol^2, 0.25*thresh2/rho64) * rho32.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_cycle_start(const double* __restrict__ r64,
                                                        float* __restrict__ r32,
                                                        float* __restrict__ d32,
                                                        float* __restrict__ p32, VecArgs va,
                                                        CgScalars* sc, double* partials,
                                                        const int* __restrict__ gate) {
  if (gate && *gate) return;  // uniform over the launch
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!(r < va.nr && sc[r].done == 0)) continue;
    const double sc_r = 1.0 / sqrt(sc[r].rho);
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride) {
      const float v = (float)(r64[r * va.vs + i] * sc_r);
      r32[r * va.vs + i] = v;
      d32[r * va.vs + i] = 0.0f;
      p32[r * va.vs + i] = 0.0f;
      s += (double)v * (double)v;
    }
    acc[r] = s;
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr) store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
  }
}

// x64 += (d32 + alpha p32)/scale for RHS whose cycle ended (then the
// refreshed residual): alpha p32 is the cycle's last step, whose x update
// was deferred to a direction update that did not come
template <int NR>
__global__ __launch_bounds__(kBlock) void k_cycle_accumulate(double* __restrict__ x64,
                                                             const float* __restrict__ d32,
                                                             const float* __restrict__ p32,
                                                             VecArgs va, const CgScalars* sc,
                                                             const int* __restrict__ gate) {
  if (gate && *gate) return;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!(r < va.nr && sc[r].done == 0)) continue;
    const double inv = 1.0 / sc[r].scale;
    const float a = (float)sc[r].alpha;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride)
      x64[r * va.vs + i] += (double)(d32[r * va.vs + i] + a * p32[r * va.vs + i]) * inv;
  }
}

// After k_residual wrote r64 and aux3 = ||r64||^2: new rho, refresh count,
// convergence test (done=1), iteration cap (done=2).
__global__ void k_refresh_finish(int nr, CgScalars* sc, const int* __restrict__ gate) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (gate && *gate) return;
  for (int r = 0; r < nr; ++r) {
    CgScalars& c = sc[r];
    if (c.done != 0) continue;
    c.rho = c.aux3;
    c.refreshes += 1;
    if (c.rho < c.thresh2) c.done = 1;
    else if (c.iter >= c.maxiter) c.done = 2;
    else if (!isfinite(c.rho)) c.done = 3;
  }
}

// --------------------------------------------------------------------------
// device control of a batched solve (one thread each)
// --------------------------------------------------------------------------
// After a chunk: refresh when some right-hand side is still running and every
// running one has ended its FP32 cycle.
__global__ void k_batch_refresh_gate(const CgScalars* sc, int nr, BatchCtl* c) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  bool live = false, waiting = true;
  for (int r = 0; r < nr; ++r) {
    live = live || sc[r].done == 0;
    waiting = waiting && (sc[r].done != 0 || sc[r].cycle_done != 0);
  }
  c->refresh_skip = live && waiting ? 0 : 1;
}

// After the (gated) refresh: start a new cycle when the refresh ran and left
// a right-hand side running; stamp the first replay that finds every
// right-hand side done.
__global__ void k_batch_start_gate(const CgScalars* sc, int nr, int cycles, BatchCtl* c) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  bool live = false;
  for (int r = 0; r < nr; ++r) live = live || sc[r].done == 0;
  c->start_skip = cycles && c->refresh_skip == 0 && live ? 0 : 1;
  if (!live && !c->finished) {
    c->finished = 1;
    c->stamp = (unsigned long long)wall_clock64();
  }
}

// finished systems of the batch -> *out
__global__ void k_batch_count(const BatchCtl* c, int ni, int* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int n = 0;
  for (int i = 0; i < ni; ++i) n += c[i].finished;
  *out = n;
}

__global__ void k_batch_clock(unsigned long long* t) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *t = (unsigned long long)wall_clock64();
}

void launch_batch_refresh_gate(fs_sys* s, int nr, BatchCtl* c) {
  hipLaunchKernelGGL(k_batch_refresh_gate, dim3(1), dim3(64), 0, s->stream, s->sc, nr, c);
}
void launch_batch_start_gate(fs_sys* s, int nr, int cycles, BatchCtl* c) {
  hipLaunchKernelGGL(k_batch_start_gate, dim3(1), dim3(64), 0, s->stream, s->sc, nr, cycles, c);
}
void launch_batch_count(hipStream_t st, const BatchCtl* c, int ni, int* out) {
  hipLaunchKernelGGL(k_batch_count, dim3(1), dim3(64), 0, st, c, ni, out);
}
void launch_batch_clock(hipStream_t st, unsigned long long* t) {
  hipLaunchKernelGGL(k_batch_clock, dim3(1), dim3(64), 0, st, t);
}

// --------------------------------------------------------------------------
// capacitance: C[i][r] = sum_{n in terminal i} (L phi_r)_n
// --------------------------------------------------------------------------
// grid: (blocks_x, nbins); node lists per bin (x-fast indices) in CSR form.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_charge(const fs_op op, const double* __restrict__ phi,
                                                   const int64_t* __restrict__ nodes,
                                                   const int64_t* __restrict__ offs, VecArgs va,
                                                   double* __restrict__ partials) {
  const Dims d = op.d;
  const int bin = blockIdx.y;
  const int64_t lo = offs[bin], hi = offs[bin + 1];
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  for (int64_t t = lo + (int64_t)blockIdx.x * kBlock + threadIdx.x; t < hi;
       t += (int64_t)gridDim.x * kBlock) {
    const int64_t n = nodes[t];
    int i, j, k;
    node_coords(d, n, i, j, k);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r >= va.nr) continue;
      acc[r] += raw_row(phi + r * va.vs, op.w[0], op.w[1], op.w[2], op.extra, d, i, j, k, n);
    }
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr)
        partials[((int64_t)bin * va.nr + r) * gridDim.x + blockIdx.x] = acc[r];
  }
}

// one block: out[bin*nr + r] = ordered sum of the k_charge partials
__global__ __launch_bounds__(kBlock) void k_charge_final(const double* __restrict__ partials,
                                                         int nbins, int nr, int nbx,
                                                         double* __restrict__ out) {
  for (int br = 0; br < nbins * nr; ++br) {
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nbx; b += kBlock) v[0] += partials[(int64_t)br * nbx + b];
    block_sum<1>(v);
    if (threadIdx.x == 0) out[br] = v[0];
  }
}

// --------------------------------------------------------------------------
// Streaming probe (measurement aid): NI FP32 input streams read and NO output
// streams written once, 16 bytes per lane and access, grid-stride, the
// outputs stored non-temporally as the stencil stores q.  out[o] =
// in[o % NI] * sum(in) keeps every load live.
// --------------------------------------------------------------------------
struct StreamArrays {
  const float* in[16];
  float* out[16];
};

template <int NI, int NO>
__global__ __launch_bounds__(kBlock) void k_stream_probe(StreamArrays a, int64_t n4) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n4;
       t += (int64_t)gridDim.x * kBlock) {
    float4 v[NI];
#pragma unroll
    for (int k = 0; k < NI; ++k) v[k] = *reinterpret_cast<const float4*>(a.in[k] + 4 * t);
    float4 sum = v[0];
#pragma unroll
    for (int k = 1; k < NI; ++k) {
      sum.x += v[k].x;
      sum.y += v[k].y;
      sum.z += v[k].z;
      sum.w += v[k].w;
    }
#pragma unroll
    for (int o = 0; o < NO; ++o) {
      const float4 w = v[o % NI];
      f4n y;
      y.x = w.x * sum.x;
      y.y = w.y * sum.y;
      y.z = w.z * sum.z;
      y.w = w.w * sum.w;
      __builtin_nontemporal_store(y, reinterpret_cast<f4n*>(a.out[o] + 4 * t));
    }
  }
}

// the stencil's mix for k live right-hand sides (3 + k read, k written), and a copy
int launch_stream_probe(hipStream_t st, const float* const* in, float* const* out, int nread,
                        int nwrite, int64_t n, int nblocks) {
  StreamArrays a{};
  for (int k = 0; k < nread; ++k) a.in[k] = in[k];
  for (int k = 0; k < nwrite; ++k) a.out[k] = out[k];
  const int64_t n4 = n / 4;
#define FS_SP(I, O)                                                                     \
  if (nread == (I) && nwrite == (O)) {                                                  \
    hipLaunchKernelGGL((k_stream_probe<I, O>), dim3(nblocks), dim3(kBlock), 0, st, a, n4); \
    return 0;                                                                           \
  }
  FS_SP(1, 1)
  FS_SP(4, 1)
  FS_SP(5, 2)
  FS_SP(6, 3)
  FS_SP(7, 4)
  FS_SP(8, 5)
  FS_SP(9, 6)
  FS_SP(10, 7)
  FS_SP(11, 8)
#undef FS_SP
  return -1;
}

// --------------------------------------------------------------------------
// host launchers
// --------------------------------------------------------------------------
// The wide FP32 stencil holds ~2W+6 registers per right-hand side and lane:
// it is instantiated for the exact RHS count, where the other kernels use the
// 1/2/4/8 buckets, and for each width W in {2, 3, 4}.
template <int NR>
static void dotw_t(hipStream_t st, fs_sys* s, const Tiling& tl, int w, const VecArgs& va,
                   int mode, unsigned long long* prof, hipEvent_t ea, hipEvent_t eb) {
#define FS_DW(W_)                                                                               \
  launch_timed(k_stencil_dotw<NR, W_>, dim3(tl.nblocks), st, ea, eb, (const float*)s->p32,       \
               s->q32, (const float*)s->c32[0], (const float*)s->c32[1],                          \
               (const float*)s->c32[2], s->d, tl, va, mode, s->sc, s->partials, prof)
  if (w == 2) FS_DW(2);
  else if (w == 3) FS_DW(3);
  else FS_DW(4);
#undef FS_DW
}
static void launch_dotw(hipStream_t st, fs_sys* s, const Tiling& tl, int w, const VecArgs& va,
                        int mode, unsigned long long* prof, hipEvent_t ea = nullptr,
                        hipEvent_t eb = nullptr) {
  switch (va.nr) {
    case 1: dotw_t<1>(st, s, tl, w, va, mode, prof, ea, eb); break;
    case 2: dotw_t<2>(st, s, tl, w, va, mode, prof, ea, eb); break;
    case 3: dotw_t<3>(st, s, tl, w, va, mode, prof, ea, eb); break;
    case 4: dotw_t<4>(st, s, tl, w, va, mode, prof, ea, eb); break;
    case 5: dotw_t<5>(st, s, tl, w, va, mode, prof, ea, eb); break;
    case 6: dotw_t<6>(st, s, tl, w, va, mode, prof, ea, eb); break;
    case 7: dotw_t<7>(st, s, tl, w, va, mode, prof, ea, eb); break;
    default: dotw_t<8>(st, s, tl, w, va, mode, prof, ea, eb); break;
  }
}

template <int NR>
static void dot2_t(hipStream_t st, fs_sys* s, const VecArgs& va, int mode,
                   unsigned long long* prof, hipEvent_t ea, hipEvent_t eb) {
  launch_timed(k_stencil_dot2<NR>, dim3(s->tl2.nblocks), st, ea, eb, (const double*)s->p, s->q,
               (const double*)s->c64[0], (const double*)s->c64[1], (const double*)s->c64[2],
               s->d, s->tl2, va, mode, s->sc, s->partials, prof);
}
// FP64 stencil of the Krylov loop, instantiated for the exact RHS count
static void launch_dot2(hipStream_t st, fs_sys* s, const VecArgs& va, int mode,
                        unsigned long long* prof, hipEvent_t ea, hipEvent_t eb) {
  switch (va.nr) {
    case 1: dot2_t<1>(st, s, va, mode, prof, ea, eb); break;
    case 2: dot2_t<2>(st, s, va, mode, prof, ea, eb); break;
    case 3: dot2_t<3>(st, s, va, mode, prof, ea, eb); break;
    case 4: dot2_t<4>(st, s, va, mode, prof, ea, eb); break;
    case 5: dot2_t<5>(st, s, va, mode, prof, ea, eb); break;
    case 6: dot2_t<6>(st, s, va, mode, prof, ea, eb); break;
    case 7: dot2_t<7>(st, s, va, mode, prof, ea, eb); break;
    default: dot2_t<8>(st, s, va, mode, prof, ea, eb); break;
  }
}

template <int NR>
static int launch_iteration_t(fs_sys* s, int nr, int mode, ProfSlot* sl) {
  hipStream_t st = s->stream;
  VecArgs va{s->d.n, s->vstride, nr};
  const int vb = s->vec_blocks;
  const int sb = s->tl.nblocks;
  if (mode == 0) {
    // with an island space: p = (I + Z E^-1 Z^T) r + beta p, alpha = zr/pq
    const bool defl = s->defl_n > 0;
    // x += alpha p of the previous step rides on the direction update (DX);
    // the last pending step is added by launch_flush_x when the loop ends
    launch_timed(k_pupdate<double, NR, true>, dim3(vb), st, prof_start(sl, FS_PROF_DIRECTION),
                 prof_stop(sl, FS_PROF_DIRECTION), s->p, (const double*)s->r, va, mode,
                 (const CgScalars*)s->sc, s->x, prof_ctr(s, sl, FS_PROF_DIRECTION));
    if (defl) launch_defl_scatter(s, nr, mode);
    int sbd = sb;
    hipEvent_t ea = prof_start(sl, FS_PROF_STENCIL), eb = prof_stop(sl, FS_PROF_STENCIL);
    unsigned long long* pst = prof_ctr(s, sl, FS_PROF_STENCIL);
    if (s->use_x2) {
      sbd = s->tl2.nblocks;
      launch_dot2(st, s, va, mode, pst, ea, eb);
    } else {
      launch_timed(k_stencil_dot<double, NR>, dim3(sb), st, ea, eb, (const double*)s->p, s->q,
                   (const double*)s->c64[0], (const double*)s->c64[1], (const double*)s->c64[2],
                   s->d, s->tl, va, mode, s->sc, s->partials, pst);
    }
    finalize<NR>(st, s->partials, defl ? FIN_ALPHA_ZR : FIN_ALPHA, nr, sbd, mode, s->sc);
    launch_timed(k_update<double, NR, true>, dim3(vb), st, prof_start(sl, FS_PROF_RESIDUAL),
                 prof_stop(sl, FS_PROF_RESIDUAL), s->x, s->r, (const double*)s->p,
                 (const double*)s->q, va, mode, s->sc, s->partials,
                 prof_ctr(s, sl, FS_PROF_RESIDUAL));
    finalize<NR>(st, s->partials, FIN_BETA, nr, vb, mode, s->sc);
    if (defl) launch_defl_coef(s, nr, mode, 0);
    return 0;
  }
  // FP32 cycle: CG searches along r32, flexible CG along z32 = M(r32)
  const float* dir = mode == 2 ? s->z32 : s->r32;
  if (mode == 2 && pc_z_half(s->pc))
    launch_timed(k_pupdate_h<NR>, dim3(vb), st, prof_start(sl, FS_PROF_DIRECTION),
                 prof_stop(sl, FS_PROF_DIRECTION), s->p32,
                 reinterpret_cast<const _Float16*>(s->z32), va, mode, (const CgScalars*)s->sc,
                 s->x32, prof_ctr(s, sl, FS_PROF_DIRECTION));
  else
    launch_timed(k_pupdate<float, NR, true>, dim3(vb), st, prof_start(sl, FS_PROF_DIRECTION),
                 prof_stop(sl, FS_PROF_DIRECTION), s->p32, dir, va, mode,
                 (const CgScalars*)s->sc, s->x32, prof_ctr(s, sl, FS_PROF_DIRECTION));
  int sbf = sb;
  hipEvent_t ea = prof_start(sl, FS_PROF_STENCIL), eb = prof_stop(sl, FS_PROF_STENCIL);
  unsigned long long* pst = prof_ctr(s, sl, FS_PROF_STENCIL);
  if (s->xw) {
    sbf = s->tlw.nblocks;
    launch_dotw(st, s, s->tlw, s->xw, va, mode, pst, ea, eb);
  } else {
    launch_timed(k_stencil_dot<float, NR>, dim3(sb), st, ea, eb, (const float*)s->p32, s->q32,
                 (const float*)s->c32[0], (const float*)s->c32[1], (const float*)s->c32[2],
                 s->d, s->tl, va, mode, s->sc, s->partials, pst);
  }
  finalize<NR>(st, s->partials, FIN_ALPHA, nr, sbf, mode, s->sc);
  launch_timed(k_update<float, NR, true>, dim3(vb), st, prof_start(sl, FS_PROF_RESIDUAL),
               prof_stop(sl, FS_PROF_RESIDUAL), s->x32, s->r32, (const float*)s->p32,
               (const float*)s->q32, va, mode, s->sc, s->partials,
               prof_ctr(s, sl, FS_PROF_RESIDUAL));
  finalize<NR>(st, s->partials, FIN_BETA, nr, vb, mode, s->sc);
  if (mode == 2) {
    int nbp = 0;
    if (int rc = pc_apply_solver(s, nr, &nbp, sl)) return rc;
    finalize<NR>(st, s->partials, FIN_FCG, nr, nbp, mode, s->sc);
  }
  return 0;
}

#define FS_DISPATCH_NR(nr, FN, ...)        \
  do {                                     \
    if ((nr) <= 1) FN<1>(__VA_ARGS__);     \
    else if ((nr) <= 2) FN<2>(__VA_ARGS__); \
    else if ((nr) <= 4) FN<4>(__VA_ARGS__); \
    else FN<8>(__VA_ARGS__);               \
  } while (0)

void launch_flush_x(fs_sys* s, int nr) {
  VecArgs va{s->d.n, s->vstride, nr};
  hipLaunchKernelGGL(k_flush_x, dim3(s->vec_blocks), dim3(kBlock), 0, s->stream, s->x,
                     (const double*)s->p, va, (const CgScalars*)s->sc);
}

// After k_flush_x and an FP64 residual r = b_s - A_s x (aux3 = ||r||^2):
// the pending step is consumed (alpha = 0 everywhere, so a later flush adds
// nothing), and the right-hand sides in `restart` continue from the true
// residual with a fresh search direction (beta = 0: p = r).
__global__ void k_replace(int nr, unsigned restart, CgScalars* sc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int r = 0; r < nr; ++r) {
    CgScalars& c = sc[r];
    c.alpha = 0.0;
    if (!((restart >> r) & 1u)) continue;
    c.rho = c.aux3;
    c.beta = 0.0;
    c.done = 0;
    c.refreshes += 1;
  }
}

void launch_replace(fs_sys* s, int nr, unsigned restart) {
  hipLaunchKernelGGL(k_replace, dim3(1), dim3(64), 0, s->stream, nr, restart, s->sc);
}

int launch_iteration(fs_sys* s, int nr, int mode, ProfSlot* slot) {
  if (s->defl_n > 0 && mode == 1) return launch_iteration_defl(s, nr, mode, slot);
  if (nr <= 1) return launch_iteration_t<1>(s, nr, mode, slot);
  if (nr <= 2) return launch_iteration_t<2>(s, nr, mode, slot);
  if (nr <= 4) return launch_iteration_t<4>(s, nr, mode, slot);
  return launch_iteration_t<8>(s, nr, mode, slot);
}

template <int NR>
static void launch_init_t(fs_sys* s, int nr, double tol, int maxiter) {
  VecArgs va{s->d.n, s->vstride, nr};
  hipLaunchKernelGGL((k_cg_init<NR>), dim3(s->vec_blocks), dim3(kBlock), 0, s->stream,
                     s->bs, s->x, s->r, s->p, va, s->partials);
  finalize<NR>(s->stream, s->partials, FIN_INIT, nr, s->vec_blocks, 0, s->sc, nullptr, tol,
               maxiter);
}
void launch_init(fs_sys* s, int nr, double tol, int maxiter) {
  FS_DISPATCH_NR(nr, launch_init_t, s, nr, tol, maxiter);
}

template <int NR>
static void launch_lift_general_t(fs_sys* s, int nr, const double* b) {
  VecArgs va{s->d.n, s->vstride, nr};
  hipLaunchKernelGGL((k_lift_general<NR>), dim3(s->vec_blocks), dim3(kBlock), 0,
                     s->stream, b, s->mask, s->dinv, s->bs, va, s->partials);
  finalize<NR>(s->stream, s->partials, FIN_BNORM, nr, s->vec_blocks, 0, s->sc);
}
void launch_lift_general(fs_sys* s, int nr, const double* b) {
  FS_DISPATCH_NR(nr, launch_lift_general_t, s, nr, b);
}

template <int NR>
static void launch_lift_cap_t(fs_sys* s, int nr, const fs_op* L, const int32_t* tid, int j0) {
  VecArgs va{s->d.n, s->vstride, nr};
  hipLaunchKernelGGL((k_lift_cap<NR>), dim3(s->vec_blocks), dim3(kBlock), 0, s->stream,
                     *L, tid, s->mask, s->dinv, s->bs, va, j0, s->partials);
  finalize<NR>(s->stream, s->partials, FIN_BNORM, nr, s->vec_blocks, 0, s->sc);
}
void launch_lift_cap(fs_sys* s, int nr, const fs_op* L, const int32_t* tid, int j0) {
  FS_DISPATCH_NR(nr, launch_lift_cap_t, s, nr, L, tid, j0);
}

template <int NR>
static void launch_expand_t(fs_sys* s, int nr, const double* b, const int32_t* tid,
                            int j0, double* phi) {
  VecArgs va{s->d.n, s->vstride, nr};
  hipLaunchKernelGGL((k_expand<NR>), dim3(s->vec_blocks), dim3(kBlock), 0, s->stream,
                     s->x, s->mask, s->dinv, b, tid, phi, va, j0);
}
void launch_expand(fs_sys* s, int nr, const double* b, const int32_t* tid, int j0,
                   double* phi) {
  FS_DISPATCH_NR(nr, launch_expand_t, s, nr, b, tid, j0, phi);
}

template <int NR>
static void residual2_t(fs_sys* s, const VecArgs& va, double* res) {
  hipLaunchKernelGGL((k_residual2<NR>), dim3(s->tl2.nblocks), dim3(kBlock), 0, s->stream,
                     (const double*)s->bs, (const double*)s->x, res, (const double*)s->c64[0],
                     (const double*)s->c64[1], (const double*)s->c64[2],
                     (const double*)s->dhalf, s->d, s->tl2, va, s->partials, s->gate);
}

template <int NR>
static void launch_residual_t(fs_sys* s, int nr, double* res) {
  VecArgs va{s->d.n, s->vstride, nr};
  if (s->use_x2) {
    // 2 nodes per lane on the stencil tiles, instantiated for the exact count
    switch (nr) {
      case 1: residual2_t<1>(s, va, res); break;
      case 2: residual2_t<2>(s, va, res); break;
      case 3: residual2_t<3>(s, va, res); break;
      case 4: residual2_t<4>(s, va, res); break;
      case 5: residual2_t<5>(s, va, res); break;
      case 6: residual2_t<6>(s, va, res); break;
      case 7: residual2_t<7>(s, va, res); break;
      default: residual2_t<8>(s, va, res); break;
    }
    finalize<NR>(s->stream, s->partials, FIN_RESID, nr, s->tl2.nblocks, 0, s->sc, nullptr, 0.0,
                 0, 0.0, s->gate);
    return;
  }
  hipLaunchKernelGGL((k_residual<NR>), dim3(s->vec_blocks), dim3(kBlock), 0, s->stream,
                     s->bs, s->x, res, s->c64[0], s->c64[1], s->c64[2], s->dhalf, s->d, va,
                     s->partials, s->gate);
  finalize<NR>(s->stream, s->partials, FIN_RESID, nr, s->vec_blocks, 0, s->sc, nullptr, 0.0, 0,
               0.0, s->gate);
}
void launch_residual(fs_sys* s, int nr, double* res) {
  FS_DISPATCH_NR(nr, launch_residual_t, s, nr, res);
}

template <int NR>
static void launch_cycle_start_t(fs_sys* s, int nr, double inner_tol2) {
  VecArgs va{s->d.n, s->vstride, nr};
  hipLaunchKernelGGL((k_cycle_start<NR>), dim3(s->vec_blocks), dim3(kBlock), 0,
                     s->stream, s->r, s->r32, s->x32, s->p32, va, s->sc, s->partials, s->gate);
  finalize<NR>(s->stream, s->partials, FIN_CYCLE, nr, s->vec_blocks, 1, s->sc, nullptr,
               0.0, 0, inner_tol2, s->gate);
}
template <int NR>
static int launch_fcg_start_t(fs_sys* s, int nr, ProfSlot* slot) {
  int nbp = 0;
  if (int rc = pc_apply_solver(s, nr, &nbp, slot)) return rc;
  finalize<NR>(s->stream, s->partials, FIN_FCG0, nr, nbp, 2, s->sc, nullptr, 0.0, 0, 0.0,
               s->gate);
  return 0;
}
int launch_fcg_start(fs_sys* s, int nr, ProfSlot* slot) {
  if (nr <= 1) return launch_fcg_start_t<1>(s, nr, slot);
  if (nr <= 2) return launch_fcg_start_t<2>(s, nr, slot);
  if (nr <= 4) return launch_fcg_start_t<4>(s, nr, slot);
  return launch_fcg_start_t<8>(s, nr, slot);
}
template <int NR>
static void launch_stencil_only_t(fs_sys* s, int nr, const Tiling& tl, int w) {
  VecArgs va{s->d.n, s->vstride, nr};
  if (w > 1)
    launch_dotw(s->stream, s, tl, w, va, 1, nullptr);
  else
    hipLaunchKernelGGL((k_stencil_dot<float, NR>), dim3(tl.nblocks), dim3(kBlock), 0, s->stream,
                       s->p32, s->q32, s->c32[0], s->c32[1], s->c32[2], s->d, tl, va, 1, s->sc,
                       s->partials, nullptr);
}
void launch_stencil_only(fs_sys* s, int nr, const Tiling& tl, int w) {
  FS_DISPATCH_NR(nr, launch_stencil_only_t, s, nr, tl, w);
}

void launch_cycle_start(fs_sys* s, int nr, double inner_tol2) {
  if (s->defl_n > 0) {
    launch_cycle_start_defl(s, nr, inner_tol2);
    return;
  }
  FS_DISPATCH_NR(nr, launch_cycle_start_t, s, nr, inner_tol2);
}

template <int NR>
static void launch_refresh_t(fs_sys* s, int nr) {
  VecArgs va{s->d.n, s->vstride, nr};
  hipLaunchKernelGGL((k_cycle_accumulate<NR>), dim3(s->vec_blocks), dim3(kBlock), 0,
                     s->stream, s->x, s->x32, s->p32, va, s->sc, s->gate);
  launch_residual_t<NR>(s, nr, s->r);
  hipLaunchKernelGGL(k_refresh_finish, dim3(1), dim3(64), 0, s->stream, nr, s->sc, s->gate);
}
void launch_refresh_finish(fs_sys* s, int nr) {
  hipLaunchKernelGGL(k_refresh_finish, dim3(1), dim3(64), 0, s->stream, nr, s->sc, s->gate);
}
void launch_refresh(fs_sys* s, int nr) {
  if (s->defl_n > 0) {
    launch_refresh_defl(s, nr);
    return;
  }
  FS_DISPATCH_NR(nr, launch_refresh_t, s, nr);
}

template <int NR>
static void launch_charge_t(fs_sys* s, int nr, const fs_op* L, const double* phi,
                            const int64_t* nodes, const int64_t* offs, int nbins, int nbx,
                            double* out) {
  VecArgs va{s->d.n, s->vstride, nr};
  hipLaunchKernelGGL((k_charge<NR>), dim3(nbx, nbins), dim3(kBlock), 0, s->stream, *L,
                     phi, nodes, offs, va, s->partials);
  hipLaunchKernelGGL(k_charge_final, dim3(1), dim3(kBlock), 0, s->stream, s->partials,
                     nbins, nr, nbx, out);
}
void launch_charge(fs_sys* s, int nr, const fs_op* L, const double* phi, const int64_t* nodes,
                   const int64_t* offs, int nbins, int nbx, double* out) {
  FS_DISPATCH_NR(nr, launch_charge_t, s, nr, L, phi, nodes, offs, nbins, nbx, out);
}

}  // namespace fs
