// Kernels of the implicit time loops (EQS and thermal steps) and of the
// Joule-heat product.  One field per loop; all node vectors x-fast.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fs_device.hpp"
#include "fs_internal.hpp"
#include "fs_launch.hpp"

namespace fs {

namespace {

__device__ __forceinline__ void node_ijk(const Dims& d, int64_t n, int& i, int& j, int& k) {
  node_coords(d, n, i, j, k);
}

// rhs = Aexp x + sum_s scale[s] * src[s]   (Aexp may be absent)
__global__ __launch_bounds__(kBlock) void k_explicit(const fs_op A, int has_a,
                                                     const double* __restrict__ x,
                                                     const double* __restrict__ src,
                                                     const double* __restrict__ scale, int nsrc,
                                                     int64_t vs, double* __restrict__ rhs) {
  const Dims d = A.d;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < d.n;
       n += (int64_t)gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(d, n, i, j, k);
    double y = has_a ? raw_row(x, A.w[0], A.w[1], A.w[2], A.extra, d, i, j, k, n) : 0.0;
    for (int s = 0; s < nsrc; ++s) y += scale[s] * src[(int64_t)s * vs + n];
    rhs[n] = y;
  }
}

// Dirichlet lifting with per-group values (apply_dirichlet,
// fieldspice/operators.py:395-407): free rows b = rhs + sum over fixed
// neighbours of w*v, b_s = b*dinv; ||b||^2 (Dirichlet rows carry their value,
// pinned rows 0) into the partials.
__global__ __launch_bounds__(kBlock) void k_lift_groups(const fs_op K,
                                                        const double* __restrict__ rhs,
                                                        const int32_t* __restrict__ group,
                                                        const double* __restrict__ vals,
                                                        const uint8_t* __restrict__ mask,
                                                        const double* __restrict__ dinv,
                                                        double* __restrict__ bs,
                                                        double* __restrict__ partials) {
  const Dims d = K.d;
  double acc[1] = {0.0};
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < d.n;
       n += (int64_t)gridDim.x * kBlock) {
    const uint8_t m = mask[n];
    if (m == 2) {
      bs[n] = 0.0;
      continue;
    }
    if (m == 1) {
      const double v = vals[group[n]];
      acc[0] += v * v;
      bs[n] = 0.0;
      continue;
    }
    int i, j, k;
    node_ijk(d, n, i, j, k);
    double b = rhs ? rhs[n] : 0.0;
    int g;
    if (i + 1 < d.nx && (g = group[n + 1]) >= 0) b += K.w[0][n] * vals[g];
    if (i > 0 && (g = group[n - 1]) >= 0) b += K.w[0][n - 1] * vals[g];
    if (j + 1 < d.ny && (g = group[n + d.nx]) >= 0) b += K.w[1][n] * vals[g];
    if (j > 0 && (g = group[n - d.nx]) >= 0) b += K.w[1][n - d.nx] * vals[g];
    if (k + 1 < d.nz && (g = group[n + d.sxy]) >= 0) b += K.w[2][n] * vals[g];
    if (k > 0 && (g = group[n - d.sxy]) >= 0) b += K.w[2][n - d.sxy] * vals[g];
    acc[0] += b * b;
    bs[n] = b * dinv[n];
  }
  block_sum<1>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

// Warm start: x_s = g*dhalf on free rows, r = b_s - A_s x_s, p = 0, with the
// guess g = x_prev, or with x_prev2 given the linear extrapolation
// g = 2 x_prev - x_prev2 of the last two fields.  scaled: x_prev is already
// x_s (scaled variables, zero off the free rows).
// Partials: ||r||^2 (quantity 0) and ||b_s||^2 (quantity 1).
__global__ __launch_bounds__(kBlock) void k_init_warm(const double* __restrict__ xprev,
                                                      const double* __restrict__ xprev2,
                                                      int scaled,
                                                      const double* __restrict__ dhalf,
                                                      const double* __restrict__ cx,
                                                      const double* __restrict__ cy,
                                                      const double* __restrict__ cz, Dims d,
                                                      const double* __restrict__ bs,
                                                      double* __restrict__ xs,
                                                      double* __restrict__ r,
                                                      double* __restrict__ p,
                                                      double* __restrict__ partials) {
  double acc[2] = {0.0, 0.0};
  // dhalf = 0 off the free rows
  auto guess = [&](int64_t m) {
    if (scaled) return xprev[m];
    const double g = xprev2 ? 2.0 * xprev[m] - xprev2[m] : xprev[m];
    return g * dhalf[m];
  };
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < d.n;
       n += (int64_t)gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(d, n, i, j, k);
    const double xc = guess(n);
    double s = 0.0;
    if (i + 1 < d.nx) s += cx[n] * guess(n + 1);
    if (i > 0) s += cx[n - 1] * guess(n - 1);
    if (j + 1 < d.ny) s += cy[n] * guess(n + d.nx);
    if (j > 0) s += cy[n - d.nx] * guess(n - d.nx);
    if (k + 1 < d.nz) s += cz[n] * guess(n + d.sxy);
    if (k > 0) s += cz[n - d.sxy] * guess(n - d.sxy);
    const double b = bs[n];
    const double rv = b - (xc - s);
    xs[n] = xc;
    r[n] = rv;
    p[n] = 0.0;
    acc[0] += rv * rv;
    acc[1] += b * b;
  }
  block_sum<2>(acc);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = acc[0];
    partials[gridDim.x + blockIdx.x] = acc[1];
  }
}

__global__ __launch_bounds__(kBlock) void k_init_warm_final(const double* __restrict__ partials,
                                                            int nb, double tol, int maxiter,
                                                            CgScalars* sc) {
  const double s0 = sum_partials(partials, nb);
  const double s1 = sum_partials(partials + nb, nb);
  if (threadIdx.x != 0) return;
  CgScalars& c = sc[0];
  c.rho = s0;
  c.bnorm2 = s1;
  c.thresh2 = tol * tol * s1;
  c.beta = 0.0;
  c.alpha = 0.0;
  c.iter = 0;
  c.cycle_iter = 0;
  c.maxiter = maxiter;
  c.done = (s1 == 0.0 || s0 < c.thresh2) ? 1 : 0;
  c.cycle_done = 0;
  c.refreshes = 0;
}

__global__ __launch_bounds__(kBlock) void k_bnorm_final(const double* __restrict__ partials,
                                                        int nb, CgScalars* sc) {
  const double s0 = sum_partials(partials, nb);
  if (threadIdx.x == 0) sc[0].aux2 = s0;
}

// x = x_s*dinv on free rows, the group value on Dirichlet rows, 0 pinned.
__global__ __launch_bounds__(kBlock) void k_expand_groups(const double* __restrict__ xs,
                                                          const uint8_t* __restrict__ mask,
                                                          const double* __restrict__ dinv,
                                                          const int32_t* __restrict__ group,
                                                          const double* __restrict__ vals,
                                                          int64_t n, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const uint8_t m = mask[i];
    x[i] = m == 0 ? xs[i] * dinv[i] : (m == 1 ? vals[group[i]] : 0.0);
  }
}

// Probe currents and voltages: grid (nbx, nprobe); partials
// [(probe*2 + q)*nbx + bx], q = 0 current, 1 voltage.
__global__ __launch_bounds__(kBlock) void k_probe(const fs_op A1, const fs_op A2, int has_b,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ xprev,
                                                  const double* __restrict__ w,
                                                  const int64_t* __restrict__ nodes,
                                                  const int64_t* __restrict__ offs,
                                                  double* __restrict__ partials) {
  const Dims d = A1.d;
  const int pb = blockIdx.y;
  double acc[2] = {0.0, 0.0};
  for (int64_t t = offs[pb] + (int64_t)blockIdx.x * kBlock + threadIdx.x; t < offs[pb + 1];
       t += (int64_t)gridDim.x * kBlock) {
    const int64_t n = nodes[t];
    int i, j, k;
    node_ijk(d, n, i, j, k);
    double cur = raw_row_diff(x, xprev, A1.w[0], A1.w[1], A1.w[2], A1.extra, d, i, j, k, n);
    if (has_b) cur += raw_row(x, A2.w[0], A2.w[1], A2.w[2], A2.extra, d, i, j, k, n);
    acc[0] += cur;
    acc[1] += w[n] * x[n];
  }
  block_sum<2>(acc);
  if (threadIdx.x == 0) {
    partials[((int64_t)pb * 2 + 0) * gridDim.x + blockIdx.x] = acc[0];
    partials[((int64_t)pb * 2 + 1) * gridDim.x + blockIdx.x] = acc[1];
  }
}

// one block per probe: out_i[pb], out_v[pb] of this step
__global__ __launch_bounds__(kBlock) void k_probe_final(const double* __restrict__ partials,
                                                        int nbx, double* out_i,
                                                        double* out_v) {
  const int pb = blockIdx.x;
  const double si = sum_partials(partials + ((int64_t)pb * 2 + 0) * nbx, nbx);
  const double sv = sum_partials(partials + ((int64_t)pb * 2 + 1) * nbx, nbx);
  if (threadIdx.x == 0) {
    if (out_i) out_i[pb] = si;
    out_v[pb] = sv;
  }
}

// Per-step reductions: sum x, max x, x^T Qa x, x^T Qb x (each edge once).
__device__ __forceinline__ double quad_row(const double* __restrict__ x, const fs_op& Q,
                                           const Dims& d, int i, int j, int k, int64_t n) {
  const double xc = x[n];
  double v = 0.0;
  if (i + 1 < d.nx) { const double t = xc - x[n + 1]; v += Q.w[0][n] * t * t; }
  if (j + 1 < d.ny) { const double t = xc - x[n + d.nx]; v += Q.w[1][n] * t * t; }
  if (k + 1 < d.nz) { const double t = xc - x[n + d.sxy]; v += Q.w[2][n] * t * t; }
  if (Q.extra) v += Q.extra[n] * xc * xc;
  return v;
}

__global__ __launch_bounds__(kBlock) void k_step_stats(const double* __restrict__ x, Dims d,
                                                       const fs_op Qa, int has_a, const fs_op Qb,
                                                       int has_b, double* __restrict__ partials) {
  double acc[3] = {0.0, 0.0, 0.0};
  double mx = -INFINITY;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < d.n;
       n += (int64_t)gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(d, n, i, j, k);
    const double xc = x[n];
    acc[0] += xc;
    mx = fmax(mx, xc);
    if (has_a) acc[1] += quad_row(x, Qa, d, i, j, k, n);
    if (has_b) acc[2] += quad_row(x, Qb, d, i, j, k, n);
  }
  block_sum<3>(acc);
  for (int off = kWave / 2; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, kWave));
  __shared__ double smx[kBlock / kWave];
  if (threadIdx.x % kWave == 0) smx[threadIdx.x / kWave] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = smx[0];
    for (int w = 1; w < kBlock / kWave; ++w) m = fmax(m, smx[w]);
    partials[0 * gridDim.x + blockIdx.x] = acc[0];
    partials[1 * gridDim.x + blockIdx.x] = m;
    partials[2 * gridDim.x + blockIdx.x] = acc[1];
    partials[3 * gridDim.x + blockIdx.x] = acc[2];
  }
}

__global__ __launch_bounds__(kBlock) void k_step_stats_final(const double* __restrict__ partials,
                                                             int nb, double* out) {
  const double s0 = sum_partials(partials, nb);
  const double s2 = sum_partials(partials + 2 * (int64_t)nb, nb);
  const double s3 = sum_partials(partials + 3 * (int64_t)nb, nb);
  double m = -INFINITY;
  for (int b = threadIdx.x; b < nb; b += kBlock) m = fmax(m, partials[nb + b]);
  for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, kWave));
  __shared__ double smx[kBlock / kWave];
  if (threadIdx.x % kWave == 0) smx[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = smx[0];
    for (int w = 1; w < kBlock / kWave; ++w) mm = fmax(mm, smx[w]);
    out[0] = s0;
    out[1] = mm;
    out[2] = s2;
    out[3] = s3;
  }
}

// q_n = 1/2 sum over the <= 6 edges at n of w_e (x_n - x_m)^2
__global__ __launch_bounds__(kBlock) void k_edge_power(const fs_op A, const double* __restrict__ x,
                                                       double* __restrict__ q) {
  const Dims d = A.d;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < d.n;
       n += (int64_t)gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(d, n, i, j, k);
    const double xc = x[n];
    double v = 0.0;
    if (i + 1 < d.nx) { const double t = xc - x[n + 1]; v += A.w[0][n] * t * t; }
    if (i > 0) { const double t = xc - x[n - 1]; v += A.w[0][n - 1] * t * t; }
    if (j + 1 < d.ny) { const double t = xc - x[n + d.nx]; v += A.w[1][n] * t * t; }
    if (j > 0) { const double t = xc - x[n - d.nx]; v += A.w[1][n - d.nx] * t * t; }
    if (k + 1 < d.nz) { const double t = xc - x[n + d.sxy]; v += A.w[2][n] * t * t; }
    if (k > 0) { const double t = xc - x[n - d.sxy]; v += A.w[2][n - d.sxy] * t * t; }
    q[n] = 0.5 * v;
  }
}

int grid_for(int64_t n, int cap) {
  const int64_t nb = (n + kBlock - 1) / kBlock;
  return (int)(nb < cap ? (nb > 0 ? nb : 1) : cap);
}


// ----------------------------------------------------------------------------
// Initial guesses from earlier solutions (projection method of P. F. Fischer,
// "Projection techniques for iterative solution of Ax = b with successive
// right-hand sides", CMAME 163, 1998): X holds nb A_s-orthonormal vectors of
// the scaled space (zero off the free rows) and AX = A_s X.  The guess for a
// right-hand side b is x0 = X X^T b, the A_s-norm best approximation of the
// solution in span(X).
// ----------------------------------------------------------------------------

// partials[i * gridDim.x + block] = V_i . w  for i < nb
__global__ __launch_bounds__(kBlock) void k_proj_dots(const double* __restrict__ V, int64_t vs,
                                                      int nb, const double* __restrict__ w,
                                                      int64_t n, double* __restrict__ partials) {
  double acc[kProjMax];
#pragma unroll
  for (int i = 0; i < kProjMax; ++i) acc[i] = 0.0;
  for (int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x; m < n;
       m += (int64_t)gridDim.x * kBlock) {
    const double wv = w[m];
#pragma unroll
    for (int i = 0; i < kProjMax; ++i)
      if (i < nb) acc[i] += V[i * vs + m] * wv;
  }
  block_sum<kProjMax>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kProjMax; ++i)
      if (i < nb) partials[(int64_t)i * gridDim.x + blockIdx.x] = acc[i];
  }
}

// coef[i] = sum of the nblocks partials of quantity i (one block per i)
__global__ __launch_bounds__(kBlock) void k_proj_coef(const double* __restrict__ partials,
                                                      int nblocks, double* __restrict__ coef) {
  const double v = sum_partials(partials + (int64_t)blockIdx.x * nblocks, nblocks);
  if (threadIdx.x == 0) coef[blockIdx.x] = v;
}

// out = sum_i coef[i] X_i   (sub: out -= sum_i coef[i] X_i)
__global__ __launch_bounds__(kBlock) void k_proj_combine(const double* __restrict__ X, int64_t vs,
                                                         int nb, const double* __restrict__ coef,
                                                         int sub, int64_t n,
                                                         double* __restrict__ out) {
  double c[kProjMax];
#pragma unroll
  for (int i = 0; i < kProjMax; ++i) c[i] = i < nb ? coef[i] : 0.0;
  for (int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x; m < n;
       m += (int64_t)gridDim.x * kBlock) {
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < kProjMax; ++i)
      if (i < nb) v += c[i] * X[i * vs + m];
    out[m] = sub ? out[m] - v : v;
  }
}

// out = a - b (b may be null: out = a)
__global__ __launch_bounds__(kBlock) void k_diff(const double* __restrict__ a,
                                                 const double* __restrict__ b, int64_t n,
                                                 double* __restrict__ out) {
  for (int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x; m < n;
       m += (int64_t)gridDim.x * kBlock)
    out[m] = b ? a[m] - b[m] : a[m];
}

// y = A_s v (equilibrated: unit diagonal on free rows, masked couplings;
// v is zero off the free rows) and partials of v . y
__global__ __launch_bounds__(kBlock) void k_apply_s(const double* __restrict__ cx,
                                                    const double* __restrict__ cy,
                                                    const double* __restrict__ cz, Dims d,
                                                    const double* __restrict__ v,
                                                    double* __restrict__ y,
                                                    double* __restrict__ partials) {
  double acc[1] = {0.0};
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < d.n;
       n += (int64_t)gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(d, n, i, j, k);
    double s = 0.0;
    if (i + 1 < d.nx) s += cx[n] * v[n + 1];
    if (i > 0) s += cx[n - 1] * v[n - 1];
    if (j + 1 < d.ny) s += cy[n] * v[n + d.nx];
    if (j > 0) s += cy[n - d.nx] * v[n - d.nx];
    if (k + 1 < d.nz) s += cz[n] * v[n + d.sxy];
    if (k > 0) s += cz[n - d.sxy] * v[n - d.sxy];
    const double yv = v[n] - s;
    y[n] = yv;
    acc[0] += v[n] * yv;
  }
  block_sum<1>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

// A_s-normalise a new basis vector in place: x /= sqrt(norm2), ax likewise,
// when norm2 = x . A_s x (device slot) is positive and finite
__global__ __launch_bounds__(kBlock) void k_proj_normalise(const double* __restrict__ norm2,
                                                           int64_t n, double* __restrict__ x,
                                                           double* __restrict__ ax) {
  const double q = *norm2;
  if (!(q > 0.0) || !(q < INFINITY)) return;  // uniform over the launch
  const double f = 1.0 / sqrt(q);
  for (int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x; m < n;
       m += (int64_t)gridDim.x * kBlock) {
    x[m] *= f;
    ax[m] *= f;
  }
}

}  // namespace

void launch_explicit(fs_sys* s, const fs_op* A, const double* x, const double* src,
                     const double* scale, int nsrc, int64_t vs, double* rhs) {
  fs_op dummy = A ? *A : *s->op;
  hipLaunchKernelGGL(k_explicit, dim3(s->vec_blocks), dim3(kBlock), 0, s->stream, dummy,
                     A ? 1 : 0, x, src, scale, nsrc, vs, rhs);
}

void launch_lift_groups(fs_sys* s, const double* rhs, const int32_t* group, const double* vals) {
  hipLaunchKernelGGL(k_lift_groups, dim3(s->vec_blocks), dim3(kBlock), 0, s->stream,
                     *s->op, rhs, group, vals, s->mask, s->dinv, s->bs, s->partials);
  hipLaunchKernelGGL(k_bnorm_final, dim3(1), dim3(kBlock), 0, s->stream, s->partials,
                     s->vec_blocks, s->sc);
}

void launch_init_warm(fs_sys* s, const double* xprev, const double* xprev2, double tol,
                      int maxiter, int scaled, int rhs) {
  const int64_t o = (int64_t)rhs * s->vstride;
  hipLaunchKernelGGL(k_init_warm, dim3(s->vec_blocks), dim3(kBlock), 0, s->stream, xprev,
                     xprev2, scaled, s->dhalf, s->c64[0], s->c64[1], s->c64[2], s->d,
                     s->bs + o, s->x + o, s->r + o, s->p + o, s->partials);
  hipLaunchKernelGGL(k_init_warm_final, dim3(1), dim3(kBlock), 0, s->stream, s->partials,
                     s->vec_blocks, tol, maxiter, s->sc + rhs);
}


// x0 = X X^T b_s, then the usual start from x0 (r = b_s - A_s x0 computed
// from the operator, not from AX)
void launch_proj_guess(fs_sys* s, const double* X, const double* AX, int64_t vs, int nb,
                       double* coef, double* x0, double tol, int maxiter) {
  (void)AX;
  const int vb = s->vec_blocks;
  hipLaunchKernelGGL(k_proj_dots, dim3(vb), dim3(kBlock), 0, s->stream, X, vs, nb,
                     (const double*)s->bs, s->d.n, s->partials);
  hipLaunchKernelGGL(k_proj_coef, dim3(nb), dim3(kBlock), 0, s->stream,
                     (const double*)s->partials, vb, coef);
  hipLaunchKernelGGL(k_proj_combine, dim3(vb), dim3(kBlock), 0, s->stream, X, vs, nb,
                     (const double*)coef, 0, s->d.n, x0);
  launch_init_warm(s, x0, nullptr, tol, maxiter, 1);
}

// Append v (overwritten) to the basis as slot nb: two classical Gram-Schmidt
// passes in the A_s inner product (h = AX^T v), then AX[nb] = A_s v from the
// operator and norm2 = v . A_s v into *norm2; the vector is normalised on
// the device when norm2 > 0.  The caller reads norm2 to count the slot.
void launch_proj_append(fs_sys* s, double* X, double* AX, int64_t vs, int nb, double* coef,
                        double* norm2) {
  const int vb = s->vec_blocks;
  double* v = X + (int64_t)nb * vs;
  double* av = AX + (int64_t)nb * vs;
  for (int pass = 0; pass < 2 && nb > 0; ++pass) {
    hipLaunchKernelGGL(k_proj_dots, dim3(vb), dim3(kBlock), 0, s->stream, (const double*)AX, vs,
                       nb, (const double*)v, s->d.n, s->partials);
    hipLaunchKernelGGL(k_proj_coef, dim3(nb), dim3(kBlock), 0, s->stream,
                       (const double*)s->partials, vb, coef);
    hipLaunchKernelGGL(k_proj_combine, dim3(vb), dim3(kBlock), 0, s->stream, (const double*)X,
                       vs, nb, (const double*)coef, 1, s->d.n, v);
  }
  hipLaunchKernelGGL(k_apply_s, dim3(vb), dim3(kBlock), 0, s->stream, (const double*)s->c64[0],
                     (const double*)s->c64[1], (const double*)s->c64[2], s->d, (const double*)v,
                     av, s->partials);
  hipLaunchKernelGGL(k_proj_coef, dim3(1), dim3(kBlock), 0, s->stream, (const double*)s->partials,
                     vb, norm2);
  hipLaunchKernelGGL(k_proj_normalise, dim3(vb), dim3(kBlock), 0, s->stream,
                     (const double*)norm2, s->d.n, v, av);
}

void launch_diff(fs_sys* s, const double* a, const double* b, double* out) {
  hipLaunchKernelGGL(k_diff, dim3(s->vec_blocks), dim3(kBlock), 0, s->stream, a, b, s->d.n, out);
}

void launch_expand_groups(fs_sys* s, const int32_t* group, const double* vals, double* x) {
  hipLaunchKernelGGL(k_expand_groups, dim3(s->vec_blocks), dim3(kBlock), 0, s->stream,
                     s->x, s->mask, s->dinv, group, vals, s->d.n, x);
}

void launch_probe(fs_sys* s, const fs_op* A1, const fs_op* A2, const double* x,
                  const double* xprev, const double* w, const int64_t* nodes,
                  const int64_t* offs, int nprobe, int nbx, double* out_i, double* out_v) {
  fs_op b = A2 ? *A2 : *A1;
  hipLaunchKernelGGL(k_probe, dim3(nbx, nprobe), dim3(kBlock), 0, s->stream, *A1, b,
                     A2 ? 1 : 0, x, xprev, w, nodes, offs, s->partials);
  hipLaunchKernelGGL(k_probe_final, dim3(nprobe), dim3(kBlock), 0, s->stream, s->partials,
                     nbx, out_i, out_v);
}

void launch_step_stats(fs_sys* s, const double* x, const fs_op* Qa, const fs_op* Qb,
                       double* out) {
  fs_op a = Qa ? *Qa : *s->op, b = Qb ? *Qb : *s->op;
  hipLaunchKernelGGL(k_step_stats, dim3(s->vec_blocks), dim3(kBlock), 0, s->stream, x,
                     s->d, a, Qa ? 1 : 0, b, Qb ? 1 : 0, s->partials);
  hipLaunchKernelGGL(k_step_stats_final, dim3(1), dim3(kBlock), 0, s->stream, s->partials,
                     s->vec_blocks, out);
}

void launch_edge_power(hipStream_t st, const fs_op* A, const double* x, double* q, int ncu) {
  hipLaunchKernelGGL(k_edge_power, dim3(grid_for(A->d.n, 8 * ncu)), dim3(kBlock), 0, st, *A, x,
                     q);
}

}  // namespace fs
