// Island coarse space for high-contrast systems.
//
// A conductor that touches no Dirichlet row (a floating copper island in
// SiO2 under an EQS step, a high-conductivity inclusion in a thermal model)
// is a near-null mode of the node Laplacian: inside the island the couplings
// are ~1e6 times those to its surroundings, so Jacobi-CG needs a number of
// iterations that grows with that contrast, and FP32 arithmetic cannot even
// represent the mode's Rayleigh quotient (it lies below the rounding of
// x - sum c x).  Neither fieldspice nor SciPy's CG has a counterpart; the
// reference's Jacobi-CG hits its max(500, 10 sqrt(n)) cap on such systems
// (fieldspice/solvers/poisson.py:487-589).
//
// The coarse space holds one vector per island, Z_I = D^1/2 1_I in the
// equilibrated space, and E = Z^T A_s Z (= 1_I^T L_ff 1_J of the unscaled
// masked operator), inverted once on the host.  The preconditioner is the
// additive M = I + Z E^-1 Z^T (symmetric positive definite, so plain PCG
// applies): per iteration one segmented reduction over the island nodes
// (chunk partials, then one wave per island), the dense E^-1 product (one
// wave per row), a fixed-order sum for zr and beta, and one scatter.
//
// FP32 cycles of a deflated system run on the unscaled operator in
// difference form, y_n = e_n u_n + sum_e w_e (u_n - u_m) with u = D^-1/2 x_s.
// Rounding then perturbs each edge flux relatively, which keeps the
// near-null structure of the islands, whereas rounding x - sum c x perturbs
// the row sums by the FP32 unit and hides it.  In those variables the
// preconditioner is M_u = D^-1 + 1_I E^-1 1_I^T (the same M).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "fs_common.hpp"
#include "fs_device.hpp"
#include "fs_launch.hpp"

namespace fs {

// v . Z per chunk of island nodes: weights dhalf (equilibrated space, ZW) or
// 1 (unscaled).  One block per chunk; a chunk never spans two islands.
template <typename T, int NR, bool ZW>
__global__ __launch_bounds__(kBlock) void k_isl_sums(const T* __restrict__ v,
                                                     const double* __restrict__ dhalf,
                                                     const int32_t* __restrict__ nodes,
                                                     const int32_t* __restrict__ cstart,
                                                     int nchunks, VecArgs va, int mode,
                                                     const CgScalars* sc,
                                                     double* __restrict__ part) {
  const int c = blockIdx.x;
  const int lo = cstart[c], hi = cstart[c + 1];
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  bool live[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) live[r] = r < va.nr && sc[r].done == 0;
  for (int t = lo + (int)threadIdx.x; t < hi; t += kBlock) {
    const int64_t n = nodes[t];
    const double w = ZW ? dhalf[n] : 1.0;
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (live[r]) acc[r] += w * (double)v[r * va.vs + n];
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr) part[(int64_t)r * nchunks + c] = acc[r];
  }
}

// s = Z^T r: one wave per island (4 per block) sums its chunk partials in a
// fixed order (lanes stride the chunks, then a wave reduction).
// grid (ceil(ni / 4), nr).
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__global__ __launch_bounds__(kBlock) void k_isl_reduce(const double* __restrict__ part,
                                                       int nchunks,
                                                       const int32_t* __restrict__ ichunk,
                                                       int ni, double* __restrict__ sv,
                                                       int mode, int first,
                                                       const CgScalars* sc) {
  const int r = blockIdx.y;
  if (first ? sc[r].done != 0 : !rhs_live(sc, r, mode)) return;  // uniform over the block
  const int lane = threadIdx.x % kWave;
  const int isl = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (isl >= ni) return;
  double a = 0.0;
  for (int c = ichunk[isl] + lane; c < ichunk[isl + 1]; c += kWave)
    a += part[(int64_t)r * nchunks + c];
  a = wave_sum_f64(a);
  if (lane == 0) sv[(int64_t)r * kMaxIslands + isl] = a;
}

// c = E^-1 s, one wave per row (coalesced row reads), and the products c_I s_I.
// grid (ceil(ni / 4), nr).
__global__ __launch_bounds__(kBlock) void k_defl_matvec(const double* __restrict__ einv, int ni,
                                                        const double* __restrict__ sv,
                                                        double* __restrict__ coef,
                                                        double* __restrict__ cs, int mode,
                                                        int first, const CgScalars* sc) {
  const int r = blockIdx.y;
  if (first ? sc[r].done != 0 : !rhs_live(sc, r, mode)) return;  // uniform over the block
  const int lane = threadIdx.x % kWave;
  const int row = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (row >= ni) return;
  const double* e = einv + (int64_t)row * ni;
  const double* s = sv + (int64_t)r * kMaxIslands;
  double a = 0.0;
  for (int j = lane; j < ni; j += kWave) a += e[j] * s[j];
  a = wave_sum_f64(a);
  if (lane == 0) {
    coef[(int64_t)r * kMaxIslands + row] = a;
    cs[(int64_t)r * kMaxIslands + row] = a * s[row];
  }
}

// One block per RHS: zr = <M r, r> = rho + sum_I c_I s_I (fixed order).
// first: start of a (FP64 or FP32-cycle) iteration sequence, beta stays 0;
// otherwise beta = zr_new / zr_old (the stop tests of k_finalize already ran
// on rho).
__global__ __launch_bounds__(kBlock) void k_defl_finish(const double* __restrict__ cs, int ni,
                                                        int mode, int first, CgScalars* sc) {
  const int r = blockIdx.x;
  if (first ? sc[r].done != 0 : !rhs_live(sc, r, mode)) return;  // uniform over the block
  double acc[1] = {0.0};
  for (int i = threadIdx.x; i < ni; i += kBlock) acc[0] += cs[(int64_t)r * kMaxIslands + i];
  block_sum<1>(acc);
  if (threadIdx.x != 0) return;
  CgScalars& c = sc[r];
  const double rho = mode == 0 ? c.rho : c.aux0;
  const double zr = rho + acc[0];
  if (first) {
    c.beta = 0.0;
  } else {
    c.beta = zr / c.zr;
  }
  c.zr = zr;
  if (!(zr > 0.0) || !isfinite(zr) || !isfinite(c.beta)) c.done = 3;
}

// v += Z c on the island nodes (after p = M_jacobi r + beta p).
template <typename T, int NR, bool ZW>
__global__ __launch_bounds__(kBlock) void k_isl_scatter(T* __restrict__ v,
                                                        const double* __restrict__ dhalf,
                                                        const int32_t* __restrict__ nodes,
                                                        const int32_t* __restrict__ cstart,
                                                        const int32_t* __restrict__ cisl,
                                                        const double* __restrict__ coef,
                                                        VecArgs va, int mode,
                                                        const CgScalars* sc) {
  const int c = blockIdx.x;
  const int lo = cstart[c], hi = cstart[c + 1], isl = cisl[c];
  double cr[NR];
  bool live[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    live[r] = r < va.nr && rhs_live(sc, r, mode);
    cr[r] = live[r] ? coef[(int64_t)r * kMaxIslands + isl] : 0.0;
  }
  for (int t = lo + (int)threadIdx.x; t < hi; t += kBlock) {
    const int64_t n = nodes[t];
    const double w = ZW ? dhalf[n] : 1.0;
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (live[r]) v[r * va.vs + n] += (T)(cr[r] * w);
  }
}

// q = L_u p in difference form (FP32), fused with the partials of p.q.
// Rows of Dirichlet and pinned nodes produce values nobody reads: p is 0
// there, and every later use of r is weighted by u_dinv (0 there).
template <int NR>
__global__ __launch_bounds__(kBlock) void k_stencil_diff(const float* __restrict__ p,
                                                         float* __restrict__ q,
                                                         const float* __restrict__ wx,
                                                         const float* __restrict__ wy,
                                                         const float* __restrict__ wz,
                                                         const float* __restrict__ ex, Dims d,
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
    const int i = tx * kTileX + (int)(threadIdx.x % kTileX);
    const int j = ty * kTileY + (int)(threadIdx.x / kTileX);
    const int k0 = tz * tl.kz;
    const int k1 = min(d.nz, k0 + tl.kz);
    if (i < d.nx && j < d.ny) {
      const int64_t base = (int64_t)j * d.nx + i;
      const bool has_e = i + 1 < d.nx, has_w = i > 0, has_n = j + 1 < d.ny, has_s = j > 0;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (!live[r]) continue;
        const float* __restrict__ pr = p + r * va.vs;
        float* __restrict__ qr = q + r * va.vs;
        int64_t nn = (int64_t)k0 * d.sxy + base;
        float wzm = k0 > 0 ? wz[nn - d.sxy] : 0.f;
        float pm = k0 > 0 ? pr[nn - d.sxy] : 0.f;
        float pc = pr[nn];
        double a = 0.0;
        for (int k = k0; k < k1; ++k, nn += d.sxy) {
          const float wxc = wx[nn], wyc = wy[nn], wzc = wz[nn];
          const float wxm = has_w ? wx[nn - 1] : 0.f;
          const float wym = has_s ? wy[nn - d.nx] : 0.f;
          const float pu = k + 1 < d.nz ? pr[nn + d.sxy] : 0.f;
          const float pe = has_e ? pr[nn + 1] : 0.f;
          const float pw = has_w ? pr[nn - 1] : 0.f;
          const float pn = has_n ? pr[nn + d.nx] : 0.f;
          const float ps = has_s ? pr[nn - d.nx] : 0.f;
          // edges past the boundary carry weight 0 (node-padded arrays)
          float y = wxc * (pc - pe) + wxm * (pc - pw) + wyc * (pc - pn) + wym * (pc - ps) +
                    wzc * (pc - pu) + wzm * (pc - pm);
          if (ex) y += ex[nn] * pc;
          qr[nn] = y;
          a += (double)pc * (double)y;
          pm = pc;
          pc = pu;
          wzm = wzc;
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

// k_stencil_diff with 4 consecutive x nodes per lane (rows a multiple of 4):
// one 16-byte load per array and plane, every right-hand side advanced in the
// same z-march, x neighbours at a lane's ends from the adjacent lanes
// (shuffles) or from memory at the edges of the wave; grid of 256-node row
// segments.
typedef float f4d __attribute__((ext_vector_type(4)));
template <int NR>
__global__ __launch_bounds__(kBlock) void k_stencil_diff4(const float* __restrict__ p,
                                                          float* __restrict__ q,
                                                          const float* __restrict__ wx,
                                                          const float* __restrict__ wy,
                                                          const float* __restrict__ wz,
                                                          const float* __restrict__ ex, Dims d,
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
    const int i = (tx * kWave + lane) * 4;
    const int j = ty * kTileY + (int)(threadIdx.x / kWave);
    const int k0 = tz * tl.kz;
    const int k1 = min(d.nz, k0 + tl.kz);
    const bool valid = i < d.nx && j < d.ny;  // i .. i+3 inside: nx % 4 == 0
    const int64_t base = valid ? (int64_t)j * d.nx + i : 0;
    const bool has_w = i > 0, has_e = i + 4 < d.nx, has_n = j + 1 < d.ny, has_s = j > 0;
    int64_t nn = (int64_t)k0 * d.sxy + base;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 pm[NR], pc[NR];
    float4 wzm = z4;
    if (valid && k0 > 0) wzm = *reinterpret_cast<const float4*>(wz + nn - d.sxy);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      pm[r] = z4;
      pc[r] = z4;
      if (valid && live[r]) {
        if (k0 > 0) pm[r] = *reinterpret_cast<const float4*>(p + r * va.vs + nn - d.sxy);
        pc[r] = *reinterpret_cast<const float4*>(p + r * va.vs + nn);
      }
    }
    for (int k = k0; k < k1; ++k, nn += d.sxy) {
      float4 wxc = z4, wyc = z4, wzc = z4, wym = z4, e4 = z4;
      if (valid) {
        wxc = *reinterpret_cast<const float4*>(wx + nn);
        wyc = *reinterpret_cast<const float4*>(wy + nn);
        wzc = *reinterpret_cast<const float4*>(wz + nn);
        if (has_s) wym = *reinterpret_cast<const float4*>(wy + nn - d.nx);
        if (ex) e4 = *reinterpret_cast<const float4*>(ex + nn);
      }
      // weight of the edge west of this lane's first node
      const float wxw_l = __shfl_up(wxc.w, 1, kWave);
      float wxw = 0.f;
      if (valid && has_w) wxw = lane > 0 ? wxw_l : wx[nn - 1];
      const bool up = k + 1 < d.nz;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (!live[r]) continue;  // uniform over the launch
        const float* __restrict__ pr = p + r * va.vs;
        float4 pu = z4, pn = z4, ps = z4;
        if (valid) {
          if (up) pu = *reinterpret_cast<const float4*>(pr + nn + d.sxy);
          if (has_n) pn = *reinterpret_cast<const float4*>(pr + nn + d.nx);
          if (has_s) ps = *reinterpret_cast<const float4*>(pr + nn - d.nx);
        }
        const float4 c = pc[r];
        float pw = __shfl_up(c.w, 1, kWave), pe = __shfl_down(c.x, 1, kWave);
        if (valid) {
          if (!has_w) pw = 0.f;
          else if (lane == 0) pw = pr[nn - 1];
          if (!has_e) pe = 0.f;
          else if (lane == kWave - 1) pe = pr[nn + 4];
        }
        const float4 m = pm[r];
        // edges past the boundary carry weight 0 (node-padded arrays)
        f4d y;
        y.x = e4.x * c.x + wxc.x * (c.x - c.y) + wxw * (c.x - pw) + wyc.x * (c.x - pn.x) +
              wym.x * (c.x - ps.x) + wzc.x * (c.x - pu.x) + wzm.x * (c.x - m.x);
        y.y = e4.y * c.y + wxc.y * (c.y - c.z) + wxc.x * (c.y - c.x) + wyc.y * (c.y - pn.y) +
              wym.y * (c.y - ps.y) + wzc.y * (c.y - pu.y) + wzm.y * (c.y - m.y);
        y.z = e4.z * c.z + wxc.z * (c.z - c.w) + wxc.y * (c.z - c.y) + wyc.z * (c.z - pn.z) +
              wym.z * (c.z - ps.z) + wzc.z * (c.z - pu.z) + wzm.z * (c.z - m.z);
        y.w = e4.w * c.w + wxc.w * (c.w - pe) + wxc.z * (c.w - c.z) + wyc.w * (c.w - pn.w) +
              wym.w * (c.w - ps.w) + wzc.w * (c.w - pu.w) + wzm.w * (c.w - m.w);
        if (valid) {
          __builtin_nontemporal_store(y, reinterpret_cast<f4d*>(q + r * va.vs + nn));
          acc[r] += (double)(c.x * y.x + c.y * y.y + c.z * y.z + c.w * y.w);
        }
        pm[r] = c;
        pc[r] = pu;
      }
      wzm = wzc;
    }
  }
  block_sum<NR>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < va.nr) store_partial(partials + (int64_t)r * gridDim.x + blockIdx.x, acc[r]);
  }
}

// r -= alpha q; partial sum wdot r^2 (the equilibrated norm of an unscaled
// residual, wdot = 1/D).  x += alpha p is deferred to the next k_pupdate_w
// (which reads p anyway); the refresh adds the pending step of a cycle.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_update_w(float* __restrict__ r_,
                                                     const float* __restrict__ q,
                                                     const float* __restrict__ wdot, VecArgs va,
                                                     int mode, CgScalars* sc, double* partials) {
  if (all_idle(sc, va.nr, mode)) return;
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!(r < va.nr && rhs_live(sc, r, mode))) continue;
    const float a = (float)sc[r].alpha;
    float* rr = r_ + r * va.vs;
    const float* qr = q + r * va.vs;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride) {
      const float rn = rr[i] - a * qr[i];
      rr[i] = rn;
      s += (double)wdot[i] * (double)rn * (double)rn;
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

// x += alpha p (the previous step's, deferred; alpha is 0 at a cycle start),
// then p = wdir r + beta p
template <int NR>
__global__ __launch_bounds__(kBlock) void k_pupdate_w(float* __restrict__ p,
                                                      float* __restrict__ x,
                                                      const float* __restrict__ r_,
                                                      const float* __restrict__ wdir,
                                                      VecArgs va, int mode,
                                                      const CgScalars* sc) {
  if (all_idle(sc, va.nr, mode)) return;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!(r < va.nr && rhs_live(sc, r, mode))) continue;
    const float b = (float)sc[r].beta, a = (float)sc[r].alpha;
    float* pr = p + r * va.vs;
    float* xr = x + r * va.vs;
    const float* rr = r_ + r * va.vs;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride) {
      const float pv = pr[i];
      xr[i] += a * pv;
      pr[i] = wdir[i] * rr[i] + b * pv;
    }
  }
}

// FP32 cycle start in unscaled variables: r32 = dhalf * r64 / ||r64||,
// d32 = p32 = 0; partials of ||r64/||r64||||^2 (the cycle's rho).
template <int NR>
__global__ __launch_bounds__(kBlock) void k_cycle_start_u(const double* __restrict__ r64,
                                                          const double* __restrict__ dhalf,
                                                          float* __restrict__ r32,
                                                          float* __restrict__ d32,
                                                          float* __restrict__ p32, VecArgs va,
                                                          const CgScalars* sc,
                                                          double* partials) {
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
      const double v = r64[r * va.vs + i] * sc_r;
      r32[r * va.vs + i] = (float)(v * dhalf[i]);
      d32[r * va.vs + i] = 0.0f;
      p32[r * va.vs + i] = 0.0f;
      s += v * v;
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

// x64 += dhalf * (d32 + alpha p32) / scale (the cycle solved for unscaled
// potentials; alpha p32 is its last step, whose x update was deferred)
template <int NR>
__global__ __launch_bounds__(kBlock) void k_cycle_accumulate_u(double* __restrict__ x64,
                                                               const float* __restrict__ d32,
                                                               const float* __restrict__ p32,
                                                               const double* __restrict__ dhalf,
                                                               VecArgs va, const CgScalars* sc) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!(r < va.nr && sc[r].done == 0)) continue;
    const double inv = 1.0 / sc[r].scale;
    const float a = (float)sc[r].alpha;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < va.n; i += stride)
      x64[r * va.vs + i] +=
          dhalf[i] * (double)(d32[r * va.vs + i] + a * p32[r * va.vs + i]) * inv;
  }
}

__global__ __launch_bounds__(kBlock) void k_to_f32(const double* __restrict__ in,
                                                   float* __restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    out[i] = (float)in[i];
}

// 1/D = 1/dhalf^2 on free rows, 0 elsewhere
__global__ __launch_bounds__(kBlock) void k_udinv(const double* __restrict__ dhalf,
                                                  float* __restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double h = dhalf[i];
    out[i] = h > 0.0 ? (float)(1.0 / (h * h)) : 0.0f;
  }
}

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------
template <int NR>
static void defl_coef_t(fs_sys* s, int nr, int mode, int first) {
  VecArgs va{s->d.n, s->vstride, nr};
  const int nc = s->defl_nchunks;
  if (mode == 0)
    hipLaunchKernelGGL((k_isl_sums<double, NR, true>), dim3(nc), dim3(kBlock), 0, s->stream,
                       s->r, s->dhalf, s->defl_nodes, s->defl_cstart, nc, va, mode, s->sc,
                       s->defl_part);
  else
    hipLaunchKernelGGL((k_isl_sums<float, NR, false>), dim3(nc), dim3(kBlock), 0, s->stream,
                       s->r32, s->dhalf, s->defl_nodes, s->defl_cstart, nc, va, mode, s->sc,
                       s->defl_part);
  const int ni = s->defl_n;
  const dim3 g2((ni + kBlock / kWave - 1) / (kBlock / kWave), nr);
  double* sv = s->defl_coef + (size_t)kMaxRhs * kMaxIslands;      // s = Z^T r
  double* cs = s->defl_coef + (size_t)2 * kMaxRhs * kMaxIslands;  // c_I s_I
  hipLaunchKernelGGL(k_isl_reduce, g2, dim3(kBlock), 0, s->stream, s->defl_part, nc,
                     s->defl_ichunk, ni, sv, mode, first, s->sc);
  hipLaunchKernelGGL(k_defl_matvec, g2, dim3(kBlock), 0, s->stream, s->defl_einv, ni, sv,
                     s->defl_coef, cs, mode, first, s->sc);
  hipLaunchKernelGGL(k_defl_finish, dim3(nr), dim3(kBlock), 0, s->stream, cs, ni, mode, first,
                     s->sc);
}

template <int NR>
static void defl_scatter_t(fs_sys* s, int nr, int mode) {
  VecArgs va{s->d.n, s->vstride, nr};
  const int nc = s->defl_nchunks;
  if (mode == 0)
    hipLaunchKernelGGL((k_isl_scatter<double, NR, true>), dim3(nc), dim3(kBlock), 0, s->stream,
                       s->p, s->dhalf, s->defl_nodes, s->defl_cstart, s->defl_cisl,
                       s->defl_coef, va, mode, s->sc);
  else
    hipLaunchKernelGGL((k_isl_scatter<float, NR, false>), dim3(nc), dim3(kBlock), 0, s->stream,
                       s->p32, s->dhalf, s->defl_nodes, s->defl_cstart, s->defl_cisl,
                       s->defl_coef, va, mode, s->sc);
}

template <int NR>
static int iteration_defl_t(fs_sys* s, int nr, int mode, ProfSlot* sl) {
  hipStream_t st = s->stream;
  VecArgs va{s->d.n, s->vstride, nr};
  const int vb = s->vec_blocks, sb = s->tl.nblocks;
  hipEvent_t ea = prof_start(sl, FS_PROF_STENCIL), eb = prof_stop(sl, FS_PROF_STENCIL);
  unsigned long long* prof = prof_ctr(s, sl, FS_PROF_STENCIL);
  // FP32 cycle on unscaled potentials: p = (D^-1 + 1_I E^-1 1_I^T) r + beta p
  hipLaunchKernelGGL((k_pupdate_w<NR>), dim3(vb), dim3(kBlock), 0, st, s->p32, s->x32, s->r32,
                     s->u_dinv, va, mode, s->sc);
  defl_scatter_t<NR>(s, nr, mode);
  int sbf = sb;
  if (s->use_x4) {
    sbf = s->tl4.nblocks;
    launch_timed(k_stencil_diff4<NR>, dim3(sbf), st, ea, eb, (const float*)s->p32, s->q32,
                 (const float*)s->u_w[0], (const float*)s->u_w[1], (const float*)s->u_w[2],
                 (const float*)s->u_extra, s->d, s->tl4, va, mode, s->sc, s->partials, prof);
  } else {
    launch_timed(k_stencil_diff<NR>, dim3(sb), st, ea, eb, (const float*)s->p32, s->q32,
                 (const float*)s->u_w[0], (const float*)s->u_w[1], (const float*)s->u_w[2],
                 (const float*)s->u_extra, s->d, s->tl, va, mode, s->sc, s->partials, prof);
  }
  launch_finalize(st, s->partials, FIN_ALPHA_ZR, nr, sbf, mode, s->sc);
  hipLaunchKernelGGL((k_update_w<NR>), dim3(vb), dim3(kBlock), 0, st, s->r32, s->q32,
                     s->u_dinv, va, mode, s->sc, s->partials);
  launch_finalize(st, s->partials, FIN_BETA, nr, vb, mode, s->sc);
  defl_coef_t<NR>(s, nr, mode, 0);
  return 0;
}

#define FS_DEFL_NR(nr, FN, ...)              \
  do {                                       \
    if ((nr) <= 1) FN<1>(__VA_ARGS__);       \
    else if ((nr) <= 2) FN<2>(__VA_ARGS__);  \
    else if ((nr) <= 4) FN<4>(__VA_ARGS__);  \
    else FN<8>(__VA_ARGS__);                 \
  } while (0)

int launch_iteration_defl(fs_sys* s, int nr, int mode, ProfSlot* slot) {
  if (nr <= 1) return iteration_defl_t<1>(s, nr, mode, slot);
  if (nr <= 2) return iteration_defl_t<2>(s, nr, mode, slot);
  if (nr <= 4) return iteration_defl_t<4>(s, nr, mode, slot);
  return iteration_defl_t<8>(s, nr, mode, slot);
}

void launch_defl_coef(fs_sys* s, int nr, int mode, int first) {
  FS_DEFL_NR(nr, defl_coef_t, s, nr, mode, first);
}
void launch_defl_scatter(fs_sys* s, int nr, int mode) {
  FS_DEFL_NR(nr, defl_scatter_t, s, nr, mode);
}

template <int NR>
static void cycle_start_defl_t(fs_sys* s, int nr, double inner_tol2) {
  VecArgs va{s->d.n, s->vstride, nr};
  hipLaunchKernelGGL((k_cycle_start_u<NR>), dim3(s->vec_blocks), dim3(kBlock), 0, s->stream,
                     s->r, s->dhalf, s->r32, s->x32, s->p32, va, s->sc, s->partials);
  launch_finalize(s->stream, s->partials, FIN_CYCLE, nr, s->vec_blocks, 1, s->sc, inner_tol2);
  defl_coef_t<NR>(s, nr, 1, 1);
}
void launch_cycle_start_defl(fs_sys* s, int nr, double inner_tol2) {
  FS_DEFL_NR(nr, cycle_start_defl_t, s, nr, inner_tol2);
}

template <int NR>
static void refresh_defl_t(fs_sys* s, int nr) {
  VecArgs va{s->d.n, s->vstride, nr};
  hipLaunchKernelGGL((k_cycle_accumulate_u<NR>), dim3(s->vec_blocks), dim3(kBlock), 0,
                     s->stream, s->x, s->x32, s->p32, s->dhalf, va, s->sc);
  launch_residual(s, nr, s->r);
  launch_refresh_finish(s, nr);
}
void launch_refresh_defl(fs_sys* s, int nr) { FS_DEFL_NR(nr, refresh_defl_t, s, nr); }

void free_defl(fs_sys* s) {
  dfree(s->defl_nodes);
  dfree(s->defl_cstart);
  dfree(s->defl_cisl);
  dfree(s->defl_ichunk);
  dfree(s->defl_einv);
  dfree(s->defl_part);
  dfree(s->defl_coef);
  for (auto& w : s->u_w) dfree(w);
  dfree(s->u_extra);
  dfree(s->u_dinv);
  s->defl_n = 0;
  s->defl_nchunks = 0;
}

// --------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------
namespace {

// x-fast index -> C-order (z fastest) index
inline int64_t c_index(const Dims& d, int64_t n) {
  const int64_t i = n % d.nx, j = (n / d.nx) % d.ny, k = n / d.sxy;
  return (i * d.ny + j) * d.nz + k;
}

int download(hipStream_t st, const void* dev, void* host, size_t bytes) {
  FS_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
  FS_HIP(hipStreamSynchronize(st));
  return FS_OK;
}

// edge weights of the system's operator and its row mask, x-fast
int host_operator(fs_sys* s, std::vector<double> (&w)[3], std::vector<double>& extra,
                  std::vector<uint8_t>& mask) {
  const int64_t n = s->d.n;
  hipStream_t st = s->ctx->stream;
  for (int a = 0; a < 3; ++a) {
    w[a].resize((size_t)n);
    if (int rc = download(st, s->op->w[a], w[a].data(), (size_t)n * sizeof(double))) return rc;
  }
  extra.clear();
  if (s->op->extra) {
    extra.resize((size_t)n);
    if (int rc = download(st, s->op->extra, extra.data(), (size_t)n * sizeof(double))) return rc;
  }
  mask.resize((size_t)n);
  return download(st, s->mask, mask.data(), (size_t)n);
}

struct UnionFind {
  std::vector<int32_t> parent;
  explicit UnionFind(int64_t n) : parent((size_t)n) {
    std::iota(parent.begin(), parent.end(), 0);
  }
  int32_t find(int32_t a) {
    while (parent[a] != a) {
      parent[a] = parent[parent[a]];
      a = parent[a];
    }
    return a;
  }
  void unite(int32_t a, int32_t b) {
    a = find(a);
    b = find(b);
    if (a == b) return;
    if (a < b) std::swap(a, b);
    parent[a] = b;  // root = smallest index
  }
};

}  // namespace

}  // namespace fs

using namespace fs;

int fs_sys_find_islands(fs_sys* sys, double contrast, int max_islands, int32_t* island,
                        int* n_islands) {
  FS_GUARD_BEGIN
  if (!sys || !island || !n_islands) return fail(FS_EINVAL, "NULL argument");
  if (!sys->op) return fail(FS_EINVAL, "the system's operator is not available");
  if (sys->d.per) return fail(FS_EINVAL, "periodic operators run fs_sys_solve and fs_capacitance only (FS_PREC_FP64 or FS_PREC_MIXED)");
  if (max_islands < 1 || max_islands > kMaxIslands)
    return fail(FS_EINVAL, "max_islands must lie in [1, %d], got %d", kMaxIslands, max_islands);
  if (!(contrast >= 1.0)) return fail(FS_EINVAL, "contrast must be >= 1, got %g", contrast);
  *n_islands = 0;
  FS_HIP(hipSetDevice(sys->ctx->device));
  const Dims& d = sys->d;
  const int64_t n = d.n;
  std::vector<double> w[3], extra;
  std::vector<uint8_t> mask;
  if (int rc = host_operator(sys, w, extra, mask)) return rc;
  double wmin = INFINITY, wmax = 0.0;
  for (int a = 0; a < 3; ++a)
    for (double v : w[a])
      if (v > 0.0) {
        wmin = std::min(wmin, v);
        wmax = std::max(wmax, v);
      }
  for (int64_t t = 0; t < n; ++t) island[t] = -1;
  if (!(wmax > 0.0) || wmax < contrast * wmin) return FS_OK;
  const double thr = std::sqrt(wmin * wmax);
  UnionFind uf(n);
  const int64_t step[3] = {1, d.nx, d.sxy};
  for (int64_t m = 0; m < n; ++m) {
    const int64_t i = m % d.nx, j = (m / d.nx) % d.ny, k = m / d.sxy;
    const bool inside[3] = {i + 1 < d.nx, j + 1 < d.ny, k + 1 < d.nz};
    for (int a = 0; a < 3; ++a)
      if (inside[a] && w[a][m] > thr) uf.unite((int32_t)m, (int32_t)(m + step[a]));
  }
  // components with >= 2 nodes and no Dirichlet or pinned row
  std::vector<int64_t> size((size_t)n, 0);
  std::vector<uint8_t> anchored((size_t)n, 0);
  for (int64_t m = 0; m < n; ++m) {
    const int32_t rt = uf.find((int32_t)m);
    size[rt] += 1;
    if (mask[m] != 0) anchored[rt] = 1;
  }
  std::vector<int32_t> roots;
  for (int64_t m = 0; m < n; ++m)
    if (uf.parent[m] == m && size[m] >= 2 && !anchored[m]) roots.push_back((int32_t)m);
  if ((int)roots.size() > max_islands) {
    // keep the largest islands (ties: lowest node), then restore node order
    std::stable_sort(roots.begin(), roots.end(),
                     [&](int32_t a, int32_t b) { return size[a] > size[b]; });
    roots.resize((size_t)max_islands);
    std::sort(roots.begin(), roots.end());
  }
  std::vector<int32_t> id_of_root((size_t)n, -1);
  for (size_t t = 0; t < roots.size(); ++t) id_of_root[roots[t]] = (int32_t)t;
  for (int64_t m = 0; m < n; ++m) island[c_index(d, m)] = id_of_root[uf.find((int32_t)m)];
  *n_islands = (int)roots.size();
  return FS_OK;
  FS_GUARD_END
}

int fs_sys_set_deflation(fs_sys* sys, int n_islands, const int32_t* island) {
  FS_GUARD_BEGIN
  if (!sys) return fail(FS_EINVAL, "NULL argument");
  if (sys->d.per && n_islands > 0) return fail(FS_EINVAL, "periodic operators run fs_sys_solve and fs_capacitance only (FS_PREC_FP64 or FS_PREC_MIXED)");
  FS_HIP(hipSetDevice(sys->ctx->device));
  if (n_islands == 0 || !island) {
    FS_HIP(hipStreamSynchronize(sys->ctx->stream));
    free_defl(sys);
    return FS_OK;
  }
  if (!sys->op) return fail(FS_EINVAL, "the system's operator is not available");
  if (n_islands < 0 || n_islands > kMaxIslands)
    return fail(FS_EINVAL, "n_islands must lie in [0, %d], got %d", kMaxIslands, n_islands);
  const Dims& d = sys->d;
  const int64_t n = d.n;
  const int ni = n_islands;
  std::vector<double> w[3], extra;
  std::vector<uint8_t> mask;
  if (int rc = host_operator(sys, w, extra, mask)) return rc;
  // island of each x-fast node
  std::vector<int32_t> isl((size_t)n);
  std::vector<int64_t> count((size_t)ni, 0);
  for (int64_t m = 0; m < n; ++m) {
    const int32_t v = island[c_index(d, m)];
    if (v < -1 || v >= ni)
      return fail(FS_EINVAL, "island id %d at node %lld outside [-1, %d)", v,
                  (long long)c_index(d, m), ni);
    if (v >= 0 && mask[m] != 0)
      return fail(FS_EINVAL, "node %lld of island %d is a Dirichlet or pinned row",
                  (long long)c_index(d, m), v);
    isl[m] = v;
    if (v >= 0) count[v] += 1;
  }
  for (int I = 0; I < ni; ++I)
    if (count[I] == 0) return fail(FS_EINVAL, "island %d has no node", I);
  // E = 1_I^T L_ff 1_J: every edge leaving island I adds to E_II; an edge
  // between islands I and J adds -w to E_IJ and E_JI
  std::vector<double> E((size_t)ni * ni, 0.0);
  const int64_t step[3] = {1, d.nx, d.sxy};
  for (int64_t m = 0; m < n; ++m) {
    const int64_t i = m % d.nx, j = (m / d.nx) % d.ny, k = m / d.sxy;
    const bool inside[3] = {i + 1 < d.nx, j + 1 < d.ny, k + 1 < d.nz};
    const int32_t I = isl[m];
    if (I >= 0 && !extra.empty()) E[(size_t)I * ni + I] += extra[m];
    for (int a = 0; a < 3; ++a) {
      if (!inside[a]) continue;
      const double we = w[a][m];
      if (we == 0.0) continue;
      const int32_t J = isl[m + step[a]];
      if (I == J) continue;
      if (I >= 0) E[(size_t)I * ni + I] += we;
      if (J >= 0) E[(size_t)J * ni + J] += we;
      if (I >= 0 && J >= 0) {
        E[(size_t)I * ni + J] -= we;
        E[(size_t)J * ni + I] -= we;
      }
    }
  }
  // E^-1 through a Cholesky factorisation (E is SPD when every island is
  // coupled to a Dirichlet row through the operator)
  std::vector<double> Lc(E);
  for (int c = 0; c < ni; ++c) {
    double diag = Lc[(size_t)c * ni + c];
    for (int t = 0; t < c; ++t) diag -= Lc[(size_t)c * ni + t] * Lc[(size_t)c * ni + t];
    if (!(diag > 1e-14 * E[(size_t)c * ni + c]) || !std::isfinite(diag))
      return fail(FS_EINVAL,
                  "island coarse matrix is not positive definite at island %d (an island "
                  "not coupled to any Dirichlet row)", c);
    const double lcc = std::sqrt(diag);
    Lc[(size_t)c * ni + c] = lcc;
    for (int rr = c + 1; rr < ni; ++rr) {
      double v = Lc[(size_t)rr * ni + c];
      for (int t = 0; t < c; ++t) v -= Lc[(size_t)rr * ni + t] * Lc[(size_t)c * ni + t];
      Lc[(size_t)rr * ni + c] = v / lcc;
    }
  }
  std::vector<double> Einv((size_t)ni * ni, 0.0);
  std::vector<double> col((size_t)ni);
  for (int c = 0; c < ni; ++c) {
    std::fill(col.begin(), col.end(), 0.0);
    col[c] = 1.0;
    for (int rr = 0; rr < ni; ++rr) {  // L y = e_c
      double v = col[rr];
      for (int t = 0; t < rr; ++t) v -= Lc[(size_t)rr * ni + t] * col[t];
      col[rr] = v / Lc[(size_t)rr * ni + rr];
    }
    for (int rr = ni - 1; rr >= 0; --rr) {  // L^T x = y
      double v = col[rr];
      for (int t = rr + 1; t < ni; ++t) v -= Lc[(size_t)t * ni + rr] * col[t];
      col[rr] = v / Lc[(size_t)rr * ni + rr];
    }
    for (int rr = 0; rr < ni; ++rr) Einv[(size_t)rr * ni + c] = col[rr];
  }
  // node list grouped by island, cut into chunks of at most kIslChunk nodes
  std::vector<int64_t> off((size_t)ni + 1, 0);
  for (int I = 0; I < ni; ++I) off[I + 1] = off[I] + count[I];
  const int64_t total = off[ni];
  if (total > ((int64_t)1 << 31) - 1) return fail(FS_EINVAL, "too many island nodes");
  std::vector<int32_t> nodes((size_t)total);
  {
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    for (int64_t m = 0; m < n; ++m)
      if (isl[m] >= 0) nodes[pos[isl[m]]++] = (int32_t)m;
  }
  std::vector<int32_t> cstart, cisl, ichunk((size_t)ni + 1, 0);
  for (int I = 0; I < ni; ++I) {
    ichunk[I] = (int32_t)cisl.size();
    for (int64_t t = off[I]; t < off[I + 1]; t += kIslChunk) {
      cstart.push_back((int32_t)t);
      cisl.push_back(I);
    }
  }
  ichunk[ni] = (int32_t)cisl.size();
  cstart.push_back((int32_t)total);
  const int nchunks = (int)cisl.size();

  FS_HIP(hipStreamSynchronize(sys->ctx->stream));
  free_defl(sys);
  hipStream_t st = sys->ctx->stream;
  int rc = FS_OK;
  auto up = [&](auto* dev, const auto& host) {
    if (rc != FS_OK) return;
    if (hipMemcpyAsync(dev, host.data(), host.size() * sizeof(host[0]), hipMemcpyHostToDevice,
                       st) != hipSuccess)
      rc = fail(FS_EHIP, "island space upload failed");
  };
  if ((rc = dalloc(&sys->defl_nodes, (size_t)total)) ||
      (rc = dalloc(&sys->defl_cstart, cstart.size())) ||
      (rc = dalloc(&sys->defl_cisl, cisl.size())) ||
      (rc = dalloc(&sys->defl_ichunk, ichunk.size())) ||
      (rc = dalloc(&sys->defl_einv, Einv.size())) ||
      (rc = dalloc(&sys->defl_part, (size_t)kMaxRhs * nchunks)) ||
      (rc = dalloc(&sys->defl_coef, (size_t)3 * kMaxRhs * kMaxIslands)) ||
      (rc = dalloc(&sys->u_w[0], (size_t)n)) || (rc = dalloc(&sys->u_w[1], (size_t)n)) ||
      (rc = dalloc(&sys->u_w[2], (size_t)n)) || (rc = dalloc(&sys->u_dinv, (size_t)n)) ||
      (sys->op->extra && (rc = dalloc(&sys->u_extra, (size_t)n)))) {
    free_defl(sys);
    return rc;
  }
  up(sys->defl_nodes, nodes);
  up(sys->defl_cstart, cstart);
  up(sys->defl_cisl, cisl);
  up(sys->defl_ichunk, ichunk);
  up(sys->defl_einv, Einv);
  if (rc != FS_OK) {
    free_defl(sys);
    return rc;
  }
  const int blocks = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, 4096);
  for (int a = 0; a < 3; ++a)
    hipLaunchKernelGGL(k_to_f32, dim3(blocks), dim3(kBlock), 0, st, sys->op->w[a], sys->u_w[a], n);
  if (sys->op->extra)
    hipLaunchKernelGGL(k_to_f32, dim3(blocks), dim3(kBlock), 0, st, sys->op->extra,
                       sys->u_extra, n);
  hipLaunchKernelGGL(k_udinv, dim3(blocks), dim3(kBlock), 0, st, sys->dhalf, sys->u_dinv, n);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    free_defl(sys);
    return fail(FS_EHIP, "island space set-up failed");
  }
  sys->defl_nchunks = nchunks;
  sys->defl_n = ni;
  return FS_OK;
  FS_GUARD_END
}

int fs_sys_get_deflation(fs_sys* sys, int* n_islands, double* einv) {
  FS_GUARD_BEGIN
  if (!sys || !n_islands) return fail(FS_EINVAL, "NULL argument");
  *n_islands = sys->defl_n;
  if (einv && sys->defl_n > 0) {
    FS_HIP(hipSetDevice(sys->ctx->device));
    FS_HIP(hipMemcpy(einv, sys->defl_einv, (size_t)sys->defl_n * sys->defl_n * sizeof(double),
                     hipMemcpyDeviceToHost));
  }
  return FS_OK;
  FS_GUARD_END
}
