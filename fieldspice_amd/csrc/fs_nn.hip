// Network kernels of the learned preconditioner: a three-level convolutional
// network (16/32/64 channels, 3x3x3 kernels, BF16 weights and activations,
// FP32 accumulation) evaluated and trained with hand-written MFMA
// implicit-GEMM convolutions (v_mfma_f32_16x16x32_bf16 on gfx950).
//
// Lane layout of the 16x16x32 BF16 MFMA (checked on the device):
//   lane l holds A[l%16][8*(l/16) + j], B[8*(l/16) + j][l%16], j = 0..7,
//   and D[4*(l/16) + j][l%16], j = 0..3.
//
// Convolution as implicit GEMM: rows = 16 output voxels consecutive in x,
// columns = 16 output channels, K = (tap, input channel) with the channel
// fastest, padded to a multiple of 32.  Activations are channels-last
// [batch][voxel][C] so the 8 consecutive K values a lane needs are one 16-byte
// load; weights are packed [cout][K] for the same reason.  The data-gradient
// of a convolution is the same kernel run with mirrored taps and transposed
// channel roles; the weight-gradient contracts over voxels on the MFMA.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fs_device.hpp"
#include "fs_internal.hpp"
#include "fs_nn.hpp"

namespace fs {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (inputs are finite)
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ void voxel_ijk(const Grid3& g, int64_t v, int& i, int& j, int& k) {
  if (g.n <= 0xffffffffLL) {  // 32-bit divisions (64-bit ones are long emulated sequences)
    const uint32_t u = (uint32_t)v, nx = (uint32_t)g.nx, ny = (uint32_t)g.ny;
    const uint32_t row = u / nx;
    i = (int)(u - row * nx);
    k = (int)(row / ny);
    j = (int)(row - (uint32_t)k * ny);
    return;
  }
  i = (int)(v % g.nx);
  j = (int)((v / g.nx) % g.ny);
  k = (int)(v / ((int64_t)g.nx * g.ny));
}

#define FS_GRID_STRIDE(t, cnt) \
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < (cnt); t += (int64_t)gridDim.x * 256)

// element t = (b n + v) C + c of a channels-last tensor of cnt elements ->
// (b, v, c), with 32-bit divisions when cnt fits
__device__ __forceinline__ void split_bvc(int64_t t, int C, int64_t n, int64_t cnt, int64_t& b,
                                          int64_t& v, int& c) {
  if (cnt <= 0xffffffffLL) {
    const uint32_t u = (uint32_t)t, q = u / (uint32_t)C, bb = q / (uint32_t)n;
    c = (int)(u - q * (uint32_t)C);
    v = (int64_t)(q - bb * (uint32_t)n);
    b = (int64_t)bb;
    return;
  }
  c = (int)(t % C);
  v = (t / C) % n;
  b = t / ((int64_t)C * n);
}

// ---------------------------------------------------------------------------
// forward convolution: Y = mask(act(conv(X, W) + b + skip))
// block = 4 waves over 16 voxels along x each; for short rows (nx <= 32) the
// waves of a block split the K loop (27*CIN) into ks slices over the same
// voxels instead (ks = 64 / width of the block's x range), and the partial
// tiles meet in LDS.  grid (ceil(nx/(64/ks)), ny, nz*nb)
// ---------------------------------------------------------------------------
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv_fwd(const uint16_t* __restrict__ X,
                                                  const uint16_t* __restrict__ Wt, int kpad,
                                                  const float* __restrict__ bias,
                                                  const uint16_t* __restrict__ skip,
                                                  const uint16_t* __restrict__ mask,
                                                  uint16_t* __restrict__ Y, Grid3 g, int relu,
    const int* __restrict__ gate, int ks) {
  if (gate && *gate) return;  // every right-hand side of the batch is idle (uniform)
  constexpr int NT = COUT / 16;
  __shared__ floatx4 red[4][NT][64];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int vt = wave / ks, slice = wave % ks;
  const int i0 = blockIdx.x * (64 / ks) + vt * 16;
  const int j = blockIdx.y;
  const int k = blockIdx.z % g.nz, b = blockIdx.z / g.nz;
  const bool active = i0 < g.nx;  // uniform per wave
  const int64_t xoff = (int64_t)b * g.n * CIN, yoff = (int64_t)b * g.n * COUT;
  const int m = lane % 16, kq = lane / 16;
  const int i = i0 + m;
  floatx4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int nit = kpad / 32, per = (nit + ks - 1) / ks;
  const int kb0 = slice * per * 32, kb1 = min(nit, (slice + 1) * per) * 32;
  if (active) {
    for (int kb = kb0; kb < kb1; kb += 32) {
      const int kk = kb + 8 * kq;
      const int tap = kk / CIN, ci0 = kk % CIN;
      uint4 araw = make_uint4(0u, 0u, 0u, 0u);
      if (tap < 27) {
        const int si = i + tap % 3 - 1, sj = j + (tap / 3) % 3 - 1, sk = k + tap / 9 - 1;
        if (i < g.nx && si >= 0 && si < g.nx && sj >= 0 && sj < g.ny && sk >= 0 && sk < g.nz) {
          const int64_t sv = ((int64_t)sk * g.ny + sj) * g.nx + si;
          araw = *reinterpret_cast<const uint4*>(X + xoff + sv * CIN + ci0);
        }
      }
      const bf16x8 a = __builtin_bit_cast(bf16x8, araw);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const uint4 braw = *reinterpret_cast<const uint4*>(Wt + (int64_t)(16 * t + m) * kpad + kk);
        const bf16x8 bw = __builtin_bit_cast(bf16x8, braw);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[t], 0, 0, 0);
      }
    }
  }
  if (ks > 1) {  // uniform over the block
#pragma unroll
    for (int t = 0; t < NT; ++t) red[wave][t][lane] = acc[t];
    __syncthreads();
    if (slice != 0) return;
    for (int w = 1; w < ks; ++w) {
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const floatx4 o = red[wave + w][t][lane];
        acc[t] = floatx4{acc[t][0] + o[0], acc[t][1] + o[1], acc[t][2] + o[2], acc[t][3] + o[3]};
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = 16 * t + m;
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int io = i0 + 4 * kq + r;
      if (io >= g.nx) continue;
      const int64_t o = yoff + (((int64_t)k * g.ny + j) * g.nx + io) * COUT + co;
      float y = acc[t][r] + bv;
      if (skip) y += bf2f(skip[o]);
      if (relu) y = fmaxf(y, 0.f);
      if (mask && !(bf2f(mask[o]) > 0.f)) y = 0.f;
      Y[o] = f2bf(y);
    }
  }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[tap][ci][co] = sum_v X[v + off(tap)][ci] * dY[v][co]
// (v over all batch elements).  One wave per (voxel range, tap, ci tile,
// co tile); the partial tiles are summed by k_wgrad_reduce in a fixed order.
// ---------------------------------------------------------------------------
template <int CIN, int COUT>
__global__ __launch_bounds__(64) void k_conv_wgrad(const uint16_t* __restrict__ X,
                                                   const uint16_t* __restrict__ dY, Grid3 g,
                                                   int64_t range_len, float* __restrict__ part) {
  constexpr int CT = CIN / 16, OT = COUT / 16;
  const int lane = threadIdx.x;
  const int range = blockIdx.x;
  const int tile = blockIdx.y;  // tap * CT * OT + cit * OT + cot
  const int tap = tile / (CT * OT);
  const int cit = (tile / OT) % CT, cot = tile % OT;
  const int di = tap % 3 - 1, dj = (tap / 3) % 3 - 1, dk = tap / 9 - 1;
  const int64_t ntot = g.n * g.nb;
  const int64_t vb0 = (int64_t)range * range_len;
  const int64_t vb1 = vb0 + range_len < ntot ? vb0 + range_len : ntot;
  const int m = lane % 16, kq = lane / 16;
  const int ci = cit * 16 + m, co = cot * 16 + m;
  floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int64_t vb = vb0; vb < vb1; vb += 32) {
    bf16x8 a, bv;
    // the lane's 8 consecutive voxels: coordinates of the first by division,
    // the others by stepping x with carries into y, z and the batch
    const int64_t v0 = vb + 8 * kq;
    int64_t bt = v0 / g.n;
    int i, j, k;
    voxel_ijk(g, v0 - bt * g.n, i, j, k);
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const int64_t v = v0 + jj;
      uint16_t xa = 0, yb = 0;
      if (v < vb1) {
        const int si = i + di, sj = j + dj, sk = k + dk;
        if (si >= 0 && si < g.nx && sj >= 0 && sj < g.ny && sk >= 0 && sk < g.nz)
          xa = X[(bt * g.n + ((int64_t)sk * g.ny + sj) * g.nx + si) * CIN + ci];
        yb = dY[v * COUT + co];
      }
      a[jj] = __builtin_bit_cast(__bf16, xa);
      bv[jj] = __builtin_bit_cast(__bf16, yb);
      if (++i == g.nx) {
        i = 0;
        if (++j == g.ny) {
          j = 0;
          if (++k == g.nz) {
            k = 0;
            ++bt;
          }
        }
      }
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc, 0, 0, 0);
  }
  // D[row = ci local][col = co local]: lane holds rows 4*kq + r of column m
  float* out = part + ((int64_t)range * gridDim.y + tile) * 256;
#pragma unroll
  for (int r = 0; r < 4; ++r) out[(4 * kq + r) * 16 + m] = acc[r];
}

__global__ __launch_bounds__(256) void k_wgrad_reduce(const float* __restrict__ part, int nranges,
                                                      int ntiles, int cin, int cout,
                                                      float* __restrict__ dW) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;  // tile*256 + e
  if (idx >= (int64_t)ntiles * 256) return;
  const int tile = (int)(idx / 256), e = (int)(idx % 256);
  const int CT = cin / 16, OT = cout / 16;
  const int tap = tile / (CT * OT), cit = (tile / OT) % CT, cot = tile % OT;
  float s = 0.f;
  for (int r = 0; r < nranges; ++r) s += part[((int64_t)r * ntiles + tile) * 256 + e];
  const int ci = cit * 16 + e / 16, co = cot * 16 + e % 16;
  dW[((int64_t)tap * cin + ci) * cout + co] = s;
}

// bias gradient: db[co] = sum_v dY[v][co] (one block per channel, fixed order)
__global__ __launch_bounds__(256) void k_bgrad(const uint16_t* __restrict__ dY, int64_t nvox,
                                               int cout, float* __restrict__ db) {
  const int co = blockIdx.x;
  double a[1] = {0.0};
  for (int64_t v = threadIdx.x; v < nvox; v += 256) a[0] += bf2f(dY[v * cout + co]);
  block_sum<1>(a);
  if (threadIdx.x == 0) db[co] = (float)a[0];
}

// ---------------------------------------------------------------------------
// level transfers inside the network (2:1 per axis, last cell may be single)
// pool: coarse = mean of the fine voxels 2I + {0,1}^3 that exist
// up:   fine i = coarse floor(i/2)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int child_count(const Grid3& gf, int I, int J, int K) {
  return (min(2 * I + 1, gf.nx - 1) - 2 * I + 1) * (min(2 * J + 1, gf.ny - 1) - 2 * J + 1) *
         (min(2 * K + 1, gf.nz - 1) - 2 * K + 1);
}

__global__ __launch_bounds__(256) void k_pool(const uint16_t* __restrict__ Xf, Grid3 gf,
                                              uint16_t* __restrict__ Xc, Grid3 gc, int C,
    const int* __restrict__ gate) {
  if (gate && *gate) return;  // every right-hand side of the batch is idle
  const int64_t cnt = gc.nb * gc.n * C;
  FS_GRID_STRIDE(t, cnt) {
    int c;
    int64_t v, b;
    split_bvc(t, C, gc.n, cnt, b, v, c);
    int I, J, K;
    voxel_ijk(gc, v, I, J, K);
    float s = 0.f;
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
          if (i < gf.nx && j < gf.ny && k < gf.nz)
            s += bf2f(Xf[(b * gf.n + ((int64_t)k * gf.ny + j) * gf.nx + i) * C + c]);
        }
    Xc[t] = f2bf(s / (float)child_count(gf, I, J, K));
  }
}

// adjoint of k_pool (+ add), then the ReLU mask of the fine activation
__global__ __launch_bounds__(256) void k_pool_bwd(const uint16_t* __restrict__ dXc, Grid3 gc,
                                                  uint16_t* __restrict__ dXf, Grid3 gf, int C,
                                                  const uint16_t* __restrict__ add,
                                                  const uint16_t* __restrict__ mask) {
  const int64_t cnt = gf.nb * gf.n * C;
  FS_GRID_STRIDE(t, cnt) {
    int c;
    int64_t v, b;
    split_bvc(t, C, gf.n, cnt, b, v, c);
    int i, j, k;
    voxel_ijk(gf, v, i, j, k);
    const int I = i / 2, J = j / 2, K = k / 2;
    float g = bf2f(dXc[(b * gc.n + ((int64_t)K * gc.ny + J) * gc.nx + I) * C + c]) /
              (float)child_count(gf, I, J, K);
    if (add) g += bf2f(add[t]);
    if (mask && !(bf2f(mask[t]) > 0.f)) g = 0.f;
    dXf[t] = f2bf(g);
  }
}

__global__ __launch_bounds__(256) void k_up(const uint16_t* __restrict__ Xc, Grid3 gc,
                                            uint16_t* __restrict__ Xf, Grid3 gf, int C,
    const int* __restrict__ gate) {
  if (gate && *gate) return;  // every right-hand side of the batch is idle
  const int64_t cnt = gf.nb * gf.n * C;
  FS_GRID_STRIDE(t, cnt) {
    int c;
    int64_t v, b;
    split_bvc(t, C, gf.n, cnt, b, v, c);
    int i, j, k;
    voxel_ijk(gf, v, i, j, k);
    Xf[t] = Xc[(b * gc.n + ((int64_t)(k / 2) * gc.ny + j / 2) * gc.nx + i / 2) * C + c];
  }
}

// adjoint of k_up (coarse voxel sums its children), then the ReLU mask
__global__ __launch_bounds__(256) void k_up_bwd(const uint16_t* __restrict__ dXf, Grid3 gf,
                                                uint16_t* __restrict__ dXc, Grid3 gc, int C,
                                                const uint16_t* __restrict__ mask) {
  const int64_t cnt = gc.nb * gc.n * C;
  FS_GRID_STRIDE(t, cnt) {
    int c;
    int64_t v, b;
    split_bvc(t, C, gc.n, cnt, b, v, c);
    int I, J, K;
    voxel_ijk(gc, v, I, J, K);
    float s = 0.f;
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
          if (i < gf.nx && j < gf.ny && k < gf.nz)
            s += bf2f(dXf[(b * gf.n + ((int64_t)k * gf.ny + j) * gf.nx + i) * C + c]);
        }
    if (mask && !(bf2f(mask[t]) > 0.f)) s = 0.f;
    dXc[t] = f2bf(s);
  }
}

// ---------------------------------------------------------------------------
// input/output of the network
// ---------------------------------------------------------------------------
// one block per batch element: s = sqrt(mean((r dh)^2)), sc = {s, 1/s}
__global__ __launch_bounds__(256) void k_rms(const float* __restrict__ r, int64_t vs,
                                             const float* __restrict__ dh, int64_t n,
                                             float* __restrict__ sc,
    const int* __restrict__ gate) {
  if (gate && *gate) return;  // every right-hand side of the batch is idle
  const int b = blockIdx.x;
  double a[1] = {0.0};
  for (int64_t v = threadIdx.x; v < n; v += 256) {
    const double x = (double)r[b * vs + v] * dh[v];
    a[0] += x * x;
  }
  block_sum<1>(a);
  if (threadIdx.x == 0) {
    const double s = sqrt(a[0] / (double)(n > 0 ? n : 1));
    sc[2 * b] = (float)s;
    sc[2 * b + 1] = s > 0 ? (float)(1.0 / s) : 0.f;
  }
}

__global__ __launch_bounds__(256) void k_features(const float* __restrict__ r, int64_t vs,
                                                  const float* __restrict__ dh,
                                                  const float* __restrict__ feat4,
                                                  const float* __restrict__ sc, Grid3 g,
                                                  uint16_t* __restrict__ X,
    const int* __restrict__ gate) {
  if (gate && *gate) return;  // every right-hand side of the batch is idle
  const int64_t cnt = g.nb * g.n * 16;
  FS_GRID_STRIDE(t, cnt) {
    int c;
    int64_t v, b;
    split_bvc(t, 16, g.n, cnt, b, v, c);
    float val = 0.f;
    if (c == 0) val = r[b * vs + v] * dh[v] * sc[2 * b + 1];
    else if (c <= 4) val = feat4[v * 4 + (c - 1)];
    X[t] = f2bf(val);
  }
}

__global__ __launch_bounds__(256) void k_take0(const uint16_t* __restrict__ Y,
                                               const float* __restrict__ dh,
                                               const float* __restrict__ sc, Grid3 g,
                                               float* __restrict__ e, int64_t vs,
    const int* __restrict__ gate) {
  if (gate && *gate) return;  // every right-hand side of the batch is idle
  const int64_t cnt = g.nb * g.n;
  FS_GRID_STRIDE(t, cnt) {
    int c;
    int64_t v, b;
    split_bvc(t, 1, g.n, cnt, b, v, c);
    e[b * vs + v] = sc[2 * b] * dh[v] * bf2f(Y[t * 16]);
  }
}

__global__ __launch_bounds__(256) void k_put0(const float* __restrict__ ge, int64_t vs,
                                              const float* __restrict__ dh,
                                              const float* __restrict__ sc, Grid3 g,
                                              uint16_t* __restrict__ dY) {
  const int64_t cnt = g.nb * g.n * 16;
  FS_GRID_STRIDE(t, cnt) {
    int c;
    int64_t v, b;
    split_bvc(t, 16, g.n, cnt, b, v, c);
    dY[t] = c == 0 ? f2bf(sc[2 * b] * dh[v] * ge[b * vs + v]) : (uint16_t)0;
  }
}

// ---------------------------------------------------------------------------
// weight packing, optimiser, conversions
// ---------------------------------------------------------------------------
// forward pack: Wt[co][k], k = tap*cin + ci (zero padded to kpad), from
// W[tap][ci][co] FP32
__global__ __launch_bounds__(256) void k_pack_fwd(const float* __restrict__ W, int cin, int cout,
                                                  int kpad, uint16_t* __restrict__ Wt) {
  FS_GRID_STRIDE(t, (int64_t)cout * kpad) {
    const int co = (int)(t / kpad), kk = (int)(t % kpad);
    const int tap = kk / cin, ci = kk % cin;
    Wt[t] = tap < 27 ? f2bf(W[((int64_t)tap * cin + ci) * cout + co]) : (uint16_t)0;
  }
}

// data-gradient pack: the transposed convolution is a convolution with the
// tap mirrored (26 - tap) and the channel roles swapped: Wb[ci][tap'*cout + co]
__global__ __launch_bounds__(256) void k_pack_bwd(const float* __restrict__ W, int cin, int cout,
                                                  int kpad_b, uint16_t* __restrict__ Wb) {
  FS_GRID_STRIDE(t, (int64_t)cin * kpad_b) {
    const int ci = (int)(t / kpad_b), kk = (int)(t % kpad_b);
    const int tapb = kk / cout, co = kk % cout;
    Wb[t] = tapb < 27 ? f2bf(W[((int64_t)(26 - tapb) * cin + ci) * cout + co]) : (uint16_t)0;
  }
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              int64_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
  FS_GRID_STRIDE(t, n) {
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
  }
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ in, int64_t cnt,
                                                 uint16_t* __restrict__ out) {
  FS_GRID_STRIDE(t, cnt) out[t] = f2bf(in[t]);
}
__global__ __launch_bounds__(256) void k_from_bf16(const uint16_t* __restrict__ in, int64_t cnt,
                                                   float* __restrict__ out) {
  FS_GRID_STRIDE(t, cnt) out[t] = bf2f(in[t]);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static int blocks_for(int64_t cnt) {
  const int64_t b = (cnt + 255) / 256;
  return (int)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

template <int CIN, int COUT>
static void conv_fwd_t(hipStream_t s, const uint16_t* X, const uint16_t* Wt, int kpad,
                       const float* bias, const uint16_t* skip, const uint16_t* mask,
                       uint16_t* Y, const Grid3& g, int relu, const int* gate) {
  // short rows: the 4 waves of a block split K instead of covering 64 voxels
  const int ks = g.nx <= 16 ? 4 : (g.nx <= 32 ? 2 : 1);
  dim3 grid((g.nx + 64 / ks - 1) / (64 / ks), g.ny, g.nz * g.nb);
  hipLaunchKernelGGL((k_conv_fwd<CIN, COUT>), grid, dim3(256), 0, s, X, Wt, kpad, bias, skip,
                     mask, Y, g, relu, gate, ks);
}

#define FS_NN_SHAPES(X) X(16, 16) X(16, 32) X(32, 16) X(32, 32) X(32, 64) X(64, 32) X(64, 64)

int conv_fwd(hipStream_t s, int cin, int cout, const uint16_t* X, const uint16_t* Wt, int kpad,
             const float* bias, const uint16_t* skip, const uint16_t* mask, uint16_t* Y,
             const Grid3& g, int relu, const int* gate) {
  if (kpad != kpad_of(cin)) return -1;
#define FS_CONV(CI, CO)                                                   \
  if (cin == CI && cout == CO) {                                          \
    conv_fwd_t<CI, CO>(s, X, Wt, kpad, bias, skip, mask, Y, g, relu, gate);     \
    return 0;                                                             \
  }
  FS_NN_SHAPES(FS_CONV)
#undef FS_CONV
  return -1;
}

template <int CIN, int COUT>
static void conv_wgrad_t(hipStream_t s, const uint16_t* X, const uint16_t* dY, const Grid3& g,
                         float* part, int nranges, float* dW) {
  const int ntiles = 27 * (CIN / 16) * (COUT / 16);
  const int64_t ntot = g.n * g.nb;
  const int64_t range_len = ((ntot + nranges - 1) / nranges + 31) / 32 * 32;
  hipLaunchKernelGGL((k_conv_wgrad<CIN, COUT>), dim3(nranges, ntiles), dim3(64), 0, s, X, dY, g,
                     range_len, part);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3((ntiles * 256 + 255) / 256), dim3(256), 0, s, part,
                     nranges, ntiles, CIN, COUT, dW);
}

int conv_wgrad(hipStream_t s, int cin, int cout, const uint16_t* X, const uint16_t* dY,
               const Grid3& g, float* part, int nranges, float* dW) {
#define FS_WG(CI, CO)                                        \
  if (cin == CI && cout == CO) {                             \
    conv_wgrad_t<CI, CO>(s, X, dY, g, part, nranges, dW);    \
    return 0;                                                \
  }
  FS_NN_SHAPES(FS_WG)
#undef FS_WG
  return -1;
}

void bgrad(hipStream_t s, const uint16_t* dY, int64_t nvox, int cout, float* db) {
  hipLaunchKernelGGL(k_bgrad, dim3(cout), dim3(256), 0, s, dY, nvox, cout, db);
}
void pool(hipStream_t s, const uint16_t* Xf, const Grid3& gf, uint16_t* Xc, const Grid3& gc,
          int C, const int* gate) {
  hipLaunchKernelGGL(k_pool, dim3(blocks_for(gc.nb * gc.n * C)), dim3(256), 0, s, Xf, gf, Xc, gc,
                     C, gate);
}
void pool_bwd(hipStream_t s, const uint16_t* dXc, const Grid3& gc, uint16_t* dXf,
              const Grid3& gf, int C, const uint16_t* add, const uint16_t* mask) {
  hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for(gf.nb * gf.n * C)), dim3(256), 0, s, dXc, gc,
                     dXf, gf, C, add, mask);
}
void up(hipStream_t s, const uint16_t* Xc, const Grid3& gc, uint16_t* Xf, const Grid3& gf,
        int C, const int* gate) {
  hipLaunchKernelGGL(k_up, dim3(blocks_for(gf.nb * gf.n * C)), dim3(256), 0, s, Xc, gc, Xf, gf,
                     C, gate);
}
void up_bwd(hipStream_t s, const uint16_t* dXf, const Grid3& gf, uint16_t* dXc,
            const Grid3& gc, int C, const uint16_t* mask) {
  hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for(gc.nb * gc.n * C)), dim3(256), 0, s, dXf, gf, dXc,
                     gc, C, mask);
}
void rms(hipStream_t s, const float* r, int64_t vs, const float* dh, const Grid3& g,
         float* sc, const int* gate) {
  hipLaunchKernelGGL(k_rms, dim3(g.nb), dim3(256), 0, s, r, vs, dh, g.n, sc, gate);
}
void features(hipStream_t s, const float* r, int64_t vs, const float* dh, const float* feat4,
              const float* sc, const Grid3& g, uint16_t* X, const int* gate) {
  hipLaunchKernelGGL(k_features, dim3(blocks_for(g.nb * g.n * 16)), dim3(256), 0, s, r, vs, dh,
                     feat4, sc, g, X, gate);
}
void take0(hipStream_t s, const uint16_t* Y, const float* dh, const float* sc, const Grid3& g,
           float* e, int64_t vs, const int* gate) {
  hipLaunchKernelGGL(k_take0, dim3(blocks_for(g.nb * g.n)), dim3(256), 0, s, Y, dh, sc, g, e, vs,
                     gate);
}
void put0(hipStream_t s, const float* ge, int64_t vs, const float* dh, const float* sc,
          const Grid3& g, uint16_t* dY) {
  hipLaunchKernelGGL(k_put0, dim3(blocks_for(g.nb * g.n * 16)), dim3(256), 0, s, ge, vs, dh, sc,
                     g, dY);
}
void pack_fwd(hipStream_t s, const float* W, int cin, int cout, int kpad, uint16_t* Wt) {
  hipLaunchKernelGGL(k_pack_fwd, dim3(blocks_for((int64_t)cout * kpad)), dim3(256), 0, s, W, cin,
                     cout, kpad, Wt);
}
void pack_bwd(hipStream_t s, const float* W, int cin, int cout, int kpad_b, uint16_t* Wb) {
  hipLaunchKernelGGL(k_pack_bwd, dim3(blocks_for((int64_t)cin * kpad_b)), dim3(256), 0, s, W, cin,
                     cout, kpad_b, Wb);
}
void adam(hipStream_t s, float* p, const float* g, float* m, float* v, int64_t n, float lr,
          float b1, float b2, float eps, int step) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps,
                     c1, c2);
}
void to_bf16(hipStream_t s, const float* in, int64_t cnt, uint16_t* out) {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(cnt)), dim3(256), 0, s, in, cnt, out);
}
void from_bf16(hipStream_t s, const uint16_t* in, int64_t cnt, float* out) {
  hipLaunchKernelGGL(k_from_bf16, dim3(blocks_for(cnt)), dim3(256), 0, s, in, cnt, out);
}

}  // namespace nn
}  // namespace fs
