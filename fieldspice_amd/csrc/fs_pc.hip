// Host driver of the learned multilevel preconditioner: the network object
// (parameters, Adam state, packed BF16 weights), its forward and backward
// passes over the MFMA kernels, the aggregation hierarchy of a prepared
// system, the cycle used inside flexible CG, and training on device-resident
// residual/error pairs.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "fieldspice_hip.h"
#include "fs_common.hpp"
#include "fs_device.hpp"
#include "fs_internal.hpp"
#include "fs_launch.hpp"
#include "fs_nn.hpp"

using fs::dalloc;
using fs::dfree;
using fs::fail;
namespace nn = fs::nn;
namespace mg = fs::mg;

// --------------------------------------------------------------------------
// network layout
// --------------------------------------------------------------------------
namespace {

constexpr int kLayers = 11;
constexpr int kCin[kLayers] = {16, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16};
constexpr int kCout[kLayers] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 16};
// network level (1 = bottom grid of the hierarchy, 2, 3 = pooled) of each layer
constexpr int kLevel[kLayers] = {1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 1};

// activations (A) and gradients (H) of one batch, channels-last BF16
enum Tensor : int {
  T_X0, T_A1, T_A2, T_P2, T_A3, T_A4, T_P4, T_A5, T_A6, T_U6, T_A7, T_A8, T_U8, T_A9, T_A10,
  T_Y,
  T_DY, T_H10, T_H9, T_DU8, T_H8, T_H7, T_DU6, T_H6, T_H5, T_DP4, T_H4, T_H3, T_DP2, T_H2, T_H1,
  T_COUNT
};
// (level, channels) of each tensor
constexpr int kTLevel[T_COUNT] = {1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 2, 1, 1, 1, 1,
                                  1, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 2, 1, 1};
constexpr int kTChan[T_COUNT] = {16, 16, 16, 16, 32, 32, 32, 64, 64, 64, 32, 32, 32, 16, 16, 16,
                                 16, 16, 16, 32, 32, 32, 64, 64, 64, 32, 32, 32, 16, 16, 16};

int wgrad_ranges(int li, int64_t nvox) {
  const int ntiles = 27 * (kCin[li] / 16) * (kCout[li] / 16);
  const int64_t by_work = std::max<int64_t>(1, nvox / 512);
  return (int)std::max<int64_t>(1, std::min<int64_t>(std::max(1, 4096 / ntiles), by_work));
}

}  // namespace

struct NetWs {
  nn::Grid3 g[4]{};  // g[1..3]
  int cap_nb = 0;
  int64_t cap_n1 = 0;
  uint16_t* t[T_COUNT] = {nullptr};
  float* part = nullptr;
  int64_t part_cap = 0;
};

struct fs_net {
  fs_ctx* ctx = nullptr;
  int64_t off_w[kLayers]{}, off_b[kLayers]{};
  int kpad[kLayers]{}, kpadb[kLayers]{};
  uint16_t* Wt[kLayers] = {nullptr};
  uint16_t* Wb[kLayers] = {nullptr};
  int64_t nparams = 0;
  float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  int adam_step = 0;
};

namespace {

void free_ws(NetWs& w) {
  for (auto& p : w.t) dfree(p);
  dfree(w.part);
  w.cap_nb = 0;
  w.cap_n1 = 0;
}

// size the workspace for a bottom grid (nx, ny, nz) and batch nb
int ensure_net_ws(NetWs& w, int nx, int ny, int nz, int nb) {
  const nn::Grid3 g1 = nn::grid3(nx, ny, nz, nb);
  const nn::Grid3 g2 = nn::half(g1), g3 = nn::half(g2);
  const bool fits = w.cap_nb >= nb && w.g[1].nx == nx && w.g[1].ny == ny && w.g[1].nz == nz;
  w.g[1] = g1;
  w.g[2] = g2;
  w.g[3] = g3;
  if (fits) return FS_OK;
  free_ws(w);
  w.g[1] = g1;
  w.g[2] = g2;
  w.g[3] = g3;
  for (int t = 0; t < T_COUNT; ++t) {
    const int64_t cnt = w.g[kTLevel[t]].n * nb * kTChan[t];
    if (int rc = dalloc(&w.t[t], (size_t)cnt)) {
      free_ws(w);
      return rc;
    }
  }
  int64_t pc = 1;
  for (int li = 0; li < kLayers; ++li) {
    const int64_t nvox = w.g[kLevel[li]].n * nb;
    const int ntiles = 27 * (kCin[li] / 16) * (kCout[li] / 16);
    pc = std::max<int64_t>(pc, (int64_t)wgrad_ranges(li, nvox) * ntiles * 256);
  }
  if (int rc = dalloc(&w.part, (size_t)pc)) {
    free_ws(w);
    return rc;
  }
  w.part_cap = pc;
  w.cap_nb = nb;
  w.cap_n1 = g1.n;
  return FS_OK;
}

void set_batch(NetWs& w, int nb) {
  for (int l = 1; l <= 3; ++l) w.g[l].nb = nb;
}

int repack(fs_net* net) {
  hipStream_t s = net->ctx->stream;
  for (int li = 0; li < kLayers; ++li) {
    nn::pack_fwd(s, net->params + net->off_w[li], kCin[li], kCout[li], net->kpad[li],
                 net->Wt[li]);
    nn::pack_bwd(s, net->params + net->off_w[li], kCin[li], kCout[li], net->kpadb[li],
                 net->Wb[li]);
  }
  FS_HIP(hipGetLastError());
  return FS_OK;
}

#define FS_NNCALL(expr)                                                        \
  do {                                                                         \
    if ((expr) != 0) return fail(FS_EINVAL, "unsupported layer shape: %s", #expr); \
  } while (0)

// forward on stream s: ws.t[T_X0] -> ws.t[T_Y]; gate (optional device flag)
// skips every launch
int net_forward(fs_net* net, NetWs& w, hipStream_t s, const int* gate = nullptr) {
  auto conv = [&](int li, Tensor in, Tensor out, const uint16_t* skip, int relu) {
    const int lv = kLevel[li];
    return nn::conv_fwd(s, kCin[li], kCout[li], w.t[in], net->Wt[li], net->kpad[li],
                        net->params + net->off_b[li], skip, nullptr, w.t[out], w.g[lv], relu,
                        gate);
  };
  FS_NNCALL(conv(0, T_X0, T_A1, nullptr, 1));
  FS_NNCALL(conv(1, T_A1, T_A2, nullptr, 1));
  nn::pool(s, w.t[T_A2], w.g[1], w.t[T_P2], w.g[2], 16, gate);
  FS_NNCALL(conv(2, T_P2, T_A3, nullptr, 1));
  FS_NNCALL(conv(3, T_A3, T_A4, nullptr, 1));
  nn::pool(s, w.t[T_A4], w.g[2], w.t[T_P4], w.g[3], 32, gate);
  FS_NNCALL(conv(4, T_P4, T_A5, nullptr, 1));
  FS_NNCALL(conv(5, T_A5, T_A6, nullptr, 1));
  nn::up(s, w.t[T_A6], w.g[3], w.t[T_U6], w.g[2], 64, gate);
  FS_NNCALL(conv(6, T_U6, T_A7, w.t[T_A4], 1));
  FS_NNCALL(conv(7, T_A7, T_A8, nullptr, 1));
  nn::up(s, w.t[T_A8], w.g[2], w.t[T_U8], w.g[1], 32, gate);
  FS_NNCALL(conv(8, T_U8, T_A9, w.t[T_A2], 1));
  FS_NNCALL(conv(9, T_A9, T_A10, nullptr, 1));
  FS_NNCALL(conv(10, T_A10, T_Y, nullptr, 0));
  FS_HIP(hipGetLastError());
  return FS_OK;
}

// backward: ws.t[T_DY] -> net->grad (overwritten); needs the activations of
// the forward pass of the same batch
int net_backward(fs_net* net, NetWs& w, hipStream_t s) {
  // data gradient of layer li: out = mask(conv^T(in) + add)
  auto bwd = [&](int li, Tensor in, Tensor out, Tensor mask_t) {
    const int lv = kLevel[li];
    return nn::conv_fwd(s, kCout[li], kCin[li], w.t[in], net->Wb[li], net->kpadb[li], nullptr,
                        nullptr, mask_t == T_COUNT ? nullptr : w.t[mask_t], w.t[out], w.g[lv], 0);
  };
  auto wg = [&](int li, Tensor x, Tensor h) {
    const int lv = kLevel[li];
    const int64_t nvox = w.g[lv].n * w.g[lv].nb;
    const int nr = wgrad_ranges(li, nvox);
    if (nn::conv_wgrad(s, kCin[li], kCout[li], w.t[x], w.t[h], w.g[lv], w.part, nr,
                       net->grad + net->off_w[li]))
      return -1;
    nn::bgrad(s, w.t[h], nvox, kCout[li], net->grad + net->off_b[li]);
    return 0;
  };
  FS_NNCALL(wg(10, T_A10, T_DY));
  FS_NNCALL(bwd(10, T_DY, T_H10, T_A10));
  FS_NNCALL(wg(9, T_A9, T_H10));
  FS_NNCALL(bwd(9, T_H10, T_H9, T_A9));
  FS_NNCALL(wg(8, T_U8, T_H9));
  FS_NNCALL(bwd(8, T_H9, T_DU8, T_COUNT));
  nn::up_bwd(s, w.t[T_DU8], w.g[1], w.t[T_H8], w.g[2], 32, w.t[T_A8]);
  FS_NNCALL(wg(7, T_A7, T_H8));
  FS_NNCALL(bwd(7, T_H8, T_H7, T_A7));
  FS_NNCALL(wg(6, T_U6, T_H7));
  FS_NNCALL(bwd(6, T_H7, T_DU6, T_COUNT));
  nn::up_bwd(s, w.t[T_DU6], w.g[2], w.t[T_H6], w.g[3], 64, w.t[T_A6]);
  FS_NNCALL(wg(5, T_A5, T_H6));
  FS_NNCALL(bwd(5, T_H6, T_H5, T_A5));
  FS_NNCALL(wg(4, T_P4, T_H5));
  FS_NNCALL(bwd(4, T_H5, T_DP4, T_COUNT));
  nn::pool_bwd(s, w.t[T_DP4], w.g[3], w.t[T_H4], w.g[2], 32, w.t[T_H7], w.t[T_A4]);
  FS_NNCALL(wg(3, T_A3, T_H4));
  FS_NNCALL(bwd(3, T_H4, T_H3, T_A3));
  FS_NNCALL(wg(2, T_P2, T_H3));
  FS_NNCALL(bwd(2, T_H3, T_DP2, T_COUNT));
  nn::pool_bwd(s, w.t[T_DP2], w.g[2], w.t[T_H2], w.g[1], 16, w.t[T_H9], w.t[T_A2]);
  FS_NNCALL(wg(1, T_A1, T_H2));
  FS_NNCALL(bwd(1, T_H2, T_H1, T_A1));
  FS_NNCALL(wg(0, T_X0, T_H1));
  FS_HIP(hipGetLastError());
  return FS_OK;
}

void free_net(fs_net* net) {
  if (!net) return;
  for (int li = 0; li < kLayers; ++li) {
    dfree(net->Wt[li]);
    dfree(net->Wb[li]);
  }
  dfree(net->params);
  dfree(net->grad);
  dfree(net->m);
  dfree(net->v);
  delete net;
}

// --------------------------------------------------------------------------
// small vector kernels of the host driver
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_f64_to_f32(const double* __restrict__ in, int64_t n,
                                                    float* __restrict__ out) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
    out[t] = (float)in[t];
}
__global__ __launch_bounds__(256) void k_f32_to_f64(const float* __restrict__ in, int64_t n,
                                                    double* __restrict__ out) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
    out[t] = (double)in[t];
}
__global__ __launch_bounds__(256) void k_mul_inplace(double* __restrict__ x,
                                                     const double* __restrict__ w, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
    x[t] *= w[t];
}
__global__ __launch_bounds__(256) void k_scale_copy(const float* __restrict__ in, int64_t n,
                                                    float sgn, float* __restrict__ out) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
    out[t] = sgn * in[t];
}
// out[row] = ordered sum of partials[row][0..nb) (one block per row)
__global__ __launch_bounds__(256) void k_sum_rows(const double* __restrict__ partials, int nb,
                                                  double* __restrict__ out) {
  const double s = fs::sum_partials(partials + (int64_t)blockIdx.x * nb, nb);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

int grid_of(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 4096); }

}  // namespace

// --------------------------------------------------------------------------
// preconditioner object
// --------------------------------------------------------------------------
struct fs_pc {
  fs_sys* sys = nullptr;  // null once the system was destroyed
  fs_ctx* ctx = nullptr;
  fs_net* net = nullptr;  // not owned
  int K = 0;              // aggregation steps; the network solves level K
  float omega = 0.8f;
  std::vector<mg::Level> lv;  // K + 1 levels
  std::vector<mg::Map> maps;  // K maps
  mg::L0Args l0{};            // level 0 on the stencil tiles
  bool l0_pairs = false;      // level-0 x/y groupings are pairs: tiled restriction
  // levels 1..K-1 on 4-wide z-march tiles (x4 set where the level qualifies)
  std::vector<mg::L0Args> lx;
  int* kstart = nullptr;
  std::vector<float*> owned;  // coarse operator arrays
  std::vector<int32_t*> owned_i;
  std::vector<float*> rv, ev;  // per level >= 1: [kMaxRhs][vs]
  float *feat4 = nullptr, *dh = nullptr, *sc = nullptr;
  float* gK = nullptr;     // bottom-loss gradient [kMaxRhs][vs_K]
  int* gate = nullptr;     // 1 while every right-hand side of the solve is idle
  double* cK = nullptr;    // bottom-loss normalisation per batch element
  NetWs ws;
  // training
  int npairs = 0, cap_pairs = 0;
  float *pe = nullptr, *pr = nullptr;  // [cap][vs0]
  std::vector<double> pE;              // ||e||_A^2 per pair
  float *bz = nullptr, *br = nullptr, *bg = nullptr;  // [kMaxRhs][vs0]
  float* gscale = nullptr;                           // [steps][kMaxRhs] during training
  int64_t gscale_cap = 0;
  double* part = nullptr;  // reduction partials
  double* sums = nullptr;  // k_sum_rows outputs
  int64_t sums_cap = 0;
  int nblk0 = 0;
};

namespace {

void free_pc(fs_pc* pc) {
  if (!pc) return;
  for (auto& p : pc->owned) dfree(p);
  for (auto& p : pc->owned_i) dfree(p);
  dfree(pc->kstart);
  for (auto& p : pc->rv) dfree(p);
  for (auto& p : pc->ev) dfree(p);
  dfree(pc->feat4);
  dfree(pc->dh);
  dfree(pc->sc);
  dfree(pc->gK);
  dfree(pc->gate);
  dfree(pc->cK);
  free_ws(pc->ws);
  dfree(pc->pe);
  dfree(pc->pr);
  dfree(pc->bz);
  dfree(pc->br);
  dfree(pc->bg);
  dfree(pc->gscale);
  dfree(pc->part);
  dfree(pc->sums);
  delete pc;
}

int ensure_sums(fs_pc* pc, int64_t rows) {
  if (pc->sums_cap >= rows) return FS_OK;
  dfree(pc->sums);
  if (int rc = dalloc(&pc->sums, (size_t)rows)) return rc;
  pc->sums_cap = rows;
  return FS_OK;
}

}  // namespace

namespace fs {

// One application of the cycle to nr level-0 vectors r0 -> z0 (stride vs0).
// With q and partials given, the level-0 prolongation also leaves the block
// partials of <z,r> and <z,q> ([2*nr][*nblocks]).  Right-hand sides that sc
// marks idle (mode as the Krylov kernels) are skipped on level 0.
int pc_cycle(fs_pc* pc, int nr, const float* r0, float* z0, const float* q, double* partials,
             int* nblocks, const CgScalars* sc, int mode, ProfSlot* slot,
             unsigned long long* prof_dev, int zhalf) {
  // level-0 timing (FS_PROF_RESTRICT / FS_PROF_PROLONG) when profiled
  unsigned long long* pr = slot && prof_dev ? prof_dev + 2 * FS_PROF_RESTRICT : nullptr;
  unsigned long long* pp = slot && prof_dev ? prof_dev + 2 * FS_PROF_PROLONG : nullptr;
  // (the tiled level-0 restriction runs when level 0 is grouped in x/y pairs)
  const hipEvent_t ra = pc->l0_pairs ? prof_start(slot, FS_PROF_RESTRICT) : nullptr;
  const hipEvent_t rb = pc->l0_pairs ? prof_stop(slot, FS_PROF_RESTRICT) : nullptr;
  const hipEvent_t pa = prof_start(slot, FS_PROF_PROLONG), pb = prof_stop(slot, FS_PROF_PROLONG);
  hipStream_t s = pc->sys->stream;
  const int K = pc->K;
  // in a solve, every kernel below skips its work once all right-hand sides
  // are idle (graph replays keep launching finished systems), or, in a
  // batch's captured cycle start, while the batch's device control says so
  // (the level-0 kernels, which otherwise skip idle right-hand sides one by
  // one, then take the gate too)
  const int* g = nullptr;
  const int* ext = sc ? pc->sys->gate : nullptr;
  if (sc) {
    mg::gate(s, sc, nr, mode, pc->gate, ext, pc->sys->pc_count);
    g = pc->gate;
  }
  const int* g0 = ext ? g : nullptr;
  for (int l = 0; l < K; ++l) {
    if (l == 0 && pc->l0_pairs)
      mg::restrict_l0(s, pc->l0, r0, pc->rv[1], pc->lv[1].vs, nr, pc->omega, sc, mode, ra, rb,
                      pr, g0);
    else if (l > 0 && pc->lx[l].x4)
      mg::restrict_l0(s, pc->lx[l], pc->rv[l], pc->rv[l + 1], pc->lv[l + 1].vs, nr, pc->omega,
                      sc, mode, nullptr, nullptr, nullptr, g0);
    else
      mg::restrict_(s, pc->lv[l], pc->maps[l], l == 0 ? r0 : pc->rv[l], pc->rv[l + 1],
                    pc->lv[l + 1].vs, nr, pc->omega, g);
  }
  const mg::Level& B = pc->lv[K];
  if (pc->net) {
    NetWs& w = pc->ws;
    set_batch(w, nr);
    nn::rms(s, pc->rv[K], B.vs, pc->dh, w.g[1], pc->sc, g);
    nn::features(s, pc->rv[K], B.vs, pc->dh, pc->feat4, pc->sc, w.g[1], w.t[T_X0], g);
    if (int rc = net_forward(pc->net, w, s, g)) return rc;
    nn::take0(s, w.t[T_Y], pc->dh, pc->sc, w.g[1], pc->ev[K], B.vs, g);
  } else {
    mg::bottom_jacobi(s, B, pc->rv[K], pc->ev[K], nr, pc->omega, g);
  }
  for (int l = K - 1; l >= 1; --l) {
    if (pc->lx[l].x4)
      mg::prolong_l0(s, pc->lx[l], pc->rv[l], pc->ev[l + 1], pc->lv[l + 1].vs, pc->ev[l], nr,
                     pc->omega, nullptr, nullptr, sc, mode, nullptr, nullptr, nullptr, g0);
    else
      mg::prolong(s, pc->lv[l], pc->maps[l], pc->rv[l], pc->ev[l + 1], pc->lv[l + 1].vs,
                  pc->ev[l], nr, pc->omega, nullptr, nullptr, 0, g);
  }
  const int nb = mg::prolong_l0(s, pc->l0, r0, pc->ev[1], pc->lv[1].vs, z0, nr, pc->omega, q,
                                partials, sc, mode, pa, pb, pp, g0, zhalf);
  if (nb < 0) return FS_EINVAL;
  if (nblocks) *nblocks = nb;
  FS_HIP(hipGetLastError());
  return FS_OK;
}

int64_t pc_level1_nodes(const fs_pc* pc) { return pc && pc->K > 0 ? pc->lv[1].n : 0; }

int pc_z_half(const fs_pc* pc) { return pc && pc->K > 0 && pc->l0.x4 ? 1 : 0; }

int pc_level0_width(const fs_pc* pc) {
  return pc && pc->l0_pairs && pc->l0.x4 ? 4 : 1;
}

// Solver hook (fs_kernels.hip, FS_PREC_FCG): z32 = M(r32) with FCG partials.
int pc_apply_solver(fs_sys* s, int nr, int* nblocks, ProfSlot* slot) {
  return pc_cycle(s->pc, nr, s->r32, s->z32, s->q32, s->partials, nblocks, s->sc, 2, slot,
                  s->prof_dev, pc_z_half(s->pc));
}

void pc_orphan(fs_pc* pc) {
  if (pc->sys && pc->sys->pc == pc) pc->sys->pc = nullptr;
  pc->sys = nullptr;
}

}  // namespace fs

namespace {

// Loss and gradient of one batch held in pc->br (r), with pc->gscale + goff
// holding 1/(E_b nb).  Leaves per-element z.Az and z.r in pc->sums[sum_off..]
// ([2*nb]) and the parameter gradient in net->grad.
int batch_loss_grad(fs_pc* pc, int nb, const float* gscale, int64_t sum_off) {
  hipStream_t s = pc->sys->stream;
  if (int rc = fs::pc_cycle(pc, nb, pc->br, pc->bz, nullptr, nullptr, nullptr, nullptr, 0))
    return rc;
  const mg::Level& L0 = pc->lv[0];
  mg::loss_grad(s, L0, pc->bz, pc->br, gscale, pc->bg, nb, pc->part, pc->nblk0);
  hipLaunchKernelGGL(k_sum_rows, dim3(2 * nb), dim3(256), 0, s, pc->part, pc->nblk0,
                     pc->sums + sum_off);
  // gradient with respect to the bottom error: the transpose of the up-sweep
  // (smoothing then prolongation) is the down-sweep restriction
  for (int l = 0; l < pc->K; ++l) {
    if (l == 0 && pc->l0_pairs)
      mg::restrict_l0(s, pc->l0, pc->bg, pc->rv[1], pc->lv[1].vs, nb, pc->omega, nullptr, 0);
    else
      mg::restrict_(s, pc->lv[l], pc->maps[l], l == 0 ? pc->bg : pc->rv[l], pc->rv[l + 1],
                    pc->lv[l + 1].vs, nb, pc->omega);
  }
  NetWs& w = pc->ws;
  nn::put0(s, pc->rv[pc->K], pc->lv[pc->K].vs, pc->dh, pc->sc, w.g[1], w.t[T_DY]);
  if (int rc = net_backward(pc->net, w, s)) return rc;
  FS_HIP(hipGetLastError());
  return FS_OK;
}

// Bottom loss of one batch held in pc->br: the down-sweep gives r_K, the
// network y = e_K, and L_b = (y^T A_K y - 2 y^T r_K) / (r_K^T D_K^-1 r_K)
// (the energy functional of A_K e = r_K, minimal at the exact bottom solve).
// Leaves y.Ay, y.r and c per element in pc->sums[sum_off..] ([3*nb]).
int batch_bottom_loss_grad(fs_pc* pc, int nb, int64_t sum_off) {
  hipStream_t s = pc->sys->stream;
  const int K = pc->K;
  for (int l = 0; l < K; ++l) {
    if (l == 0 && pc->l0_pairs)
      mg::restrict_l0(s, pc->l0, pc->br, pc->rv[1], pc->lv[1].vs, nb, pc->omega, nullptr, 0);
    else
      mg::restrict_(s, pc->lv[l], pc->maps[l], l == 0 ? pc->br : pc->rv[l], pc->rv[l + 1],
                    pc->lv[l + 1].vs, nb, pc->omega);
  }
  const mg::Level& B = pc->lv[K];
  NetWs& w = pc->ws;
  set_batch(w, nb);
  nn::rms(s, pc->rv[K], B.vs, pc->dh, w.g[1], pc->sc);
  nn::features(s, pc->rv[K], B.vs, pc->dh, pc->feat4, pc->sc, w.g[1], w.t[T_X0]);
  if (int rc = net_forward(pc->net, w, s)) return rc;
  nn::take0(s, w.t[T_Y], pc->dh, pc->sc, w.g[1], pc->ev[K], B.vs);
  const int nbk = mg::vec_blocks(B.n);
  mg::jacobi_norm(s, B, pc->rv[K], nb, pc->part, nbk);
  hipLaunchKernelGGL(k_sum_rows, dim3(nb), dim3(256), 0, s, pc->part, nbk, pc->cK);
  mg::energy_grad(s, B, pc->ev[K], pc->rv[K], pc->cK, pc->gK, nb, pc->part, nbk);
  hipLaunchKernelGGL(k_sum_rows, dim3(2 * nb), dim3(256), 0, s, pc->part, nbk,
                     pc->sums + sum_off);
  FS_HIP(hipMemcpyAsync(pc->sums + sum_off + 2 * nb, pc->cK, nb * sizeof(double),
                        hipMemcpyDeviceToDevice, s));
  nn::put0(s, pc->gK, B.vs, pc->dh, pc->sc, w.g[1], w.t[T_DY]);
  if (int rc = net_backward(pc->net, w, s)) return rc;
  FS_HIP(hipGetLastError());
  return FS_OK;
}

// e (node space, host C order) -> pair (e32, r32) at slot pointers; returns
// E = e_s . r_s through *E (synchronous)
int make_pair(fs_pc* pc, const double* e_host, float* e32, float* r32, double* E) {
  fs_sys* sys = pc->sys;
  fs_ctx* ctx = sys->ctx;
  const fs::Dims& d = sys->d;
  double* tmp = sys->node_tmp;  // [vstride] doubles at least
  double* tmp2 = sys->node_tmp + sys->vstride;
  FS_HIP(hipMemcpyAsync(tmp2, e_host, (size_t)d.n * sizeof(double), hipMemcpyHostToDevice,
                        ctx->stream));
  fs::permute_in<double>(ctx->stream, tmp2, tmp, d.nx, d.ny, d.nz, d.nx, d.ny);
  hipLaunchKernelGGL(k_mul_inplace, dim3(grid_of(d.n)), dim3(256), 0, ctx->stream, tmp,
                     sys->dhalf, d.n);
  const double* c64[3] = {sys->c64[0], sys->c64[1], sys->c64[2]};
  mg::pair_make(ctx->stream, pc->lv[0], c64, tmp, e32, r32, pc->part, pc->nblk0);
  if (int rc = ensure_sums(pc, 1)) return rc;
  hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, ctx->stream, pc->part, pc->nblk0,
                     pc->sums);
  FS_HIP(hipGetLastError());
  FS_HIP(hipMemcpyAsync(E, pc->sums, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FS_HIP(hipStreamSynchronize(ctx->stream));
  return FS_OK;
}

int adam_update(fs_net* net, float lr) {
  net->adam_step += 1;
  nn::adam(net->ctx->stream, net->params, net->grad, net->m, net->v, net->nparams, lr, 0.9f,
           0.999f, 1e-8f, net->adam_step);
  return repack(net);
}

}  // namespace

extern "C" {

// --------------------------------------------------------------------------
// network
// --------------------------------------------------------------------------
int fs_net_create(fs_ctx* ctx, uint64_t seed, fs_net** out) {
  FS_GUARD_BEGIN
  if (!ctx || !out) return fail(FS_EINVAL, "NULL argument");
  *out = nullptr;
  FS_HIP(hipSetDevice(ctx->device));
  fs_net* net = new fs_net();
  net->ctx = ctx;
  int64_t off = 0;
  for (int li = 0; li < kLayers; ++li) {
    net->off_w[li] = off;
    off += 27LL * kCin[li] * kCout[li];
    net->off_b[li] = off;
    off += kCout[li];
    net->kpad[li] = nn::kpad_of(kCin[li]);
    net->kpadb[li] = nn::kpad_of(kCout[li]);
  }
  net->nparams = off;
  int rc = FS_OK;
  if ((rc = dalloc(&net->params, (size_t)off)) || (rc = dalloc(&net->grad, (size_t)off)) ||
      (rc = dalloc(&net->m, (size_t)off)) || (rc = dalloc(&net->v, (size_t)off))) {
    free_net(net);
    return rc;
  }
  for (int li = 0; li < kLayers && rc == FS_OK; ++li) {
    rc = dalloc(&net->Wt[li], (size_t)kCout[li] * net->kpad[li]);
    if (rc == FS_OK) rc = dalloc(&net->Wb[li], (size_t)kCin[li] * net->kpadb[li]);
  }
  if (rc != FS_OK) {
    free_net(net);
    return rc;
  }
  // He-normal weights (the output layer scaled down), zero biases
  std::vector<float> h((size_t)off, 0.0f);
  std::mt19937_64 gen(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (int li = 0; li < kLayers; ++li) {
    float sd = std::sqrt(2.0f / (27.0f * kCin[li]));
    if (li == kLayers - 1) sd *= 0.1f;
    for (int64_t t = 0; t < 27LL * kCin[li] * kCout[li]; ++t) h[net->off_w[li] + t] = sd * nd(gen);
  }
  if (hipMemcpy(net->params, h.data(), off * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(net->m, 0, off * sizeof(float)) != hipSuccess ||
      hipMemset(net->v, 0, off * sizeof(float)) != hipSuccess ||
      hipMemset(net->grad, 0, off * sizeof(float)) != hipSuccess) {
    free_net(net);
    return fail(FS_EHIP, "network initialisation failed");
  }
  if ((rc = repack(net)) || (rc = (hipStreamSynchronize(ctx->stream) == hipSuccess ? FS_OK : fail(FS_EHIP, "pack failed")))) {
    free_net(net);
    return rc;
  }
  *out = net;
  return FS_OK;
  FS_GUARD_END
}

int fs_net_destroy(fs_net* net) {
  if (net) (void)hipSetDevice(net->ctx->device);
  free_net(net);
  return FS_OK;
}

int fs_net_num_params(const fs_net* net, int64_t* count) {
  if (!net || !count) return fail(FS_EINVAL, "NULL argument");
  *count = net->nparams;
  return FS_OK;
}

int fs_net_get_params(fs_net* net, float* params) {
  FS_GUARD_BEGIN
  if (!net || !params) return fail(FS_EINVAL, "NULL argument");
  FS_HIP(hipSetDevice(net->ctx->device));
  FS_HIP(hipStreamSynchronize(net->ctx->stream));
  FS_HIP(hipMemcpy(params, net->params, net->nparams * sizeof(float), hipMemcpyDeviceToHost));
  return FS_OK;
  FS_GUARD_END
}

int fs_net_set_params(fs_net* net, const float* params) {
  FS_GUARD_BEGIN
  if (!net || !params) return fail(FS_EINVAL, "NULL argument");
  for (int64_t t = 0; t < net->nparams; ++t)
    if (!std::isfinite(params[t])) return fail(FS_EINVAL, "parameter %lld is not finite", (long long)t);
  FS_HIP(hipSetDevice(net->ctx->device));
  FS_HIP(hipStreamSynchronize(net->ctx->stream));
  FS_HIP(hipMemcpy(net->params, params, net->nparams * sizeof(float), hipMemcpyHostToDevice));
  FS_HIP(hipMemset(net->m, 0, net->nparams * sizeof(float)));
  FS_HIP(hipMemset(net->v, 0, net->nparams * sizeof(float)));
  net->adam_step = 0;
  if (int rc = repack(net)) return rc;
  FS_HIP(hipStreamSynchronize(net->ctx->stream));
  return FS_OK;
  FS_GUARD_END
}

int fs_net_eval(fs_net* net, int nx, int ny, int nz, int nb, const float* X, float* Y,
                const float* dY, float* grad) {
  FS_GUARD_BEGIN
  if (!net || !X || !Y) return fail(FS_EINVAL, "NULL argument");
  if (nx < 1 || ny < 1 || nz < 1 || nb < 1 || nb > 64 || ny > 65535 ||
      (int64_t)nz * nb > 65535)
    return fail(FS_EINVAL, "grid %dx%dx%d batch %d out of range", nx, ny, nz, nb);
  if (dY && !grad) return fail(FS_EINVAL, "dY given without grad");
  fs_ctx* ctx = net->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  NetWs w;
  int rc = ensure_net_ws(w, nx, ny, nz, nb);
  const int64_t cnt = w.g[1].n * nb * 16;
  float* tmp = nullptr;
  if (rc == FS_OK) rc = dalloc(&tmp, (size_t)cnt);
  auto run = [&]() -> int {
    hipStream_t s = ctx->stream;
    FS_HIP(hipMemcpyAsync(tmp, X, cnt * sizeof(float), hipMemcpyHostToDevice, s));
    nn::to_bf16(s, tmp, cnt, w.t[T_X0]);
    if (int r = net_forward(net, w, s)) return r;
    nn::from_bf16(s, w.t[T_Y], cnt, tmp);
    FS_HIP(hipMemcpyAsync(Y, tmp, cnt * sizeof(float), hipMemcpyDeviceToHost, s));
    FS_HIP(hipStreamSynchronize(s));
    if (dY) {
      FS_HIP(hipMemcpyAsync(tmp, dY, cnt * sizeof(float), hipMemcpyHostToDevice, s));
      nn::to_bf16(s, tmp, cnt, w.t[T_DY]);
      if (int r = net_backward(net, w, s)) return r;
      FS_HIP(hipMemcpyAsync(grad, net->grad, net->nparams * sizeof(float),
                            hipMemcpyDeviceToHost, s));
      FS_HIP(hipStreamSynchronize(s));
    }
    return FS_OK;
  };
  if (rc == FS_OK) rc = run();
  dfree(tmp);
  free_ws(w);
  return rc;
  FS_GUARD_END
}

// --------------------------------------------------------------------------
// hierarchy
// --------------------------------------------------------------------------
int fs_pc_create(fs_sys* sys, fs_net* net, int nlevels, const int32_t* groups, double omega,
                 fs_pc** out) {
  FS_GUARD_BEGIN
  if (!sys || !out || (!groups && nlevels > 0)) return fail(FS_EINVAL, "NULL argument");
  *out = nullptr;
  if (sys->d.per) return fail(FS_EINVAL, "periodic operators run fs_sys_solve and fs_capacitance only (FS_PREC_FP64 or FS_PREC_MIXED)");
  if (nlevels < 1 || nlevels > 12) return fail(FS_EINVAL, "nlevels must be in 1..12");
  if (!(omega > 0.0 && omega < 2.0)) return fail(FS_EINVAL, "omega must be in (0, 2)");
  if (net && net->ctx != sys->ctx) return fail(FS_EINVAL, "network and system on different contexts");
  fs_ctx* ctx = sys->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  // validate the groupings and derive the level dims on the host
  std::vector<std::array<int, 3>> dims(nlevels + 1);
  dims[0] = {sys->d.nx, sys->d.ny, sys->d.nz};
  std::vector<std::array<std::vector<int32_t>, 3>> starts(nlevels);
  const int32_t* gp = groups;
  for (int l = 0; l < nlevels; ++l) {
    for (int a = 0; a < 3; ++a) {
      const int nf = dims[l][a];
      std::vector<int32_t>& st = starts[l][a];
      if (gp[0] != 0) return fail(FS_EINVAL, "level %d axis %d: grouping must start at 0", l, a);
      st.push_back(0);
      for (int i = 1; i < nf; ++i) {
        const int32_t dlt = gp[i] - gp[i - 1];
        if (dlt != 0 && dlt != 1)
          return fail(FS_EINVAL, "level %d axis %d: grouping must step by 0 or 1", l, a);
        if (dlt == 1) st.push_back(i);
      }
      st.push_back(nf);
      dims[l + 1][a] = (int)st.size() - 1;
      gp += nf;
    }
  }
  fs_pc* pc = new fs_pc();
  pc->sys = sys;
  pc->ctx = ctx;
  pc->net = net;
  pc->K = nlevels;
  pc->omega = (float)omega;
  pc->nblk0 = mg::vec_blocks(sys->d.n);
  auto fail_pc = [&](int rc) {
    free_pc(pc);
    return rc;
  };
  // level 0 = the prepared system (FP32 couplings, mask)
  pc->lv.resize(nlevels + 1);
  mg::Level& L0 = pc->lv[0];
  L0.nx = sys->d.nx;
  L0.ny = sys->d.ny;
  L0.nz = sys->d.nz;
  L0.n = sys->d.n;
  L0.vs = sys->vstride;
  for (int a = 0; a < 3; ++a) L0.c[a] = sys->c32[a];
  L0.dg = nullptr;
  L0.dinv = nullptr;
  L0.mask = sys->mask;
  pc->rv.assign(nlevels + 1, nullptr);
  pc->ev.assign(nlevels + 1, nullptr);
  pc->maps.resize(nlevels);
  int rc = FS_OK;
  for (int l = 0; l < nlevels && rc == FS_OK; ++l) {
    mg::Map& m = pc->maps[l];
    for (int a = 0; a < 3; ++a) {
      const std::vector<int32_t>& st = starts[l][a];
      const int nf = dims[l][a];
      std::vector<int32_t> f2c(nf);
      for (int I = 0; I + 1 < (int)st.size(); ++I)
        for (int i = st[I]; i < st[I + 1]; ++i) f2c[i] = I;
      int32_t *ds = nullptr, *df = nullptr;
      if ((rc = dalloc(&ds, st.size())) || (rc = dalloc(&df, f2c.size()))) {
        dfree(ds);
        break;
      }
      pc->owned_i.push_back(ds);
      pc->owned_i.push_back(df);
      if (hipMemcpy(ds, st.data(), st.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(df, f2c.data(), f2c.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(FS_EHIP, "hierarchy upload failed");
        break;
      }
      m.start[a] = ds;
      m.f2c[a] = df;
      m.cn[a] = dims[l + 1][a];
    }
    if (rc) break;
    mg::Level& C = pc->lv[l + 1];
    C.nx = dims[l + 1][0];
    C.ny = dims[l + 1][1];
    C.nz = dims[l + 1][2];
    C.n = (int64_t)C.nx * C.ny * C.nz;
    C.vs = (C.n + 63) / 64 * 64;
    C.mask = nullptr;
    float* arr[5] = {nullptr};
    for (int t = 0; t < 5 && rc == FS_OK; ++t) {
      rc = dalloc(&arr[t], (size_t)C.n);
      if (rc == FS_OK) pc->owned.push_back(arr[t]);
    }
    if (rc) break;
    float* dg = arr[0];
    float* dinv = arr[1];
    C.dg = dg;
    C.dinv = dinv;
    C.c[0] = arr[2];
    C.c[1] = arr[3];
    C.c[2] = arr[4];
    mg::coarsen(ctx->stream, pc->lv[l], m, dg, dinv, arr[2], arr[3], arr[4]);
    if ((rc = dalloc(&pc->rv[l + 1], (size_t)fs::kMaxRhs * C.vs)) ||
        (rc = dalloc(&pc->ev[l + 1], (size_t)fs::kMaxRhs * C.vs)))
      break;
    FS_HIP(hipMemset(pc->rv[l + 1], 0, sizeof(float) * fs::kMaxRhs * C.vs));
    FS_HIP(hipMemset(pc->ev[l + 1], 0, sizeof(float) * fs::kMaxRhs * C.vs));
  }
  if (rc) return fail_pc(rc);
  {
    // level 0 on the stencil tiles; the tiled restriction needs x/y pairs and
    // z-tiles made of whole z-groups
    mg::L0Args& a = pc->l0;
    a.d = sys->d;
    a.vs = sys->vstride;
    for (int q = 0; q < 3; ++q) a.c[q] = sys->c32[q];
    a.mask = sys->mask;
    a.tl = sys->tl;
    a.m = pc->maps[0];
    bool pairs = true;
    for (int q = 0; q < 2; ++q) {
      const std::vector<int32_t>& st = starts[0][q];
      for (size_t I = 0; I + 1 < st.size(); ++I)
        pairs = pairs && st[I] == (int32_t)(2 * I) && st[I + 1] == std::min<int32_t>(2 * I + 2, dims[0][q]);
    }
    pc->l0_pairs = pairs;
    if (pairs) {
      const std::vector<int32_t>& zs = starts[0][2];  // z-group starts + nz
      std::vector<int> ks{0};
      for (size_t G = 1; G < zs.size(); ++G)
        if (zs[G] - ks.back() >= sys->tl.kz || G + 1 == zs.size()) ks.push_back(zs[G]);
      if (ks.back() != sys->d.nz) ks.push_back(sys->d.nz);
      a.tlr = sys->tl;
      a.tlr.tiles_z = (int)ks.size() - 1;
      a.tlr.ntiles = a.tlr.tiles_x * a.tlr.tiles_y * a.tlr.tiles_z;
      a.tlr.nblocks = (a.tlr.ntiles + 7) / 8 * 8;
      if ((rc = dalloc(&pc->kstart, ks.size()))) return fail_pc(rc);
      if (hipMemcpy(pc->kstart, ks.data(), ks.size() * sizeof(int), hipMemcpyHostToDevice) !=
          hipSuccess)
        return fail_pc(fail(FS_EHIP, "hierarchy upload failed"));
      a.kstart = pc->kstart;
      // 4 nodes per lane when rows are a multiple of 4 (x pairs never cross a lane)
      // (the 4-wide kernels address nodes with 32-bit offsets)
      a.x4 = sys->d.nx % 4 == 0 && sys->d.n + sys->d.sxy < (int64_t)INT32_MAX ? 1 : 0;
      if (a.x4) {
        a.tl4 = sys->tl4;
        {
          // the prolongation marches 32 planes per tile when that still
          // leaves 512 tiles (2 waves per SIMD x 4 waves per block fill the
          // part once): the tiles' first and last planes, read twice, are
          // then a sixteenth of the reads (256^3, 5 RHS: 227 -> 220 us)
          fs::Tiling& t = a.tl4;
          const int64_t txy = (int64_t)t.tiles_x * t.tiles_y;
          if (t.kz < 32 && txy * ((sys->d.nz + 31) / 32) >= 512) {
            t.kz = 32;
            t.tiles_z = (sys->d.nz + 31) / 32;
            t.ntiles = (int)(txy * t.tiles_z);
            t.nblocks = (t.ntiles + 7) / 8 * 8;
          }
        }
        a.tlr4 = sys->tl4;
        a.tlr4.tiles_z = a.tlr.tiles_z;
        a.tlr4.ntiles = a.tlr4.tiles_x * a.tlr4.tiles_y * a.tlr4.tiles_z;
        a.tlr4.nblocks = (a.tlr4.ntiles + 7) / 8 * 8;
        const int64_t need = (int64_t)std::max(a.tl4.nblocks, a.tlr4.nblocks) * fs::kMaxRhs * 2 +
                             64 * (fs::kMaxRhs + 1) * fs::kMaxRhs;
        if (need > sys->partials_cap) a.x4 = 0;  // keep within the system's partials
      }
    }
  }
  // Coarse levels 1..K-1 on the same 4-wide z-march kernels (with their
  // diagonal) where x and y are grouped in pairs and rows hold a multiple of
  // 4 nodes, at least 128 (a wave spans 256 nodes of one row).
  pc->lx.assign(nlevels, mg::L0Args{});
  for (int l = 1; l < nlevels; ++l) {
    const mg::Level& L = pc->lv[l];
    bool ok = L.nx % 4 == 0 && L.nx >= 16 && L.n + (int64_t)L.nx * L.ny < (int64_t)INT32_MAX;
    for (int q = 0; q < 2 && ok; ++q) {
      const std::vector<int32_t>& st = starts[l][q];
      for (size_t I = 0; I + 1 < st.size(); ++I)
        ok = ok && st[I] == (int32_t)(2 * I) &&
             st[I + 1] == std::min<int32_t>(2 * I + 2, dims[l][q]);
    }
    if (!ok) continue;
    mg::L0Args& a = pc->lx[l];
    a.d = fs::Dims{L.nx, L.ny, L.nz, L.n, (int64_t)L.nx * L.ny};
    a.vs = L.vs;
    for (int q = 0; q < 3; ++q) a.c[q] = L.c[q];
    a.mask = nullptr;
    a.dg = L.dg;
    a.dinv = L.dinv;
    a.m = pc->maps[l];
    fs::Tiling t{};
    t.tiles_x = (L.nx + 255) / 256;
    t.tiles_y = (L.ny + fs::kTileY - 1) / fs::kTileY;
    const int64_t txy = (int64_t)t.tiles_x * t.tiles_y;
    t.kz = (int)std::min<int64_t>(16, std::max<int64_t>(2, (L.nz * txy + 1023) / 1024));
    t.tiles_z = (L.nz + t.kz - 1) / t.kz;
    t.ntiles = t.tiles_x * t.tiles_y * t.tiles_z;
    t.nblocks = (t.ntiles + 7) / 8 * 8;
    a.tl4 = t;
    const std::vector<int32_t>& zs = starts[l][2];
    std::vector<int> ks{0};
    for (size_t G = 1; G < zs.size(); ++G)
      if (zs[G] - ks.back() >= t.kz || G + 1 == zs.size()) ks.push_back(zs[G]);
    if (ks.back() != L.nz) ks.push_back(L.nz);
    a.tlr4 = t;
    a.tlr4.tiles_z = (int)ks.size() - 1;
    a.tlr4.ntiles = a.tlr4.tiles_x * a.tlr4.tiles_y * a.tlr4.tiles_z;
    a.tlr4.nblocks = (a.tlr4.ntiles + 7) / 8 * 8;
    int32_t* kd = nullptr;
    if ((rc = dalloc(&kd, ks.size()))) return fail_pc(rc);
    pc->owned_i.push_back(kd);
    if (hipMemcpy(kd, ks.data(), ks.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return fail_pc(fail(FS_EHIP, "hierarchy upload failed"));
    a.kstart = kd;
    a.x4 = 1;
  }
  const mg::Level& B = pc->lv[nlevels];
  if (net && (B.ny > 65535 || (int64_t)B.nz * fs::kMaxRhs > 65535))
    return fail_pc(fail(FS_EINVAL, "bottom level %dx%dx%d too large for the network", B.nx,
                        B.ny, B.nz));
  if ((rc = dalloc(&pc->feat4, (size_t)B.n * 4)) || (rc = dalloc(&pc->dh, (size_t)B.n)) ||
      (rc = dalloc(&pc->sc, (size_t)2 * fs::kMaxRhs)) ||
      (rc = dalloc(&pc->gK, (size_t)fs::kMaxRhs * B.vs)) ||
      (rc = dalloc(&pc->gate, 1)) ||
      (rc = dalloc(&pc->cK, (size_t)fs::kMaxRhs)) ||
      (rc = dalloc(&pc->part, (size_t)2 * fs::kMaxRhs * std::max(pc->nblk0, 2048) + 64)) ||
      (rc = ensure_sums(pc, 64)))
    return fail_pc(rc);
  mg::bottom_features(ctx->stream, B, pc->feat4, pc->dh);
  if (net && (rc = ensure_net_ws(pc->ws, B.nx, B.ny, B.nz, fs::kMaxRhs))) return fail_pc(rc);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
    return fail_pc(fail(FS_EHIP, "hierarchy setup kernels failed"));
  sys->pcs.push_back(pc);
  *out = pc;
  return FS_OK;
  FS_GUARD_END
}

int fs_pc_destroy(fs_pc* pc) {
  if (!pc) return FS_OK;
  (void)hipSetDevice(pc->ctx->device);
  if (fs_sys* s = pc->sys) {
    if (s->pc == pc) s->pc = nullptr;
    s->pcs.erase(std::remove(s->pcs.begin(), s->pcs.end(), pc), s->pcs.end());
  }
  free_pc(pc);
  return FS_OK;
}

int fs_pc_level_dims(fs_pc* pc, int level, int* nx, int* ny, int* nz) {
  if (!pc || !nx || !ny || !nz) return fail(FS_EINVAL, "NULL argument");
  if (level < 0 || level > pc->K) return fail(FS_EINVAL, "level %d out of range", level);
  *nx = pc->lv[level].nx;
  *ny = pc->lv[level].ny;
  *nz = pc->lv[level].nz;
  return FS_OK;
}

int fs_sys_set_precond(fs_sys* sys, fs_pc* pc) {
  if (!sys) return fail(FS_EINVAL, "NULL argument");
  if (pc && pc->sys != sys) return fail(FS_EINVAL, "preconditioner belongs to another system");
  sys->pc = pc;
  return FS_OK;
}

int fs_pc_apply(fs_pc* pc, int nrhs, const double* r, double* z) {
  FS_GUARD_BEGIN
  if (!pc || !r || !z) return fail(FS_EINVAL, "NULL argument");
  if (!pc->sys) return fail(FS_EINVAL, "the preconditioner's system was destroyed");
  if (nrhs < 1) return fail(FS_EINVAL, "nrhs must be >= 1");
  fs_sys* sys = pc->sys;
  fs_ctx* ctx = sys->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  const fs::Dims& d = sys->d;
  if (int rc = fs::ensure_solver_ws(sys, fs::kMaxRhs)) return rc;
  if (!pc->br) {
    const size_t cnt = (size_t)fs::kMaxRhs * sys->vstride;
    int rc;
    if ((rc = dalloc(&pc->br, cnt)) || (rc = dalloc(&pc->bz, cnt)) || (rc = dalloc(&pc->bg, cnt)))
      return rc;
    FS_HIP(hipMemset(pc->br, 0, cnt * sizeof(float)));
  }
  for (int r0 = 0; r0 < nrhs; r0 += fs::kMaxRhs) {
    const int nr = std::min(fs::kMaxRhs, nrhs - r0);
    double* tmp = sys->node_tmp;
    double* tmp2 = sys->node_tmp + sys->vstride;
    for (int b = 0; b < nr; ++b) {
      FS_HIP(hipMemcpyAsync(tmp2, r + (size_t)(r0 + b) * d.n, d.n * sizeof(double),
                            hipMemcpyHostToDevice, ctx->stream));
      fs::permute_in<double>(ctx->stream, tmp2, tmp, d.nx, d.ny, d.nz, d.nx, d.ny);
      hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_of(d.n)), dim3(256), 0, ctx->stream, tmp, d.n,
                         pc->br + (size_t)b * sys->vstride);
    }
    if (int rc = fs::pc_cycle(pc, nr, pc->br, pc->bz, nullptr, nullptr, nullptr, nullptr, 0))
      return rc;
    for (int b = 0; b < nr; ++b) {
      hipLaunchKernelGGL(k_f32_to_f64, dim3(grid_of(d.n)), dim3(256), 0, ctx->stream,
                         pc->bz + (size_t)b * sys->vstride, d.n, tmp);
      fs::permute_out<double>(ctx->stream, tmp, tmp2, d.nx, d.ny, d.nz, d.nx, d.ny);
      FS_HIP(hipMemcpyAsync(z + (size_t)(r0 + b) * d.n, tmp2, d.n * sizeof(double),
                            hipMemcpyDeviceToHost, ctx->stream));
    }
    FS_HIP(hipGetLastError());
    FS_HIP(hipStreamSynchronize(ctx->stream));
  }
  return FS_OK;
  FS_GUARD_END
}

// --------------------------------------------------------------------------
// training
// --------------------------------------------------------------------------
namespace {
int ensure_batch(fs_pc* pc) {
  fs_sys* sys = pc->sys;
  if (int rc = fs::ensure_solver_ws(sys, 2)) return rc;
  if (!pc->br) {
    const size_t cnt = (size_t)fs::kMaxRhs * sys->vstride;
    int rc;
    if ((rc = dalloc(&pc->br, cnt)) || (rc = dalloc(&pc->bz, cnt)) || (rc = dalloc(&pc->bg, cnt)))
      return rc;
    FS_HIP(hipMemset(pc->br, 0, cnt * sizeof(float)));
  }
  return FS_OK;
}
}  // namespace

int fs_pc_add_pairs(fs_pc* pc, int npairs, const double* e) {
  FS_GUARD_BEGIN
  if (!pc || !e) return fail(FS_EINVAL, "NULL argument");
  if (!pc->sys) return fail(FS_EINVAL, "the preconditioner's system was destroyed");
  if (npairs < 0) return fail(FS_EINVAL, "npairs must be >= 0");
  fs_sys* sys = pc->sys;
  FS_HIP(hipSetDevice(sys->ctx->device));
  if (int rc = ensure_batch(pc)) return rc;
  const int64_t vs = sys->vstride;
  if (pc->npairs + npairs > pc->cap_pairs) {
    const int cap = std::max(pc->npairs + npairs, 2 * pc->cap_pairs);
    float *ne = nullptr, *nr = nullptr;
    int rc;
    if ((rc = dalloc(&ne, (size_t)cap * vs))) return rc;
    if ((rc = dalloc(&nr, (size_t)cap * vs))) {
      dfree(ne);
      return rc;
    }
    if (pc->npairs > 0) {
      FS_HIP(hipMemcpy(ne, pc->pe, (size_t)pc->npairs * vs * sizeof(float), hipMemcpyDeviceToDevice));
      FS_HIP(hipMemcpy(nr, pc->pr, (size_t)pc->npairs * vs * sizeof(float), hipMemcpyDeviceToDevice));
    }
    dfree(pc->pe);
    dfree(pc->pr);
    pc->pe = ne;
    pc->pr = nr;
    pc->cap_pairs = cap;
  }
  for (int t = 0; t < npairs; ++t) {
    const int slot = pc->npairs;
    double E = 0.0;
    if (int rc = make_pair(pc, e + (size_t)t * sys->d.n, pc->pe + (size_t)slot * vs,
                           pc->pr + (size_t)slot * vs, &E))
      return rc;
    if (!(E > 0.0) || !std::isfinite(E))
      return fail(FS_EINVAL, "pair %d has no energy on the free nodes (E = %g)", t, E);
    pc->pE.push_back(E);
    pc->npairs += 1;
  }
  return FS_OK;
  FS_GUARD_END
}

int fs_pc_num_pairs(fs_pc* pc, int* npairs) {
  if (!pc || !npairs) return fail(FS_EINVAL, "NULL argument");
  *npairs = pc->npairs;
  return FS_OK;
}

int fs_pc_clear_pairs(fs_pc* pc) {
  if (!pc) return fail(FS_EINVAL, "NULL argument");
  (void)hipSetDevice(pc->ctx->device);
  dfree(pc->pe);
  dfree(pc->pr);
  pc->npairs = 0;
  pc->cap_pairs = 0;
  pc->pE.clear();
  return FS_OK;
}

int fs_pc_loss_grad(fs_pc* pc, int kind, int nb, const double* e, double* loss, float* grad) {
  FS_GUARD_BEGIN
  if (!pc || !e || !loss) return fail(FS_EINVAL, "NULL argument");
  if (!pc->sys) return fail(FS_EINVAL, "the preconditioner's system was destroyed");
  if (!pc->net) return fail(FS_EINVAL, "preconditioner has no network");
  if (kind != FS_LOSS_CYCLE && kind != FS_LOSS_BOTTOM) return fail(FS_EINVAL, "unknown loss %d", kind);
  if (nb < 1 || nb > fs::kMaxRhs) return fail(FS_EINVAL, "nb must be in 1..%d", fs::kMaxRhs);
  fs_sys* sys = pc->sys;
  fs_ctx* ctx = sys->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  if (int rc = ensure_batch(pc)) return rc;
  std::vector<double> E(nb);
  for (int b = 0; b < nb; ++b) {
    if (int rc = make_pair(pc, e + (size_t)b * sys->d.n, pc->bz, pc->br + (size_t)b * sys->vstride,
                           &E[b]))
      return rc;
    if (!(E[b] > 0.0)) return fail(FS_EINVAL, "error %d has no energy on the free nodes", b);
  }
  if (int rc = ensure_sums(pc, 3 * nb)) return rc;
  std::vector<double> sums(3 * nb);
  if (kind == FS_LOSS_CYCLE) {
    std::vector<float> gs(fs::kMaxRhs, 0.0f);
    for (int b = 0; b < nb; ++b) gs[b] = (float)(1.0 / (E[b] * nb));
    if (pc->gscale_cap < fs::kMaxRhs) {
      dfree(pc->gscale);
      if (int rc = dalloc(&pc->gscale, fs::kMaxRhs)) return rc;
      pc->gscale_cap = fs::kMaxRhs;
    }
    FS_HIP(hipMemcpy(pc->gscale, gs.data(), fs::kMaxRhs * sizeof(float), hipMemcpyHostToDevice));
    if (int rc = batch_loss_grad(pc, nb, pc->gscale, 0)) return rc;
  } else {
    if (int rc = batch_bottom_loss_grad(pc, nb, 0)) return rc;
  }
  FS_HIP(hipMemcpyAsync(sums.data(), pc->sums, (kind == FS_LOSS_CYCLE ? 2 : 3) * nb * sizeof(double),
                        hipMemcpyDeviceToHost, ctx->stream));
  FS_HIP(hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < nb; ++b) {
    const double norm = kind == FS_LOSS_CYCLE ? E[b] : (sums[2 * nb + b] > 0 ? sums[2 * nb + b] : 1.0);
    loss[b] = (sums[b] - 2.0 * sums[nb + b]) / norm + (kind == FS_LOSS_CYCLE ? 1.0 : 0.0);
  }
  if (grad)
    FS_HIP(hipMemcpy(grad, pc->net->grad, pc->net->nparams * sizeof(float),
                     hipMemcpyDeviceToHost));
  return FS_OK;
  FS_GUARD_END
}

int fs_pc_train(fs_pc* pc, int kind, int steps, int batch, double lr, uint64_t seed,
                double* losses) {
  FS_GUARD_BEGIN
  if (!pc) return fail(FS_EINVAL, "NULL argument");
  if (!pc->sys) return fail(FS_EINVAL, "the preconditioner's system was destroyed");
  if (!pc->net) return fail(FS_EINVAL, "preconditioner has no network");
  if (steps < 1) return fail(FS_EINVAL, "steps must be >= 1");
  if (batch < 1 || batch > fs::kMaxRhs) return fail(FS_EINVAL, "batch must be in 1..%d", fs::kMaxRhs);
  if (pc->npairs < 1) return fail(FS_EINVAL, "no training pairs (fs_pc_add_pairs)");
  if (!(lr > 0.0)) return fail(FS_EINVAL, "lr must be > 0");
  if (kind != FS_LOSS_CYCLE && kind != FS_LOSS_BOTTOM) return fail(FS_EINVAL, "unknown loss %d", kind);
  fs_sys* sys = pc->sys;
  fs_ctx* ctx = sys->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  if (int rc = ensure_batch(pc)) return rc;
  const int64_t vs = sys->vstride;
  // draw every batch and its loss scales up front (no host sync per step)
  std::mt19937_64 gen(seed);
  std::vector<int> idx((size_t)steps * batch);
  std::vector<float> sgn((size_t)steps * batch);
  std::vector<float> gs((size_t)steps * fs::kMaxRhs, 0.0f);
  for (int st = 0; st < steps; ++st)
    for (int b = 0; b < batch; ++b) {
      const int p = (int)(gen() % (uint64_t)pc->npairs);
      idx[(size_t)st * batch + b] = p;
      sgn[(size_t)st * batch + b] = (gen() & 1) ? 1.0f : -1.0f;
      gs[(size_t)st * fs::kMaxRhs + b] = (float)(1.0 / (pc->pE[p] * batch));
    }
  if (pc->gscale_cap < (int64_t)steps * fs::kMaxRhs) {
    dfree(pc->gscale);
    if (int rc = dalloc(&pc->gscale, (size_t)steps * fs::kMaxRhs)) return rc;
    pc->gscale_cap = (int64_t)steps * fs::kMaxRhs;
  }
  FS_HIP(hipMemcpy(pc->gscale, gs.data(), gs.size() * sizeof(float), hipMemcpyHostToDevice));
  const int per = kind == FS_LOSS_CYCLE ? 2 : 3;  // sums per element and step
  if (int rc = ensure_sums(pc, (int64_t)per * batch * steps)) return rc;
  const int64_t n = sys->d.n;
  for (int st = 0; st < steps; ++st) {
    for (int b = 0; b < batch; ++b) {
      const int p = idx[(size_t)st * batch + b];
      hipLaunchKernelGGL(k_scale_copy, dim3(grid_of(n)), dim3(256), 0, ctx->stream,
                         pc->pr + (size_t)p * vs, n, sgn[(size_t)st * batch + b],
                         pc->br + (size_t)b * vs);
    }
    const int64_t off = (int64_t)per * batch * st;
    if (kind == FS_LOSS_CYCLE) {
      if (int rc = batch_loss_grad(pc, batch, pc->gscale + (size_t)st * fs::kMaxRhs, off))
        return rc;
    } else if (int rc = batch_bottom_loss_grad(pc, batch, off)) {
      return rc;
    }
    if (int rc = adam_update(pc->net, (float)lr)) return rc;
  }
  FS_HIP(hipGetLastError());
  std::vector<double> sums((size_t)per * batch * steps);
  FS_HIP(hipMemcpyAsync(sums.data(), pc->sums, sums.size() * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
  FS_HIP(hipStreamSynchronize(ctx->stream));
  if (losses) {
    for (int st = 0; st < steps; ++st) {
      double acc = 0.0;
      for (int b = 0; b < batch; ++b) {
        const size_t o = (size_t)per * batch * st;
        const double zaz = sums[o + b];
        const double zr = sums[o + batch + b];
        if (kind == FS_LOSS_CYCLE) {
          acc += (zaz - 2.0 * zr) / pc->pE[idx[(size_t)st * batch + b]] + 1.0;
        } else {
          const double c = sums[o + 2 * batch + b];
          acc += (zaz - 2.0 * zr) / (c > 0 ? c : 1.0);
        }
      }
      losses[st] = acc / batch;
    }
  }
  return FS_OK;
  FS_GUARD_END
}

}  // extern "C"
