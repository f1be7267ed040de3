// Learned multilevel preconditioner: device-side declarations shared by the
// network kernels (fs_nn.hip), the aggregation hierarchy kernels (fs_mg.hip)
// and their host driver (fs_pc.hip).  Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fs_internal.hpp"

namespace fs {
namespace nn {

// Batched voxel grid of a network level: tensors are [nb][n][C] channels-last
// BF16, voxel v = (k*ny + j)*nx + i.
struct Grid3 {
  int nx, ny, nz;
  int64_t n;  // nx*ny*nz
  int nb;     // batch
};

inline Grid3 grid3(int nx, int ny, int nz, int nb) {
  return Grid3{nx, ny, nz, (int64_t)nx * ny * nz, nb};
}
inline Grid3 half(const Grid3& g) {
  return grid3((g.nx + 1) / 2, (g.ny + 1) / 2, (g.nz + 1) / 2, g.nb);
}
inline int kpad_of(int cin) { return (27 * cin + 31) / 32 * 32; }

// Y = mask(act(conv(X, W) + b + skip)); conv is 3x3x3, zero padded.  W packed
// [cout][kpad] with k = tap*cin + ci.  Returns -1 for an unsupported shape.
int conv_fwd(hipStream_t s, int cin, int cout, const uint16_t* X, const uint16_t* Wt, int kpad,
             const float* bias, const uint16_t* skip, const uint16_t* mask, uint16_t* Y,
             const Grid3& g, int relu, const int* gate = nullptr);
// dW[tap][ci][co] = sum over voxels of X(v + tap) dY(v); part holds
// nranges * 27*(cin/16)*(cout/16) * 256 floats.
int conv_wgrad(hipStream_t s, int cin, int cout, const uint16_t* X, const uint16_t* dY,
               const Grid3& g, float* part, int nranges, float* dW);
void bgrad(hipStream_t s, const uint16_t* dY, int64_t nvox, int cout, float* db);
// gate (optional, every forward kernel): a device flag; nonzero = skip the launch
void pool(hipStream_t s, const uint16_t* Xf, const Grid3& gf, uint16_t* Xc, const Grid3& gc,
          int C, const int* gate = nullptr);
void pool_bwd(hipStream_t s, const uint16_t* dXc, const Grid3& gc, uint16_t* dXf,
              const Grid3& gf, int C, const uint16_t* add, const uint16_t* mask);
void up(hipStream_t s, const uint16_t* Xc, const Grid3& gc, uint16_t* Xf, const Grid3& gf,
        int C, const int* gate = nullptr);
void up_bwd(hipStream_t s, const uint16_t* dXf, const Grid3& gf, uint16_t* dXc,
            const Grid3& gc, int C, const uint16_t* mask);
// per batch element: s = rms(r*dh) over the level, sc[2b] = s, sc[2b+1] = 1/s
void rms(hipStream_t s, const float* r, int64_t vs, const float* dh, const Grid3& g,
         float* sc, const int* gate = nullptr);
// X[b][v] = {r*dh/s, feat4[v][0..3], 0...} (16 channels)
void features(hipStream_t s, const float* r, int64_t vs, const float* dh, const float* feat4,
              const float* sc, const Grid3& g, uint16_t* X, const int* gate = nullptr);
// e[b][v] = s_b * dh[v] * Y[b][v][0]
void take0(hipStream_t s, const uint16_t* Y, const float* dh, const float* sc, const Grid3& g,
           float* e, int64_t vs, const int* gate = nullptr);
// dY[b][v][c] = c == 0 ? s_b * dh[v] * ge[b][v] : 0
void put0(hipStream_t s, const float* ge, int64_t vs, const float* dh, const float* sc,
          const Grid3& g, uint16_t* dY);
void pack_fwd(hipStream_t s, const float* W, int cin, int cout, int kpad, uint16_t* Wt);
void pack_bwd(hipStream_t s, const float* W, int cin, int cout, int kpad_b, uint16_t* Wb);
void adam(hipStream_t s, float* p, const float* g, float* m, float* v, int64_t n, float lr,
          float b1, float b2, float eps, int step);
void to_bf16(hipStream_t s, const float* in, int64_t cnt, uint16_t* out);
void from_bf16(hipStream_t s, const uint16_t* in, int64_t cnt, float* out);

}  // namespace nn

namespace mg {

// One level of the aggregation hierarchy.  Level 0 is the prepared system's
// equilibrated operator (unit diagonal, Dirichlet rows masked); deeper levels
// are Galerkin products P^T A P of piecewise-constant box aggregates, again
// 7-point: y_n = d_n x_n - sum_m c_nm x_m.
struct Level {
  int nx, ny, nz;
  int64_t n, vs;           // nodes, vector stride
  const float* c[3];       // coupling to the + neighbour along x, y, z
  const float* dg;         // diagonal (null on level 0: unit)
  const float* dinv;       // 1/diag on active nodes, 0 elsewhere (null on level 0)
  const uint8_t* mask;     // level 0 only: 0 = free
};

// Aggregation from a level to the next: tensor product of per-axis groupings
// (consecutive fine indices [start[a][I], start[a][I+1]) form coarse index I).
struct Map {
  const int32_t* start[3];  // cn[a] + 1 entries
  const int32_t* f2c[3];    // fine index -> coarse index
  int cn[3];
};

void coarsen(hipStream_t s, const Level& f, const Map& m, float* dg, float* dinv, float* c0,
             float* c1, float* c2);
void bottom_features(hipStream_t s, const Level& L, float* feat4, float* dh);
// rc = P^T (r - A (omega D^-1 r)) for nr vectors
void restrict_(hipStream_t s, const Level& f, const Map& m, const float* r, float* rc,
               int64_t vsc, int nr, float omega, const int* gate = nullptr);
// e = S (omega D^-1 r + P ec) + omega D^-1 r, S = I - omega D^-1 A; optional
// FCG partials <e,r>, <e,q> per vector (partials [2*nr][nblocks])
void prolong(hipStream_t s, const Level& f, const Map& m, const float* r, const float* ec,
             int64_t vsc, float* e, int nr, float omega, const float* q, double* partials,
             int nblocks, const int* gate = nullptr);
void bottom_jacobi(hipStream_t s, const Level& L, const float* r, float* e, int nr,
                   float omega, const int* gate = nullptr);
// *out = 1 when every right-hand side of sc[0..nr) is idle (Krylov mode)
// ext: an external skip flag (a batch's device control), count: cycles run
void gate(hipStream_t s, const CgScalars* sc, int nr, int mode, int* out,
          const int* ext = nullptr, int* count = nullptr);
// training loss on level 0: w = A z; g = 2 (w - r) * gscale[b]; partials
// [2*nr][nblocks] of z.w and z.r
void loss_grad(hipStream_t s, const Level& L, const float* z, const float* r,
               const float* gscale, float* g, int nr, double* partials, int nblocks);
// r = A e for FP64 e (pairs are made from exact errors), FP32 outputs of e and r
void pair_make(hipStream_t s, const Level& L, const double* const c64[3], const double* e,
               float* e32, float* r32, double* partials, int nblocks);
// Level 0 on the stencil tiles (fs_sys tiling for the prolongation; a
// z-tiling aligned to the z-groups for the restriction, which needs the x
// and y groupings of level 0 to be pairs {2I, 2I+1}).
struct L0Args {
  Dims d;
  int64_t vs;
  const float* c[3];
  const uint8_t* mask;  // level 0: Dirichlet mask (unit diagonal)
  const float* dg;      // a coarse level: diagonal and inverse diagonal (else null)
  const float* dinv;
  Tiling tl;            // prolongation tiles (the system's)
  Tiling tlr;           // restriction tiles
  int x4;               // rows a multiple of 4 (and x pairs): 4 nodes per lane
  Tiling tl4, tlr4;     // the same tiles in units of 256-node row segments
  const int* kstart;    // tlr.tiles_z + 1 z-starts, each a z-group start
  Map m;                // level 0 -> 1
};
// returns the number of blocks (the partials' stride)
// optional profiling: start/stop events of the launch and the counters of
// the kernel kind ([0] iterations with a live RHS, [1] live RHS summed);
// gate (optional): a device flag, nonzero = skip the launch
// zhalf: store e in FP16 scaled by zscale(sc[b]) (level 0 of a 4-wide cycle in
// a flexible-CG solve; returns -1 otherwise)
int prolong_l0(hipStream_t s, const L0Args& a, const float* r, const float* ec, int64_t vsc,
               float* e, int nr, float omega, const float* q, double* partials,
               const CgScalars* sc, int mode, hipEvent_t ea = nullptr, hipEvent_t eb = nullptr,
               unsigned long long* prof = nullptr, const int* gate = nullptr, int zhalf = 0);
void restrict_l0(hipStream_t s, const L0Args& a, const float* r, float* rc, int64_t vsc, int nr,
                 float omega, const CgScalars* sc, int mode, hipEvent_t ea = nullptr,
                 hipEvent_t eb = nullptr, unsigned long long* prof = nullptr,
                 const int* gate = nullptr);
// bottom-level training loss: c_b = r^T D^-1 r, then g = 2 (A y - r) / (c_b nr)
void jacobi_norm(hipStream_t s, const Level& L, const float* r, int nr, double* partials,
                 int nblocks);
void energy_grad(hipStream_t s, const Level& L, const float* y, const float* r, const double* c,
                 float* g, int nr, double* partials, int nblocks);
int vec_blocks(int64_t n);

}  // namespace mg
}  // namespace fs
