// Internal definitions shared by the kernels and the C ABI of the
// fieldspice MI355X library.  Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "fieldspice_hip.h"

namespace fs {

constexpr int kWave = 64;         // CDNA wavefront
constexpr int kBlock = 256;       // threads per block for every kernel here
constexpr int kMaxRhs = 8;        // right-hand sides one launch carries
constexpr int kTileX = 64;        // stencil tile: one wavefront along x ...
constexpr int kTileY = 4;         // ... times four rows
constexpr int kMaxIslands = 1024; // island coarse space: islands per system
constexpr int kIslChunk = 2048;   // island nodes per reduction block
constexpr int kProjMax = 16;      // time loops: basis of the projection initial guess

// Device layout: node (i, j, k) of an (nx, ny, nz) node grid lives at
// n = (k*ny + j)*nx + i (x fastest).  Every node vector of a system is padded
// to `vstride` elements so that right-hand sides start on 256-byte bounds.
struct Dims {
  int nx, ny, nz;
  int64_t n;       // nx*ny*nz
  int64_t sxy;     // nx*ny
  int per = 0;     // axes that wrap around (bit a for axis a): the last node
                   // couples to the first through the last edge slot
};

// Tile decomposition of the stencil kernels.  Block b of the launch is
// remapped so that the 8 XCDs (which receive blocks round robin) each sweep
// one contiguous range of tiles: z-neighbour planes and the halo rows of a
// tile are then produced and consumed inside one XCD's L2.
struct Tiling {
  int tiles_x, tiles_y, tiles_z;
  int kz;          // z planes marched per tile
  int ntiles;      // tiles_x*tiles_y*tiles_z
  int nblocks;     // ntiles rounded up to a multiple of 8
};

// Per right-hand-side Krylov scalars, written only by the single-block
// finalisation kernels and read by every block of the following kernels.
struct CgScalars {
  double rho;       // r.r of the current residual
  double pq;        // p.q
  double alpha;
  double beta;
  double thresh2;   // stop when rho < thresh2  (== (tol*||b_s||)^2)
  double bnorm2;    // ||b_s||^2
  double zr;        // flexible CG: <z, r> of the current preconditioned residual
  double scale;     // FP32 cycle scaling: r32 = r64 * scale
  double aux0, aux1, aux2, aux3;
  int iter;         // iterations completed (over all cycles)
  int cycle_iter;   // iterations in the current FP32 cycle
  int maxiter;
  int done;         // 0 running, 1 converged, 2 iteration cap, 3 breakdown
  int cycle_done;   // mixed precision: FP32 cycle reached its target
  int refreshes;
  int pad0, pad1;
};

// Flexible CG over a 4-wide level-0 cycle keeps its preconditioned residual
// z in FP16 as z / zscale: the FP32 cycle's current residual norm |r32|_2
// (aux0 = r32.r32), which bounds |z_n| / zscale by |M^-1|_2.
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float zscale(const CgScalars& c) {
  return c.aux0 > 0.0 ? (float)sqrt(c.aux0) : 1.f;
}

// Device control of one system of a batched solve, written by the control
// kernels that follow each chunk of the system's iterations in the batch's
// captured graph.
struct BatchCtl {
  int refresh_skip;          // 0: this replay refreshes the system's FP64 residual
  int start_skip;            // 0: this replay starts a new FP32 cycle
  int pc_applies;            // preconditioner applications that ran
  int finished;              // every right-hand side done
  unsigned long long stamp;  // device wall clock when first seen finished
};

// Events of one profiled Krylov iteration: one start/stop pair per kernel
// kind (FS_PROF_*), and the kinds launched with them in the current chunk.
struct ProfSlot {
  hipEvent_t a[FS_PROF_KINDS] = {};
  hipEvent_t b[FS_PROF_KINDS] = {};
  unsigned used = 0;
};

}  // namespace fs

// --------------------------------------------------------------------------
// ABI objects
// --------------------------------------------------------------------------
struct fs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_cu = 0;
  char name[256] = {0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  unsigned char* scratch = nullptr;  // this device's fs_sys_solve_noalloc scratch, or null
};

struct fs_op {
  fs_ctx* ctx = nullptr;
  int Nx = 0, Ny = 0, Nz = 0;  // cells
  int ex[3] = {0, 0, 0};       // edges per line along each axis (cells, or nodes when periodic)
  fs::Dims d{};
  double* w[3] = {nullptr, nullptr, nullptr};  // node-padded edge weights (x-fast)
  double* extra = nullptr;                     // node diagonal, or null
};

struct fs_pc;

struct fs_sys {
  fs_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;  // the context's stream; a batch solve lends its own
  const fs_op* op = nullptr;  // not owned; must outlive the system
  fs::Dims d{};
  fs::Tiling tl{};
  fs::Tiling tl4{};            // 4-nodes-per-lane layout (x tiles of 256 nodes): the
                               // preconditioner's level-0 kernels and the deflated stencil
  int use_x4 = 0;              // rows are a multiple of 4 and that layout is selected
  int xw = 0;                  // nodes per lane of the FP32 Krylov stencil (2, 3, 4; 0: one
                               // node per thread on the 64-wide tiles)
  fs::Tiling tlw{};            // its tiles (x tiles of 64*xw nodes)
  fs::Tiling tl2{};            // 2-nodes-per-lane FP64 stencil (x tiles of 128 nodes)
  int use_x2 = 0;              // rows are a multiple of 2
  int64_t vstride = 0;
  uint8_t* mask = nullptr;     // 0 free, 1 Dirichlet, 2 pinned (x-fast)
  double* dhalf = nullptr;     // sqrt(D) on free nodes, 0 elsewhere
  double* dinv = nullptr;      // 1/sqrt(D) on free nodes, 0 elsewhere
  double* c64[3] = {nullptr, nullptr, nullptr};  // equilibrated masked couplings
  float* c32[3] = {nullptr, nullptr, nullptr};
  fs_sys_info info{};
  // solve workspace, grown on demand (capacity in right-hand sides)
  int ws_nr = 0;
  double *bs = nullptr, *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;
  float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *q32 = nullptr;
  float* z32 = nullptr;              // flexible CG: preconditioned residual
  double* node_tmp = nullptr;        // host-layout staging (nr node vectors)
  double* partials = nullptr;        // reduction partials
  int64_t partials_cap = 0;
  fs::CgScalars* sc = nullptr;       // device scalars (kMaxRhs)
  fs::CgScalars* sc_host = nullptr;  // pinned host mirror
  int vec_blocks = 0;                // grid of the node-wise vector kernels
  // optional timing of the full-grid kernels of the Krylov loop (FS_PROF_*)
  int profiling = 0;
  int ev_cap = 0;                        // profiled iterations per chunk
  fs::ProfSlot* prof_slots = nullptr;    // ev_cap slots
  double prof_time[FS_PROF_KINDS] = {};
  int64_t prof_launches[FS_PROF_KINDS] = {};
  // per kind k: [2k] launches that found a live RHS, [2k+1] live RHS summed
  unsigned long long* prof_dev = nullptr;
  // flexible CG
  fs_pc* pc = nullptr;        // attached preconditioner (not owned)
  std::vector<fs_pc*> pcs;    // every preconditioner built on this system
  int acc_fallbacks = 0;      // counters of the current solve call
  int acc_pc_applies = 0;
  int acc_replacements = 0;   // FP64 residual replacements of the current solve call
  // device control of a batched solve, set while its graph is captured: the
  // refresh and cycle-start kernels then exit at once while *gate is nonzero,
  // and pc_count (when set) counts the preconditioner applications that ran
  const int* gate = nullptr;
  int* pc_count = nullptr;
  // island coarse space (fs_sys_set_deflation): Z = D^1/2 1_I per island I
  // in the equilibrated space, E = Z^T A_s Z
  int defl_n = 0;                  // islands (0: off)
  int defl_nchunks = 0;            // chunks of the island node list
  int32_t* defl_nodes = nullptr;   // island nodes (x-fast), grouped by island
  int32_t* defl_cstart = nullptr;  // chunk c covers defl_nodes[cstart[c] .. cstart[c+1])
  int32_t* defl_cisl = nullptr;    // island of chunk c
  int32_t* defl_ichunk = nullptr;  // chunks of island I: ichunk[I] .. ichunk[I+1]
  double* defl_einv = nullptr;     // E^-1 (defl_n x defl_n, row major)
  double* defl_part = nullptr;     // chunk partials [kMaxRhs][defl_nchunks]
  double* defl_coef = nullptr;     // [3][kMaxRhs][kMaxIslands]: E^-1 Z^T r, Z^T r, products
  // FP32 cycles of a deflated system run on the unscaled operator in
  // difference form: y_n = e_n u_n + sum_e w_e (u_n - u_m), u = D^-1/2 x_s
  float* u_w[3] = {nullptr, nullptr, nullptr};  // edge weights (FP32, x-fast)
  float* u_extra = nullptr;        // node diagonal (FP32) or null
  float* u_dinv = nullptr;         // 1/D on free rows, 0 elsewhere
};

struct fs_cap {
  fs_sys* sys = nullptr;
  int nt = 0;
  int nbins = 0;
  int nbx = 0;
  int has_shield = 0;
  int32_t* tid = nullptr;     // x-fast terminal map
  int64_t* nodes = nullptr;   // node lists per bin (x-fast indices)
  int64_t* offs = nullptr;
  double* out = nullptr;      // nbins x kMaxRhs charges
};
