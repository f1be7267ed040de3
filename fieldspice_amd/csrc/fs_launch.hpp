// Host-side launchers of the kernels in fs_kernels.hip.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fs_internal.hpp"

namespace fs {

// A launch whose start/stop events, when given, are timestamped at the
// kernel's own start and end (hipExtLaunchKernelGGL), so that the profiled
// duration excludes the dispatch gap an event pair around the launch holds.
// Either event may be null (a kernel split into several launches carries the
// start event on its first launch and the stop event on its last).
template <typename K, typename... A>
inline void launch_timed(K kernel, dim3 grid, hipStream_t st, hipEvent_t ea, hipEvent_t eb,
                         A... args) {
  if (ea || eb)
    hipExtLaunchKernelGGL(kernel, grid, dim3(kBlock), 0, st, ea, eb, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, dim3(kBlock), 0, st, args...);
}

// Profiling of one kernel kind of a Krylov iteration: the start and stop
// events of its launch (null without a slot; the start marks the kind as
// launched) and its device counters.
inline hipEvent_t prof_start(ProfSlot* sl, int kind) {
  if (!sl) return nullptr;
  sl->used |= 1u << kind;
  return sl->a[kind];
}
inline hipEvent_t prof_stop(ProfSlot* sl, int kind) { return sl ? sl->b[kind] : nullptr; }
inline unsigned long long* prof_ctr(const fs_sys* s, ProfSlot* sl, int kind) {
  return sl ? s->prof_dev + 2 * kind : nullptr;
}

template <typename T>
void permute_in(hipStream_t s, const T* in, T* out, int A, int B, int C, int OA, int OB);
template <typename T>
void permute_out(hipStream_t s, const T* in, T* out, int A, int B, int C, int IA, int IB);

void cell_to_edge(hipStream_t s, const double* prop, const double* const h[3],
                  const double* const hd[3], const int N[3], const Dims& d, int dir, int mode,
                  double* w);
void axpby(hipStream_t s, int64_t n, double a, const double* x, double b, const double* y,
           double* out);
void apply_raw(hipStream_t s, const fs_op* op, const double* x, double* y);

// returns 0 ok, 1 singular with pin == 0, -1 HIP failure
int setup_system(fs_sys* sys, double* dsum, void* counts_dev, void* counts_host, int pin);
constexpr size_t kSetupCountsBytes = 64;

// Operations of the single-block finalisation kernel (one block per RHS).
enum FinOp : int {
  FIN_ALPHA = 0,     // q = p.q          -> alpha = rho / (p.q)
  FIN_BETA = 1,      // q = r_new.r_new  -> beta, rho, stop tests
  FIN_INIT = 2,      // q = ||b_s||^2    -> fresh Krylov state
  FIN_BNORM = 3,     // q = ||b||^2      -> aux2
  FIN_RESID = 4,     // q0 = ||r||^2, q1 = ||r dhalf||^2 -> aux3, aux1
  FIN_CYCLE = 5,     // q = ||r32||^2    -> FP32 cycle state
  FIN_FCG = 6,       // q0 = <z,r>, q1 = <z,q> -> flexible-CG beta, zr
  FIN_FCG0 = 7,      // q0 = <z,r> at the start of a cycle -> zr
  FIN_ALPHA_ZR = 8,  // q = p.q          -> alpha = zr / (p.q) (preconditioned CG)
};
void launch_finalize(hipStream_t st, const double* partials, int op, int nr, int nb, int mode,
                     CgScalars* sc, double inner_tol2 = 0.0);

// mode 0 FP64 CG, 1 FP32 CG cycle, 2 FP32 flexible CG cycle (returns a status)
int launch_iteration(fs_sys* s, int nr, int mode, ProfSlot* slot);
// FP64 CG (mode 0) defers x += alpha p into the next direction update; this adds
// the last pending step, once, when the loop ends
void launch_flush_x(fs_sys* s, int nr);
// FP64 residual replacement: consume the flushed step on every RHS and
// restart the RHS whose bit is set in `restart` from r (rho = aux3, beta = 0)
void launch_replace(fs_sys* s, int nr, unsigned restart);
// nodes per lane of the level-0 cycle kernels of a preconditioner (1 or 4)
int pc_level0_width(const fs_pc* pc);
// 1: flexible CG stores this preconditioner's output in FP16 (4-wide level 0)
int pc_z_half(const fs_pc* pc);
void launch_init(fs_sys* s, int nr, double tol, int maxiter);
void launch_lift_general(fs_sys* s, int nr, const double* b);
void launch_lift_cap(fs_sys* s, int nr, const fs_op* L, const int32_t* tid, int j0);
void launch_expand(fs_sys* s, int nr, const double* b, const int32_t* tid, int j0,
                   double* phi);
void launch_residual(fs_sys* s, int nr, double* res);
void launch_cycle_start(fs_sys* s, int nr, double inner_tol2);
// flexible CG: z32 = M(r32) and <z,r> at the start of an FP32 cycle
int launch_fcg_start(fs_sys* s, int nr, ProfSlot* slot = nullptr);
void launch_refresh(fs_sys* s, int nr);

// island coarse space (fs_defl.hip).  Mode 0 of a system with defl_n > 0
// adds Z E^-1 Z^T r to each search direction (launch_defl_scatter after the
// p update, launch_defl_coef after the residual update); its FP32 cycles
// (mode 1) run launch_iteration_defl on the unscaled difference-form operator.
int launch_iteration_defl(fs_sys* s, int nr, int mode, ProfSlot* slot);
// Z^T r -> E^-1 Z^T r and zr (first: no beta, at the start of an iteration run)
void launch_defl_coef(fs_sys* s, int nr, int mode, int first);
void launch_defl_scatter(fs_sys* s, int nr, int mode);
void launch_cycle_start_defl(fs_sys* s, int nr, double inner_tol2);
void launch_refresh_defl(fs_sys* s, int nr);
// new rho, refresh count and stop tests after the FP64 residual of a refresh
void launch_refresh_finish(fs_sys* s, int nr);
void free_defl(fs_sys* s);
// one FP32 multi-RHS stencil launch (p32 -> q32) on tiling tl (autotuning);
// x4: the 4-nodes-per-lane kernel (tl in units of 256-node row segments)
// w: nodes per lane (1: the one-node-per-thread kernel)
void launch_stencil_only(fs_sys* s, int nr, const Tiling& tl, int w);
void launch_charge(fs_sys* s, int nr, const fs_op* L, const double* phi, const int64_t* nodes,
                   const int64_t* offs, int nbins, int nbx, double* out);

// device control of a batched solve: the gates that follow a chunk of a
// system's iterations (refresh, then cycle start and the finish stamp), the
// count of finished systems, and a device wall-clock stamp
void launch_batch_refresh_gate(fs_sys* s, int nr, BatchCtl* c);
void launch_batch_start_gate(fs_sys* s, int nr, int cycles, BatchCtl* c);
void launch_batch_count(hipStream_t st, const BatchCtl* c, int ni, int* out);
void launch_batch_clock(hipStream_t st, unsigned long long* t);

// time loops (fs_transient.hip)
void launch_explicit(fs_sys* s, const fs_op* A, const double* x, const double* src,
                     const double* scale, int nsrc, int64_t vs, double* rhs);
void launch_lift_groups(fs_sys* s, const double* rhs, const int32_t* group, const double* vals);
// start right-hand side `rhs` from a guess (node layout, or scaled)
void launch_init_warm(fs_sys* s, const double* xprev, const double* xprev2, double tol,
                      int maxiter, int scaled = 0, int rhs = 0);
// projection initial guesses of the time loops (basis of at most kProjMax)
void launch_proj_guess(fs_sys* s, const double* X, const double* AX, int64_t vs, int nb,
                       double* coef, double* x0, double tol, int maxiter);
void launch_proj_append(fs_sys* s, double* X, double* AX, int64_t vs, int nb, double* coef,
                        double* norm2);
void launch_diff(fs_sys* s, const double* a, const double* b, double* out);
void launch_expand_groups(fs_sys* s, const int32_t* group, const double* vals, double* x);
void launch_probe(fs_sys* s, const fs_op* A1, const fs_op* A2, const double* x,
                  const double* xprev, const double* w, const int64_t* nodes,
                  const int64_t* offs, int nprobe, int nbx, double* out_i, double* out_v);
void launch_step_stats(fs_sys* s, const double* x, const fs_op* Qa, const fs_op* Qb,
                       double* out);
void launch_edge_power(hipStream_t st, const fs_op* A, const double* x, double* q, int ncu);

// streaming probe: nread inputs and nwrite outputs of n floats (n % 4 == 0);
// supported mixes 1/1 and 3+k / k for k = 1..8; -1 otherwise
int launch_stream_probe(hipStream_t st, const float* const* in, float* const* out, int nread,
                        int nwrite, int64_t n, int nblocks);

// learned preconditioner (fs_pc.hip)
// With a profiling slot (and the system's counters), the level-0 restriction
// and prolongation launches are timed as FS_PROF_RESTRICT / FS_PROF_PROLONG.
int pc_cycle(fs_pc* pc, int nr, const float* r0, float* z0, const float* q, double* partials,
             int* nblocks, const CgScalars* sc, int mode, ProfSlot* slot = nullptr,
             unsigned long long* prof_dev = nullptr, int zhalf = 0);
// z32 = M(r32) with the FCG partials ([2*nr][*nblocks]); z32 holds FP16
// values (z / zscale) when pc_z_half(s->pc)
int pc_apply_solver(fs_sys* s, int nr, int* nblocks, ProfSlot* slot = nullptr);
// nodes of the level below level 0 (0 without a hierarchy)
int64_t pc_level1_nodes(const fs_pc* pc);
// the system of pc is being destroyed: pc keeps only its own memory
void pc_orphan(fs_pc* pc);
// solver workspace of a system for nr right-hand sides (fs_api.hip)
int ensure_solver_ws(fs_sys* s, int nr);

}  // namespace fs
