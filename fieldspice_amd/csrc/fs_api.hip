// C ABI of the fieldspice MI355X field solve (declared in include/fieldspice_hip.h).
//
// Every entry point validates its arguments, converts between fieldspice's
// z-fastest host layout and the x-fastest device layout, and reports failures
// through a status code plus a thread-local message.  No exception escapes.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <chrono>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <string>
#include <thread>
#include <vector>

#include "fieldspice_hip.h"
#include "fs_common.hpp"
#include "fs_internal.hpp"
#include "fs_launch.hpp"

namespace fs {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

const char* last_error() { return g_err.c_str(); }

}  // namespace fs

// Relative reduction an FP32 flexible-CG cycle aims for before the FP64
// refresh.  Measured at 256^3 (5 RHS): 1e-5 -> 169 iterations, 1e-4 -> 159,
// 1e-3 -> 128, 3e-3 -> 124, 1e-2 -> 131: deep FP32 cycles lose the
// A-orthogonality of their directions to rounding, a refresh restores it.
constexpr double kFcgInnerTol = 1e-3;

namespace {

using fs::dalloc;
using fs::dfree;
using fs::fail;

fs::Dims make_dims(int Nx, int Ny, int Nz) {
  fs::Dims d;
  d.nx = Nx + 1;
  d.ny = Ny + 1;
  d.nz = Nz + 1;
  d.sxy = (int64_t)d.nx * d.ny;
  d.n = d.sxy * d.nz;
  return d;
}

int check_dims(int Nx, int Ny, int Nz) {
  if (Nx < 1 || Ny < 1 || Nz < 1)
    return fail(FS_EINVAL, "cell counts must be >= 1, got (%d, %d, %d)", Nx, Ny, Nz);
  const int64_t n = (int64_t)(Nx + 1) * (Ny + 1) * (Nz + 1);
  if (n > ((int64_t)1 << 31) - 1)
    return fail(FS_EINVAL, "grid of %lld nodes exceeds the 2^31 node limit", (long long)n);
  return FS_OK;
}

// Upload a host array laid out C-order (A, B, C) into an x-fast device array
// with leading dims (OA, OB) (entries outside (A, B, C) are zero).
// `staging` (A*B*C elements, may be null: then allocated for the call) holds
// the host copy before the permutation.
template <typename T>
int upload_permuted(fs_ctx* ctx, const T* host, int A, int B, int C, int OA, int OB, int OC,
                    T* dev, T* staging = nullptr) {
  T* tmp = staging;
  const size_t cnt = (size_t)A * B * C;
  if (!tmp)
    if (int rc = dalloc(&tmp, cnt)) return rc;
  hipError_t e = hipMemcpyAsync(tmp, host, cnt * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemsetAsync(dev, 0, (size_t)OA * OB * OC * sizeof(T), ctx->stream);
  if (e == hipSuccess) {
    fs::permute_in<T>(ctx->stream, tmp, dev, A, B, C, OA, OB);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (!staging) dfree(tmp);
  if (e != hipSuccess) return fail(FS_EHIP, "upload failed: %s", hipGetErrorString(e));
  return FS_OK;
}

// Download an x-fast device node-shaped array to a host C-order array.
int download_nodes(fs_ctx* ctx, const fs::Dims& d, const double* dev, double* host,
                   double* tmp /* d.n doubles */) {
  fs::permute_out<double>(ctx->stream, dev, tmp, d.nx, d.ny, d.nz, d.nx, d.ny);
  FS_HIP(hipGetLastError());
  FS_HIP(hipMemcpyAsync(host, tmp, (size_t)d.n * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
  FS_HIP(hipStreamSynchronize(ctx->stream));
  return FS_OK;
}

fs_op* new_op(fs_ctx* ctx, int Nx, int Ny, int Nz, int per = 0) {
  fs_op* op = new fs_op();
  op->ctx = ctx;
  op->Nx = Nx;
  op->Ny = Ny;
  op->Nz = Nz;
  op->d = make_dims(Nx, Ny, Nz);
  op->d.per = per;
  const int cells[3] = {Nx, Ny, Nz};
  for (int a = 0; a < 3; ++a) op->ex[a] = ((per >> a) & 1) ? cells[a] + 1 : cells[a];
  return op;
}

void free_op(fs_op* op) {
  if (!op) return;
  for (auto& w : op->w) dfree(w);
  dfree(op->extra);
  delete op;
}

void free_sys_ws(fs_sys* s) {
  dfree(s->bs);
  dfree(s->x);
  dfree(s->r);
  dfree(s->p);
  dfree(s->q);
  dfree(s->x32);
  dfree(s->r32);
  dfree(s->p32);
  dfree(s->q32);
  dfree(s->z32);
  dfree(s->node_tmp);
  s->ws_nr = 0;
}

void free_sys(fs_sys* s) {
  if (!s) return;
  for (fs_pc* pc : s->pcs) fs::pc_orphan(pc);
  s->pcs.clear();
  fs::free_defl(s);
  free_sys_ws(s);
  dfree(s->mask);
  dfree(s->dhalf);
  dfree(s->dinv);
  for (auto& c : s->c64) dfree(c);
  for (auto& c : s->c32) dfree(c);
  dfree(s->partials);
  dfree(s->sc);
  dfree(s->prof_dev);
  if (s->sc_host) (void)hipHostFree(s->sc_host);
  for (int c = 0; c < s->ev_cap; ++c)
    for (int k = 0; k < FS_PROF_KINDS; ++k) {
      (void)hipEventDestroy(s->prof_slots[c].a[k]);
      (void)hipEventDestroy(s->prof_slots[c].b[k]);
    }
  delete[] s->prof_slots;
  delete s;
}

constexpr int kMaxChunk = 512;

void prof_begin(fs_sys* s) {
  for (int k = 0; k < FS_PROF_KINDS; ++k) {
    s->prof_time[k] = 0.0;
    s->prof_launches[k] = 0;
  }
  if (s->prof_dev)
    (void)hipMemsetAsync(s->prof_dev, 0, 2 * FS_PROF_KINDS * sizeof(unsigned long long),
                         s->ctx->stream);
}

// Algorithmic bytes of each profiled kernel kind from its launches with work
// (L) and live right-hand sides summed (R), n nodes, elt bytes per Krylov
// value (the cycle kernels are FP32), n1 nodes on level 1:
//   stencil   3 couplings once, p read and q written per RHS
//   direction dir, p and x read, p and x written per RHS
//   residual  r and q read, r written per RHS
//   restrict  3 couplings once; r read and the level-1 residual written per RHS
//   prolong   3 couplings and the mask once; r and q read, z written and the
//             level-1 error read per RHS
// (flexible CG over a 4-wide cycle: z is FP16, 2 bytes written and read)
void prof_end(fs_sys* s, fs_solve_stats* st, int precision) {
  if (!st || !s->profiling || !s->prof_dev) return;
  unsigned long long h[2 * FS_PROF_KINDS] = {};
  if (hipMemcpy(h, s->prof_dev, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return;
  const bool mixed = precision != FS_PREC_FP64;
  const double zb = precision == FS_PREC_FCG && fs::pc_z_half(s->pc) ? 2.0 : 4.0;
  const double n = (double)s->d.n, elt = mixed ? 4.0 : 8.0;
  const double n1 = (double)fs::pc_level1_nodes(s->pc);
  const double fixed[FS_PROF_KINDS] = {3.0 * elt * n, 0.0, 0.0, 12.0 * n, 13.0 * n};
  const double per_rhs[FS_PROF_KINDS] = {2.0 * elt * n, (mixed ? 16.0 + zb : 40.0) * n,
                                         3.0 * elt * n, 4.0 * n + 4.0 * n1,
                                         (8.0 + zb) * n + 4.0 * n1};
  for (int k = 0; k < FS_PROF_KINDS; ++k) {
    const double bytes = fixed[k] * (double)h[2 * k] + per_rhs[k] * (double)h[2 * k + 1];
    st->kernel_time[k] += s->prof_time[k];
    st->kernel_launches[k] += s->prof_launches[k];
    st->kernel_bytes[k] += bytes;
  }
  st->stencil_time += s->prof_time[FS_PROF_STENCIL];
  st->stencil_launches += s->prof_launches[FS_PROF_STENCIL];
  st->stencil_bytes += fixed[FS_PROF_STENCIL] * (double)h[0] + per_rhs[FS_PROF_STENCIL] * (double)h[1];
}

int ensure_ws(fs_sys* s, int nr) {
  if (s->ws_nr >= nr) return FS_OK;
  free_sys_ws(s);
  const size_t cnt = (size_t)nr * s->vstride;
  int rc;
  if ((rc = dalloc(&s->bs, cnt)) || (rc = dalloc(&s->x, cnt)) || (rc = dalloc(&s->r, cnt)) ||
      (rc = dalloc(&s->p, cnt)) || (rc = dalloc(&s->q, cnt)) || (rc = dalloc(&s->x32, cnt)) ||
      (rc = dalloc(&s->r32, cnt)) || (rc = dalloc(&s->p32, cnt)) ||
      (rc = dalloc(&s->q32, cnt)) || (rc = dalloc(&s->z32, cnt)) ||
      (rc = dalloc(&s->node_tmp, cnt))) {
    free_sys_ws(s);
    return rc;
  }
  // the flexible-CG start reads q32 into an unused partial: keep it finite
  if (hipMemset(s->q32, 0, cnt * sizeof(float)) != hipSuccess ||
      hipMemset(s->z32, 0, cnt * sizeof(float)) != hipSuccess) {
    free_sys_ws(s);
    return fail(FS_EHIP, "workspace initialisation failed");
  }
  s->ws_nr = nr;
  return FS_OK;
}

int default_maxiter(const fs_sys* s, const fs_solve_params* P) {
  if (P && P->maxiter > 0) return P->maxiter;
  const double nf = (double)std::max<int64_t>(s->info.n_free, 1);
  return std::max(500, (int)std::floor(10.0 * std::sqrt(nf)));
}

// Krylov loop after init: FP64 CG (mode 0), or FP32 cycles with FP64
// refreshes: CG (mode 1) or flexible CG preconditioned by the attached fs_pc
// (mode 2).  Host checks the device flags between chunks of iterations.
static int run_krylov_loop(fs_sys* s, int nr, const fs_solve_params& P, int mode);

// Residual replacements FP64 CG may make per solve.  A restart that does not
// halve the true residual means it sits at the rounding floor of the
// problem (a solution dominated by a near-null mode); the solve then stops
// and the caller sees the true residual.
constexpr int kMaxReplace = 4;

// End of an FP64 CG run: add the pending x += alpha p (deferred into the next
// direction update, which a finished right-hand side never runs), recompute
// b_s - A_s x in FP64, and restart from that residual every right-hand side
// whose recursive residual met tol while the true one did not (the
// recursion drifts from b_s - A_s x by rounding).
int fp64_finish(fs_sys* s, int nr, const fs_solve_params& P) {
  fs::launch_flush_x(s, nr);
  FS_HIP(hipGetLastError());
  const size_t scb = sizeof(fs::CgScalars) * (size_t)nr;
  std::vector<double> prev(nr, INFINITY);
  for (int rep = 0; rep <= kMaxReplace; ++rep) {
    fs::launch_residual(s, nr, s->r);
    FS_HIP(hipGetLastError());
    FS_HIP(hipMemcpyAsync(s->sc_host, s->sc, scb, hipMemcpyDeviceToHost, s->stream));
    FS_HIP(hipStreamSynchronize(s->stream));
    unsigned restart = 0;
    for (int r = 0; r < nr; ++r) {
      const fs::CgScalars& c = s->sc_host[r];
      // (thresh2 == 0: a zero right-hand side, already exact)
      if (rep < kMaxReplace && c.done == 1 && std::isfinite(c.aux3) && c.thresh2 > 0.0 &&
          c.aux3 >= c.thresh2 && c.aux3 < 0.25 * prev[r] && c.iter < c.maxiter)
        restart |= 1u << r;
      prev[r] = c.aux3;
    }
    fs::launch_replace(s, nr, restart);
    FS_HIP(hipGetLastError());
    if (!restart) return FS_OK;
    s->acc_replacements += __builtin_popcount(restart);
    if (int rc = run_krylov_loop(s, nr, P, 0)) return rc;
    fs::launch_flush_x(s, nr);
    FS_HIP(hipGetLastError());
  }
  return FS_OK;
}

int run_krylov_mode(fs_sys* s, int nr, const fs_solve_params& P, int mode) {
  const int rc = run_krylov_loop(s, nr, P, mode);
  if (mode != 0 || rc != FS_OK) {
    if (mode == 0) fs::launch_flush_x(s, nr);
    return rc;
  }
  return fp64_finish(s, nr, P);
}

static int run_krylov_loop(fs_sys* s, int nr, const fs_solve_params& P, int mode) {
  hipStream_t st = s->ctx->stream;
  const bool cyc = mode != 0;
  const double inner_tol = P.inner_tol > 0 ? P.inner_tol : (mode == 2 ? kFcgInnerTol : 1e-4);
  const size_t scb = sizeof(fs::CgScalars) * (size_t)nr;
  auto fetch = [&]() -> int {
    FS_HIP(hipMemcpyAsync(s->sc_host, s->sc, scb, hipMemcpyDeviceToHost, st));
    FS_HIP(hipStreamSynchronize(st));
    return FS_OK;
  };
  if (int rc = fetch()) return rc;
  auto all = [&](auto pred) {
    for (int r = 0; r < nr; ++r)
      if (!pred(s->sc_host[r])) return false;
    return true;
  };
  if (all([](const fs::CgScalars& c) { return c.done != 0; })) return FS_OK;
  // profiling: the first preconditioner application of each flexible-CG
  // cycle is timed in the slot after the chunk's
  fs::ProfSlot* start_sl = s->profiling && s->prof_slots ? &s->prof_slots[kMaxChunk] : nullptr;
  auto start_cycle = [&]() -> int {
    fs::launch_cycle_start(s, nr, inner_tol * inner_tol);
    if (mode == 2) {
      if (start_sl) start_sl->used = 0;
      if (int rc = fs::launch_fcg_start(s, nr, start_sl)) return rc;
      s->acc_pc_applies += 1;
    }
    FS_HIP(hipGetLastError());
    return FS_OK;
  };
  auto add_times = [&](fs::ProfSlot& sl) {
    for (int k = 0; k < FS_PROF_KINDS; ++k) {
      if (!((sl.used >> k) & 1u)) continue;
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, sl.a[k], sl.b[k]) == hipSuccess)
        s->prof_time[k] += 1e-3 * ms;
      else
        (void)hipGetLastError();  // an unrecorded pair is skipped, not left as a sticky error
      s->prof_launches[k] += 1;
    }
    sl.used = 0;
  };
  if (cyc) {
    if (int rc = start_cycle()) return rc;
  } else if (s->defl_n > 0) {
    fs::launch_defl_coef(s, nr, 0, 1);  // first search direction (I + Z E^-1 Z^T) r
    FS_HIP(hipGetLastError());
  }
  // a flexible-CG iteration costs several stencil applications: check often
  const int max_chunk = mode == 2 ? 32 : kMaxChunk;
  const int first_chunk = mode == 2 ? 4 : 16;
  int chunk = P.check_every > 0 ? P.check_every : first_chunk;
  const int maxiter = s->sc_host[0].maxiter;
  std::vector<double> last_rho(nr, 0.0);
  std::vector<int> last_it(nr, 0);
  for (int guard = 0; guard < 4 * maxiter + 64; ++guard) {
    const bool prof = s->profiling && chunk <= s->ev_cap;
    for (int c = 0; c < chunk; ++c) {
      fs::ProfSlot* sl = prof ? &s->prof_slots[c] : nullptr;
      if (sl) sl->used = 0;
      if (int rc = fs::launch_iteration(s, nr, mode, sl)) return rc;
      if (mode == 2) s->acc_pc_applies += 1;
    }
    FS_HIP(hipGetLastError());
    if (int rc = fetch()) return rc;
    if (prof)
      for (int c = 0; c < chunk; ++c) add_times(s->prof_slots[c]);
    if (start_sl && start_sl->used) add_times(*start_sl);
    if (!cyc) {
      if (all([](const fs::CgScalars& c) { return c.done != 0; })) return FS_OK;
    } else {
      if (all([](const fs::CgScalars& c) { return c.done != 0 || c.cycle_done != 0; })) {
        fs::launch_refresh(s, nr);
        FS_HIP(hipGetLastError());
        if (int rc = fetch()) return rc;
        if (all([](const fs::CgScalars& c) { return c.done != 0; })) return FS_OK;
        if (int rc = start_cycle()) return rc;
        chunk = P.check_every > 0 ? P.check_every : first_chunk;
        std::fill(last_it.begin(), last_it.end(), 0);
        continue;
      }
    }
    if (P.check_every > 0) continue;
    // Adaptive chunk: extrapolate the residual decay of the slowest live RHS
    // to its threshold (CG aims slightly past it: its idle launches exit at
    // once; a flexible-CG iteration past the target is not free).
    int want = mode == 2 ? 1 : 8;
    for (int r = 0; r < nr; ++r) {
      const fs::CgScalars& c = s->sc_host[r];
      if (c.done != 0 || (cyc && c.cycle_done)) continue;
      const double rho = cyc ? c.aux0 : c.rho;
      const double target = cyc ? c.aux1 : c.thresh2;
      const int it = cyc ? c.cycle_iter : c.iter;
      if (last_it[r] > 0 && it > last_it[r] && rho > 0 && last_rho[r] > rho) {
        const double rate = std::log(rho / last_rho[r]) / (double)(it - last_it[r]);
        const double remain = std::log(target / rho) / rate;
        if (std::isfinite(remain))
          want = std::max(want, (int)std::ceil(remain) + (mode == 2 ? 0 : 2));
      } else {
        want = std::max(want, chunk * 2);
      }
      last_rho[r] = rho;
      last_it[r] = it;
    }
    chunk = std::min(std::max(want, mode == 2 ? 1 : 4), max_chunk);
  }
  return fail(FS_EINVAL, "Krylov driver exceeded its loop guard");
}

int run_krylov(fs_sys* s, int nr, const fs_solve_params& P) {
  if (P.precision != FS_PREC_FCG)
    return run_krylov_mode(s, nr, P, P.precision == FS_PREC_MIXED ? 1 : 0);
  if (!s->pc) return fail(FS_EINVAL, "FS_PREC_FCG needs a preconditioner (fs_sys_set_precond)");
  if (s->defl_n > 0)
    return fail(FS_EINVAL, "FS_PREC_FCG does not combine with an island space; use "
                           "FS_PREC_FP64 or FS_PREC_MIXED");
  if (int rc = run_krylov_mode(s, nr, P, 2)) return rc;
  // plain FP64 CG redoes the batch when flexible CG broke down or stalled
  bool ok = true;
  for (int r = 0; r < nr; ++r) ok = ok && s->sc_host[r].done == 1;
  if (ok) return FS_OK;
  s->acc_fallbacks += 1;
  // the fallback keeps the reference's own cap, whatever maxiter bounded the
  // flexible-CG run
  fs_solve_params Q = P;
  Q.precision = FS_PREC_FP64;
  Q.maxiter = 0;
  fs::launch_init(s, nr, P.tol, default_maxiter(s, &Q));
  FS_HIP(hipGetLastError());
  return run_krylov_mode(s, nr, Q, 0);
}

// which kernels a solve in `precision` runs on this system (fs_solve_stats)
void kernel_widths(const fs_sys* s, int precision, fs_solve_stats* st) {
  if (!st) return;
  st->stencil_width = precision == FS_PREC_FP64 ? (s->use_x2 ? 2 : 1) : (s->xw ? s->xw : 1);
  st->cycle_width = precision == FS_PREC_FCG && s->pc ? fs::pc_level0_width(s->pc) : 0;
}

void collect_stats(fs_sys* s, int nr, fs_solve_stats* st, int& status) {
  if (st) {
    st->fallbacks += s->acc_fallbacks;
    st->pc_applies += s->acc_pc_applies;
    st->replacements += s->acc_replacements;
  }
  s->acc_fallbacks = 0;
  s->acc_pc_applies = 0;
  s->acc_replacements = 0;
  for (int r = 0; r < nr; ++r) {
    const fs::CgScalars& c = s->sc_host[r];
    if (st) {
      st->iterations = std::max(st->iterations, c.iter);
      st->refreshes = std::max(st->refreshes, c.refreshes);
      st->stop_reason = std::max(st->stop_reason, c.done);
      st->maxiter = std::max(st->maxiter, c.maxiter);
      double res = c.aux2 > 0 ? std::sqrt(c.aux1 / c.aux2) : 0.0;
      double req = c.bnorm2 > 0 ? std::sqrt(c.aux3 / c.bnorm2) : 0.0;
      // a non-finite residual (NaN input, breakdown) must not read as converged
      if (!(res >= 0.0) || !std::isfinite(c.aux2)) res = INFINITY;
      if (!(req >= 0.0) || !std::isfinite(c.bnorm2)) req = INFINITY;
      st->residual = std::max(st->residual, res);
      st->residual_eq = std::max(st->residual_eq, req);
    }
    if (c.done != 1) status = FS_NOT_CONVERGED;
  }
}

}  // namespace

namespace fs {
int ensure_solver_ws(fs_sys* s, int nr) { return ensure_ws(s, nr); }
}  // namespace fs

namespace {
// per-probe node lists in device order (host C-order ids -> x-fast indices)
int build_lists(const fs::Dims& d, const int32_t* id, int nbins, std::vector<int64_t>& nodes,
                std::vector<int64_t>& offs) {
  std::vector<int64_t> counts(nbins, 0);
  for (int64_t ref = 0; ref < d.n; ++ref) {
    const int t = id[ref];
    if (t < -1 || t >= nbins) return fail(FS_EINVAL, "id[%lld] = %d out of range", (long long)ref, t);
    if (t >= 0) counts[t] += 1;
  }
  offs.assign(nbins + 1, 0);
  for (int t = 0; t < nbins; ++t) offs[t + 1] = offs[t] + counts[t];
  nodes.assign((size_t)std::max<int64_t>(offs[nbins], 1), 0);
  std::vector<int64_t> fill(offs.begin(), offs.end() - 1);
  for (int64_t i = 0; i < d.nx; ++i)
    for (int64_t j = 0; j < d.ny; ++j)
      for (int64_t k = 0; k < d.nz; ++k) {
        const int t = id[(i * d.ny + j) * d.nz + k];
        if (t >= 0) nodes[fill[t]++] = (k * d.ny + j) * d.nx + i;
      }
  return FS_OK;
}

struct DevBufs {
  std::vector<void*> ptrs;
  ~DevBufs() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int get(T** p, size_t count) {
    int rc = dalloc(p, count);
    if (rc == FS_OK) ptrs.push_back((void*)*p);
    return rc;
  }
};
}  // namespace

// --------------------------------------------------------------------------
// Scratch of fs_sys_solve_noalloc: one array per GPU, reserved in device
// memory when the library's code object is loaded on that GPU and held until
// the process ends.  Its name is a host-side stand-in: host code uses only
// the device address hipGetSymbolAddress returns for it on each device.
constexpr size_t kScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current (fs_ctx_create): the address of
// this device's copy of `scratch`.
int scratch_setup(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices)
    return fail(FS_EINVAL, "device %d beyond the %d the scratch table covers", device,
                kMaxDevices);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FS_EHIP, "scratch address on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[device];
  return FS_OK;
}

// The solve workspace of a system carved out of the scratch array for one
// call; the system's own workspace pointers come back when it goes out of
// scope.
struct ScratchLoan {
  fs_sys* s;
  double *bs, *x, *r, *p, *q, *node_tmp;
  float *x32, *r32, *p32, *q32, *z32;
  int ws_nr;
  double* x0d = nullptr;  // initial guesses (nr x vstride), if asked for

  explicit ScratchLoan(fs_sys* sys)
      : s(sys), bs(sys->bs), x(sys->x), r(sys->r), p(sys->p), q(sys->q),
        node_tmp(sys->node_tmp), x32(sys->x32), r32(sys->r32), p32(sys->p32), q32(sys->q32),
        z32(sys->z32), ws_nr(sys->ws_nr) {}
  ~ScratchLoan() {
    s->bs = bs; s->x = x; s->r = r; s->p = p; s->q = q; s->node_tmp = node_tmp;
    s->x32 = x32; s->r32 = r32; s->p32 = p32; s->q32 = q32; s->z32 = z32;
    s->ws_nr = ws_nr;
  }

  // Bytes the workspace of nr right-hand sides (plus guesses) takes.
  static size_t need(const fs_sys* sys, int nr, bool with_x0) {
    const size_t vec = ((size_t)nr * sys->vstride * sizeof(double) + 255) / 256 * 256;
    const size_t vecf = ((size_t)nr * sys->vstride * sizeof(float) + 255) / 256 * 256;
    return (6 + (with_x0 ? 1 : 0)) * vec + 5 * vecf;
  }

  // Point the system's workspace at `base` (a device address).
  int lend(unsigned char* base, int nr, bool with_x0) {
    const size_t vec = ((size_t)nr * s->vstride * sizeof(double) + 255) / 256 * 256;
    const size_t vecf = ((size_t)nr * s->vstride * sizeof(float) + 255) / 256 * 256;
    unsigned char* c = base;
    auto d = [&]() { double* v = (double*)c; c += vec; return v; };
    auto f = [&]() { float* v = (float*)c; c += vecf; return v; };
    s->bs = d(); s->x = d(); s->r = d(); s->p = d(); s->q = d(); s->node_tmp = d();
    s->x32 = f(); s->r32 = f(); s->p32 = f(); s->q32 = f(); s->z32 = f();
    if (with_x0) x0d = d();
    s->ws_nr = nr;
    // the array keeps the previous call's contents: the flexible-CG start
    // reads q32 into an unused partial, and z32 must start finite
    const size_t fbytes = (size_t)nr * s->vstride * sizeof(float);
    FS_HIP(hipMemsetAsync(s->q32, 0, fbytes, s->ctx->stream));
    FS_HIP(hipMemsetAsync(s->z32, 0, fbytes, s->ctx->stream));
    return FS_OK;
  }
};

// fs_sys_solve; with `loan` every device buffer of the call comes from it
int sys_solve(fs_sys* s, int nrhs, const double* b, const double* x0, double* x,
              const fs_solve_params* params, fs_solve_stats* stats, ScratchLoan* loan);

}  // namespace

// ==========================================================================
extern "C" {

const char* fs_last_error(void) { return fs::last_error(); }

int fs_ctx_create(int device, fs_ctx** out) {
  FS_GUARD_BEGIN
  if (!out) return fail(FS_EINVAL, "out is NULL");
  *out = nullptr;
  int n = 0;
  FS_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return fail(FS_EINVAL, "device %d out of range (found %d)", device, n);
  FS_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  FS_HIP(hipGetDeviceProperties(&prop, device));
  fs_ctx* ctx = new fs_ctx();
  ctx->device = device;
  ctx->n_cu = prop.multiProcessorCount;
  snprintf(ctx->name, sizeof(ctx->name), "%s", prop.gcnArchName);
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    delete ctx;
    return fail(FS_EHIP, "stream/event creation failed");
  }
  // the no-allocation solve's scratch on this device (its failure is
  // reported by fs_sys_solve_noalloc, not here)
  const std::string keep = fs::g_err;
  if (scratch_setup(device, &ctx->scratch) != FS_OK) ctx->scratch = nullptr;
  fs::g_err = keep;
  *out = ctx;
  return FS_OK;
  FS_GUARD_END
}

int fs_ctx_destroy(fs_ctx* ctx) {
  if (!ctx) return FS_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FS_OK;
}

int fs_ctx_device_info(fs_ctx* ctx, char* name, int name_len, int* n_cu) {
  if (!ctx) return fail(FS_EINVAL, "ctx is NULL");
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", ctx->name);
  if (n_cu) *n_cu = ctx->n_cu;
  return FS_OK;
}

int fs_stream_probe(fs_ctx* ctx, int64_t n, int nread, int nwrite, int reps, double* seconds) {
  FS_GUARD_BEGIN
  if (!ctx || !seconds) return fail(FS_EINVAL, "NULL argument");
  if (n <= 0 || n % 4 != 0 || reps <= 0)
    return fail(FS_EINVAL, "n must be a positive multiple of 4 and reps positive");
  const bool mix = (nread == 1 && nwrite == 1) || (nwrite >= 1 && nwrite <= 8 && nread == 3 + nwrite);
  if (!mix) return fail(FS_EINVAL, "unsupported mix: %d read, %d written", nread, nwrite);
  FS_HIP(hipSetDevice(ctx->device));
  std::vector<float*> bufs(nread + nwrite, nullptr);
  int rc = FS_OK;
  for (auto& b : bufs)
    if ((rc = dalloc(&b, (size_t)n))) break;
  auto release = [&] {
    for (auto& b : bufs) dfree(b);
  };
  if (rc) {
    release();
    return rc;
  }
  for (int k = 0; k < nread; ++k) {
    if (hipMemsetAsync(bufs[k], 0, (size_t)n * sizeof(float), ctx->stream) != hipSuccess) {
      release();
      return fail(FS_EHIP, "probe buffer initialisation failed");
    }
  }
  const int nblocks = std::max(1, ctx->n_cu) * 8;
  const float* const* in = (const float* const*)bufs.data();
  float* const* out = bufs.data() + nread;
  fs::launch_stream_probe(ctx->stream, in, out, nread, nwrite, n, nblocks);  // warm-up
  (void)hipEventRecord(ctx->ev0, ctx->stream);
  for (int r = 0; r < reps; ++r)
    fs::launch_stream_probe(ctx->stream, in, out, nread, nwrite, n, nblocks);
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  const hipError_t e = hipEventSynchronize(ctx->ev1);
  float ms = 0.0f;
  if (e == hipSuccess) (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  const hipError_t le = hipGetLastError();
  release();
  if (e != hipSuccess || le != hipSuccess)
    return fail(FS_EHIP, "streaming probe failed: %s",
                hipGetErrorString(e != hipSuccess ? e : le));
  *seconds = 1e-3 * (double)ms / reps;
  return FS_OK;
  FS_GUARD_END
}

int fs_op_create_edges(fs_ctx* ctx, int Nx, int Ny, int Nz, const double* w_x,
                       const double* w_y, const double* w_z, const double* extra,
                       fs_op** out) {
  FS_GUARD_BEGIN
  if (!ctx || !out || !w_x || !w_y || !w_z) return fail(FS_EINVAL, "NULL argument");
  *out = nullptr;
  if (int rc = check_dims(Nx, Ny, Nz)) return rc;
  FS_HIP(hipSetDevice(ctx->device));
  fs_op* op = new_op(ctx, Nx, Ny, Nz);
  const fs::Dims& d = op->d;
  int rc = FS_OK;
  for (int a = 0; a < 3 && rc == FS_OK; ++a) rc = dalloc(&op->w[a], (size_t)d.n);
  if (rc == FS_OK) rc = upload_permuted(ctx, w_x, Nx, d.ny, d.nz, d.nx, d.ny, d.nz, op->w[0]);
  if (rc == FS_OK) rc = upload_permuted(ctx, w_y, d.nx, Ny, d.nz, d.nx, d.ny, d.nz, op->w[1]);
  if (rc == FS_OK) rc = upload_permuted(ctx, w_z, d.nx, d.ny, Nz, d.nx, d.ny, d.nz, op->w[2]);
  if (rc == FS_OK && extra) {
    rc = dalloc(&op->extra, (size_t)d.n);
    if (rc == FS_OK)
      rc = upload_permuted(ctx, extra, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz, op->extra);
  }
  if (rc != FS_OK) {
    free_op(op);
    return rc;
  }
  *out = op;
  return FS_OK;
  FS_GUARD_END
}

int fs_op_create_periodic(fs_ctx* ctx, int nx, int ny, int nz, const double* w_x,
                          const double* w_y, const double* w_z, int periodic, fs_op** out) {
  FS_GUARD_BEGIN
  if (!ctx || !out || !w_x || !w_y || !w_z) return fail(FS_EINVAL, "NULL argument");
  *out = nullptr;
  if (periodic < 0 || periodic > 7) return fail(FS_EINVAL, "periodic must be a mask of bits 0..2");
  const int nodes[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a)
    if (((periodic >> a) & 1) && nodes[a] < 2)
      return fail(FS_EINVAL, "a periodic axis needs at least 2 nodes (axis %d has %d)", a,
                  nodes[a]);
  if (int rc = check_dims(nx - 1, ny - 1, nz - 1)) return rc;
  FS_HIP(hipSetDevice(ctx->device));
  fs_op* op = new_op(ctx, nx - 1, ny - 1, nz - 1, periodic);
  const fs::Dims& d = op->d;
  const double* src[3] = {w_x, w_y, w_z};
  int rc = FS_OK;
  for (int a = 0; a < 3 && rc == FS_OK; ++a) {
    rc = dalloc(&op->w[a], (size_t)d.n);
    int shp[3] = {d.nx, d.ny, d.nz};
    shp[a] = op->ex[a];
    if (rc == FS_OK)
      rc = upload_permuted(ctx, src[a], shp[0], shp[1], shp[2], d.nx, d.ny, d.nz, op->w[a]);
  }
  if (rc != FS_OK) {
    free_op(op);
    return rc;
  }
  *out = op;
  return FS_OK;
  FS_GUARD_END
}

int fs_op_create_cells(fs_ctx* ctx, int Nx, int Ny, int Nz, const double* hx,
                       const double* hy, const double* hz, const double* prop_cell,
                       int avg_mode, fs_op** out) {
  FS_GUARD_BEGIN
  if (!ctx || !out || !hx || !hy || !hz || !prop_cell) return fail(FS_EINVAL, "NULL argument");
  *out = nullptr;
  if (int rc = check_dims(Nx, Ny, Nz)) return rc;
  if (avg_mode < FS_AVG_PARALLEL || avg_mode > FS_AVG_MAX)
    return fail(FS_EINVAL, "unknown averaging mode %d", avg_mode);
  const int N[3] = {Nx, Ny, Nz};
  const double* hh[3] = {hx, hy, hz};
  std::vector<double> host_h[3], host_hd[3];
  for (int a = 0; a < 3; ++a) {
    host_h[a].assign(hh[a], hh[a] + N[a]);
    host_hd[a].resize((size_t)N[a] + 1);
    for (int i = 0; i < N[a]; ++i)
      if (!(host_h[a][i] > 0.0) || !std::isfinite(host_h[a][i]))
        return fail(FS_EINVAL, "cell widths along axis %d must be finite and > 0", a);
    // dual widths: half cells at the two ends (fieldspice/grid.py:234-238)
    host_hd[a][0] = 0.5 * host_h[a][0];
    host_hd[a][N[a]] = 0.5 * host_h[a][N[a] - 1];
    for (int i = 1; i < N[a]; ++i) host_hd[a][i] = 0.5 * (host_h[a][i - 1] + host_h[a][i]);
  }
  FS_HIP(hipSetDevice(ctx->device));
  fs_op* op = new_op(ctx, Nx, Ny, Nz);
  const fs::Dims& d = op->d;
  double *dh[3] = {nullptr, nullptr, nullptr}, *dhd[3] = {nullptr, nullptr, nullptr};
  double* cells = nullptr;
  int rc = FS_OK;
  for (int a = 0; a < 3 && rc == FS_OK; ++a) {
    rc = dalloc(&op->w[a], (size_t)d.n);
    if (rc == FS_OK) rc = dalloc(&dh[a], (size_t)N[a]);
    if (rc == FS_OK) rc = dalloc(&dhd[a], (size_t)N[a] + 1);
    if (rc == FS_OK &&
        (hipMemcpy(dh[a], host_h[a].data(), N[a] * sizeof(double), hipMemcpyHostToDevice) !=
             hipSuccess ||
         hipMemcpy(dhd[a], host_hd[a].data(), (N[a] + 1) * sizeof(double),
                   hipMemcpyHostToDevice) != hipSuccess))
      rc = fail(FS_EHIP, "metric upload failed");
  }
  if (rc == FS_OK) rc = dalloc(&cells, (size_t)Nx * Ny * Nz);
  if (rc == FS_OK) rc = upload_permuted(ctx, prop_cell, Nx, Ny, Nz, Nx, Ny, Nz, cells);
  if (rc == FS_OK) {
    const double* ch[3] = {dh[0], dh[1], dh[2]};
    const double* chd[3] = {dhd[0], dhd[1], dhd[2]};
    for (int a = 0; a < 3; ++a) fs::cell_to_edge(ctx->stream, cells, ch, chd, N, d, a, avg_mode, op->w[a]);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(FS_EHIP, "cell_to_edge kernel failed");
  }
  for (int a = 0; a < 3; ++a) {
    dfree(dh[a]);
    dfree(dhd[a]);
  }
  dfree(cells);
  if (rc != FS_OK) {
    free_op(op);
    return rc;
  }
  *out = op;
  return FS_OK;
  FS_GUARD_END
}

int fs_op_combine(double a, const fs_op* A, double b, const fs_op* B, double c,
                  const double* extra, fs_op** out) {
  FS_GUARD_BEGIN
  if (!A || !out) return fail(FS_EINVAL, "NULL argument");
  *out = nullptr;
  if (B && (B->Nx != A->Nx || B->Ny != A->Ny || B->Nz != A->Nz || B->ctx != A->ctx ||
            B->d.per != A->d.per))
    return fail(FS_EINVAL, "operators differ in grid, periodicity or context");
  fs_ctx* ctx = A->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  fs_op* op = new_op(ctx, A->Nx, A->Ny, A->Nz, A->d.per);
  const int64_t n = op->d.n;
  int rc = FS_OK;
  for (int k = 0; k < 3 && rc == FS_OK; ++k) {
    rc = dalloc(&op->w[k], (size_t)n);
    if (rc == FS_OK) fs::axpby(ctx->stream, n, a, A->w[k], b, B ? B->w[k] : nullptr, op->w[k]);
  }
  const bool any_extra = A->extra || (B && B->extra) || extra;
  if (rc == FS_OK && any_extra) {
    rc = dalloc(&op->extra, (size_t)n);
    if (rc == FS_OK) {
      FS_HIP(hipMemsetAsync(op->extra, 0, n * sizeof(double), ctx->stream));
      if (A->extra) fs::axpby(ctx->stream, n, a, A->extra, 1.0, op->extra, op->extra);
      if (B && B->extra) fs::axpby(ctx->stream, n, b, B->extra, 1.0, op->extra, op->extra);
      if (extra) {
        double* tmp = nullptr;
        rc = dalloc(&tmp, (size_t)n);
        if (rc == FS_OK)
          rc = upload_permuted(ctx, extra, op->d.nx, op->d.ny, op->d.nz, op->d.nx, op->d.ny,
                               op->d.nz, tmp);
        if (rc == FS_OK) fs::axpby(ctx->stream, n, c, tmp, 1.0, op->extra, op->extra);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(FS_EHIP, "combine failed");
        dfree(tmp);
      }
    }
  }
  if (rc == FS_OK && (hipGetLastError() != hipSuccess ||
                      hipStreamSynchronize(ctx->stream) != hipSuccess))
    rc = fail(FS_EHIP, "combine kernels failed");
  if (rc != FS_OK) {
    free_op(op);
    return rc;
  }
  *out = op;
  return FS_OK;
  FS_GUARD_END
}

int fs_op_destroy(fs_op* op) {
  if (op) (void)hipSetDevice(op->ctx->device);
  free_op(op);
  return FS_OK;
}

int fs_op_apply(fs_op* op, const double* x, double* y) {
  FS_GUARD_BEGIN
  if (!op || !x || !y) return fail(FS_EINVAL, "NULL argument");
  fs_ctx* ctx = op->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  const fs::Dims& d = op->d;
  double *dx = nullptr, *dy = nullptr;
  int rc = dalloc(&dx, (size_t)d.n);
  if (rc == FS_OK) rc = dalloc(&dy, (size_t)d.n);
  if (rc == FS_OK) rc = upload_permuted(ctx, x, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz, dx);
  if (rc == FS_OK) {
    fs::apply_raw(ctx->stream, op, dx, dy);
    rc = download_nodes(ctx, d, dy, y, dx);
  }
  dfree(dx);
  dfree(dy);
  return rc;
  FS_GUARD_END
}

int fs_op_get_edges(fs_op* op, double* w_x, double* w_y, double* w_z) {
  FS_GUARD_BEGIN
  if (!op || !w_x || !w_y || !w_z) return fail(FS_EINVAL, "NULL argument");
  fs_ctx* ctx = op->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  const fs::Dims& d = op->d;
  double* tmp = nullptr;
  if (int rc = dalloc(&tmp, (size_t)d.n)) return rc;
  double* outs[3] = {w_x, w_y, w_z};
  const int shp[3][3] = {{op->ex[0], d.ny, d.nz}, {d.nx, op->ex[1], d.nz},
                         {d.nx, d.ny, op->ex[2]}};
  int rc = FS_OK;
  for (int a = 0; a < 3 && rc == FS_OK; ++a) {
    fs::permute_out<double>(ctx->stream, op->w[a], tmp, shp[a][0], shp[a][1], shp[a][2], d.nx,
                            d.ny);
    const size_t cnt = (size_t)shp[a][0] * shp[a][1] * shp[a][2];
    if (hipMemcpyAsync(outs[a], tmp, cnt * sizeof(double), hipMemcpyDeviceToHost,
                       ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(FS_EHIP, "edge download failed");
  }
  dfree(tmp);
  return rc;
  FS_GUARD_END
}

int fs_sys_create(fs_op* op, const uint8_t* mask, int pin, fs_sys** out, fs_sys_info* info) {
  FS_GUARD_BEGIN
  if (!op || !out) return fail(FS_EINVAL, "NULL argument");
  *out = nullptr;
  fs_ctx* ctx = op->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  fs_sys* s = new fs_sys();
  s->ctx = ctx;
  s->stream = ctx->stream;
  s->op = op;
  s->d = op->d;
  const fs::Dims& d = s->d;
  s->vstride = (d.n + 63) / 64 * 64;
  // stencil tiling: ~2K tiles, at least 8 planes per tile
  fs::Tiling& tl = s->tl;
  tl.tiles_x = (d.nx + fs::kTileX - 1) / fs::kTileX;
  tl.tiles_y = (d.ny + fs::kTileY - 1) / fs::kTileY;
  const int64_t txy = (int64_t)tl.tiles_x * tl.tiles_y;
  tl.kz = (int)std::max<int64_t>(8, (d.nz * txy + 2047) / 2048);
  tl.kz = std::min(tl.kz, d.nz);
  tl.tiles_z = (d.nz + tl.kz - 1) / tl.kz;
  tl.ntiles = (int)(txy * tl.tiles_z);
  tl.nblocks = (tl.ntiles + 7) / 8 * 8;
  s->vec_blocks = (int)std::min<int64_t>((d.n + fs::kBlock - 1) / fs::kBlock,
                                         (int64_t)std::max(ctx->n_cu, 1) * 8);
  s->vec_blocks = std::max(s->vec_blocks, 1);
  // FP32 stencil with 4 nodes per lane when rows are a multiple of 4
  // wrap-around edges run on the one-node-per-lane kernels only
  // (it addresses nodes with 32-bit offsets)
  s->use_x4 = d.nx % 4 == 0 && d.per == 0 && d.n + d.sxy < (int64_t)INT32_MAX ? 1 : 0;
  {
    fs::Tiling& t4 = s->tl4;
    t4.tiles_x = (d.nx / 4 + fs::kWave - 1) / fs::kWave;
    t4.tiles_y = (d.ny + fs::kTileY - 1) / fs::kTileY;
    const int64_t t4xy = (int64_t)t4.tiles_x * t4.tiles_y;
    t4.kz = (int)std::min<int64_t>(16, std::max<int64_t>(4, (d.nz * t4xy + 1023) / 1024));
    t4.kz = std::max(1, std::min(t4.kz, d.nz));
    t4.tiles_z = (d.nz + t4.kz - 1) / t4.kz;
    t4.ntiles = (int)(t4xy * t4.tiles_z);
    t4.nblocks = (t4.ntiles + 7) / 8 * 8;
  }
  // FP32 Krylov stencil: W nodes per lane, the widest of 4, 3, 2 that divides
  // the rows (fs_sys_autotune may pick another width or the one-node kernel)
  {
    const bool wide_ok = d.per == 0 && d.n + d.sxy < (int64_t)INT32_MAX;
    s->xw = !wide_ok ? 0 : d.nx % 4 == 0 ? 4 : d.nx % 3 == 0 ? 3 : d.nx % 2 == 0 ? 2 : 0;
    if (s->xw) {
      fs::Tiling& tw = s->tlw;
      tw.tiles_x = (d.nx / s->xw + fs::kWave - 1) / fs::kWave;
      tw.tiles_y = (d.ny + fs::kTileY - 1) / fs::kTileY;
      const int64_t twxy = (int64_t)tw.tiles_x * tw.tiles_y;
      tw.kz = (int)std::min<int64_t>(16, std::max<int64_t>(4, (d.nz * twxy + 1023) / 1024));
      tw.kz = std::max(1, std::min(tw.kz, d.nz));
      tw.tiles_z = (d.nz + tw.kz - 1) / tw.kz;
      tw.ntiles = (int)(twxy * tw.tiles_z);
      tw.nblocks = (tw.ntiles + 7) / 8 * 8;
    }
  }
  // FP64 stencil with 2 nodes per lane when rows are even
  s->use_x2 = d.nx % 2 == 0 && d.per == 0 ? 1 : 0;
  {
    fs::Tiling& t2 = s->tl2;
    t2.tiles_x = (d.nx / 2 + fs::kWave - 1) / fs::kWave;
    t2.tiles_y = (d.ny + fs::kTileY - 1) / fs::kTileY;
    const int64_t t2xy = (int64_t)t2.tiles_x * t2.tiles_y;
    t2.kz = (int)std::min<int64_t>(16, std::max<int64_t>(4, (d.nz * t2xy + 1023) / 1024));
    t2.kz = std::max(1, std::min(t2.kz, d.nz));
    t2.tiles_z = (d.nz + t2.kz - 1) / t2.kz;
    t2.ntiles = (int)(t2xy * t2.tiles_z);
    t2.nblocks = (t2.ntiles + 7) / 8 * 8;
  }
  const int64_t maxb = std::max<int64_t>(
      std::max<int64_t>(std::max<int64_t>(tl.nblocks, s->tl4.nblocks), s->tl2.nblocks),
      std::max<int64_t>(s->vec_blocks, s->xw ? s->tlw.nblocks : 0));
  s->partials_cap = maxb * fs::kMaxRhs * 2 + 64 * (fs::kMaxRhs + 1) * fs::kMaxRhs;
  int rc = FS_OK;
  double* dsum = nullptr;
  void* cnt_dev = nullptr;
  void* cnt_host = nullptr;
  if ((rc = dalloc(&s->mask, (size_t)d.n)) || (rc = dalloc(&s->dhalf, (size_t)d.n)) ||
      (rc = dalloc(&s->dinv, (size_t)d.n)) || (rc = dalloc(&dsum, (size_t)d.n)) ||
      (rc = dalloc(&s->partials, (size_t)s->partials_cap)) ||
      (rc = dalloc(&s->sc, (size_t)fs::kMaxRhs)) ||
      (rc = dalloc((char**)&cnt_dev, fs::kSetupCountsBytes))) {
    dfree(dsum);
    free_sys(s);
    return rc;
  }
  for (int a = 0; a < 3 && rc == FS_OK; ++a) {
    rc = dalloc(&s->c64[a], (size_t)d.n);
    if (rc == FS_OK) rc = dalloc(&s->c32[a], (size_t)d.n);
  }
  if (rc == FS_OK && hipHostMalloc((void**)&s->sc_host, sizeof(fs::CgScalars) * fs::kMaxRhs) != hipSuccess)
    rc = fail(FS_ENOMEM, "pinned allocation failed");
  if (rc == FS_OK && hipHostMalloc(&cnt_host, fs::kSetupCountsBytes) != hipSuccess)
    rc = fail(FS_ENOMEM, "pinned allocation failed");
  if (rc == FS_OK && hipMemset(s->sc, 0, sizeof(fs::CgScalars) * fs::kMaxRhs) != hipSuccess)
    rc = fail(FS_EHIP, "scalar init failed");
  if (rc == FS_OK) {
    if (mask) {
      rc = upload_permuted(ctx, mask, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz, s->mask);
    } else if (hipMemset(s->mask, 0, (size_t)d.n) != hipSuccess) {
      rc = fail(FS_EHIP, "mask init failed");
    }
  }
  if (rc == FS_OK) {
    const int st = fs::setup_system(s, dsum, cnt_dev, cnt_host, pin);
    if (st == 1)
      rc = fail(FS_EINVAL,
                "system is singular: every row sums to zero, so there is no Dirichlet node "
                "anywhere and the potential is only defined up to a constant.  Drive a "
                "terminal, ground a wall, or pass pin=True.");
    else if (st != 0)
      rc = fail(FS_EHIP, "system setup failed: %s", hipGetErrorString(hipGetLastError()));
  }
  if (rc == FS_OK && (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess))
    rc = fail(FS_EHIP, "system setup kernels failed");
  dfree(dsum);
  dfree(cnt_dev);
  if (cnt_host) (void)hipHostFree(cnt_host);
  if (rc != FS_OK) {
    free_sys(s);
    return rc;
  }
  if (info) *info = s->info;
  *out = s;
  return FS_OK;
  FS_GUARD_END
}

int fs_sys_destroy(fs_sys* sys) {
  if (sys) (void)hipSetDevice(sys->ctx->device);
  free_sys(sys);
  return FS_OK;
}

int fs_sys_solve(fs_sys* s, int nrhs, const double* b, const double* x0, double* x,
                 const fs_solve_params* params, fs_solve_stats* stats) {
  FS_GUARD_BEGIN
  return sys_solve(s, nrhs, b, x0, x, params, stats, nullptr);
  FS_GUARD_END
}

int fs_sys_solve_noalloc(fs_sys* s, int nrhs, const double* b, const double* x0, double* x,
                         const fs_solve_params* params, fs_solve_stats* stats) {
  FS_GUARD_BEGIN
  if (!s || !b || !x || !params) return fail(FS_EINVAL, "NULL argument");
  if (nrhs < 1) return fail(FS_EINVAL, "nrhs must be >= 1");
  if (!s->ctx->scratch)
    return fail(FS_EHIP, "no scratch address on device %d (hipGetSymbolAddress failed when the "
                "context was created)", s->ctx->device);
  FS_HIP(hipSetDevice(s->ctx->device));
  const int nr = std::min(fs::kMaxRhs, nrhs);
  const size_t need = ScratchLoan::need(s, nr, x0 != nullptr);
  if (need > kScratchBytes)
    return fail(FS_EINVAL,
                "fs_sys_solve_noalloc needs %zu bytes of scratch for %d right-hand side(s) of "
                "%lld nodes%s; its scratch holds %zu (68 B per node and right-hand side, 76 "
                "with initial guesses, up to 8 right-hand sides at a time)",
                need, nr, (long long)s->d.n, x0 ? " with initial guesses" : "", kScratchBytes);
  ScratchLoan loan(s);
  if (int rc = loan.lend(s->ctx->scratch, nr, x0 != nullptr)) return rc;
  return sys_solve(s, nrhs, b, x0, x, params, stats, &loan);
  FS_GUARD_END
}

}  // extern "C"

namespace {

int sys_solve(fs_sys* s, int nrhs, const double* b, const double* x0, double* x,
              const fs_solve_params* params, fs_solve_stats* stats, ScratchLoan* loan) {
  if (!s || !b || !x || !params) return fail(FS_EINVAL, "NULL argument");
  if (nrhs < 1) return fail(FS_EINVAL, "nrhs must be >= 1");
  if (!(params->tol > 0.0)) return fail(FS_EINVAL, "tol must be > 0");
  if (params->precision != FS_PREC_FP64 && params->precision != FS_PREC_MIXED &&
      params->precision != FS_PREC_FCG)
    return fail(FS_EINVAL, "unknown precision %d", params->precision);
  fs_ctx* ctx = s->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  const fs::Dims& d = s->d;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  const int maxiter = default_maxiter(s, params);
  int status = FS_OK;
  float ms_total = 0.0f;
  if (stats) stats->converged = 1;
  kernel_widths(s, params->precision, stats);
  for (int r0 = 0; r0 < nrhs; r0 += fs::kMaxRhs) {
    const int nr = std::min(fs::kMaxRhs, nrhs - r0);
    if (int rc = ensure_ws(s, nr)) return rc;
    // host copies go through node_tmp (written again only after the uploads)
    // when the workspace is on loan
    double* stage = loan ? s->node_tmp : nullptr;
    // full right-hand sides (x-fast) into q, initial guesses into x0d
    for (int r = 0; r < nr; ++r) {
      const double* hb = b + (size_t)(r0 + r) * d.n;
      if (int rc = upload_permuted(ctx, hb, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz,
                                   s->q + (size_t)r * s->vstride, stage))
        return rc;
    }
    DevBufs B;
    double* x0d = nullptr;
    if (x0) {
      if (loan) x0d = loan->x0d;
      else if (int rc = B.get(&x0d, (size_t)nr * s->vstride)) return rc;
      for (int r = 0; r < nr; ++r)
        if (int rc = upload_permuted(ctx, x0 + (size_t)(r0 + r) * d.n, d.nx, d.ny, d.nz, d.nx,
                                     d.ny, d.nz, x0d + (size_t)r * s->vstride, stage))
          return rc;
    }
    prof_begin(s);
    FS_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    fs::launch_lift_general(s, nr, s->q);
    // keep the full b for the Dirichlet values: move it to node_tmp
    FS_HIP(hipMemcpyAsync(s->node_tmp, s->q, (size_t)nr * s->vstride * sizeof(double),
                          hipMemcpyDeviceToDevice, ctx->stream));
    fs::launch_init(s, nr, params->tol, maxiter);
    // LinearSystem.solve(b, x0): x_s = x0 * D^1/2 on the free rows
    // (fieldspice/solvers/poisson.py:543-547), r = b_s - A_s x_s
    if (x0d)
      for (int r = 0; r < nr; ++r)
        fs::launch_init_warm(s, x0d + (size_t)r * s->vstride, nullptr, params->tol, maxiter, 0, r);
    FS_HIP(hipGetLastError());
    if (int rc = run_krylov(s, nr, *params)) return rc;
    fs::launch_residual(s, nr, s->r);
    fs::launch_expand(s, nr, s->node_tmp, nullptr, 0, s->q);
    FS_HIP(hipGetLastError());
    FS_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    FS_HIP(hipMemcpyAsync(s->sc_host, s->sc, sizeof(fs::CgScalars) * nr, hipMemcpyDeviceToHost,
                          ctx->stream));
    FS_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    FS_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ms_total += ms;
    int st = FS_OK;
    collect_stats(s, nr, stats, st);
    prof_end(s, stats, params->precision);
    if (st != FS_OK) {
      status = st;
      if (stats) stats->converged = 0;
    }
    for (int r = 0; r < nr; ++r) {
      if (int rc = download_nodes(ctx, d, s->q + (size_t)r * s->vstride,
                                  x + (size_t)(r0 + r) * d.n, s->r))
        return rc;
    }
  }
  if (stats) stats->solve_time = ms_total * 1e-3;
  if (status == FS_NOT_CONVERGED)
    fail(FS_NOT_CONVERGED, "Krylov iteration did not reach tol %.3g within %d iterations",
         params->tol, maxiter);
  return status;
}

}  // namespace

extern "C" {

int fs_cap_prepare(fs_sys* s, int nt, const int32_t* term_id, fs_cap** out) {
  FS_GUARD_BEGIN
  if (!s || !term_id || !out) return fail(FS_EINVAL, "NULL argument");
  *out = nullptr;
  if (nt < 1) return fail(FS_EINVAL, "capacitance extraction needs at least one terminal");
  fs_ctx* ctx = s->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  const fs::Dims& d = s->d;
  // node lists per bin (terminal 0..nt-1, shield nt) in device order
  const int nbins = nt + 1;
  std::vector<int64_t> counts(nbins, 0);
  for (int64_t ref = 0; ref < d.n; ++ref) {
    const int t = term_id[ref];
    if (t < -1 || t > nt)
      return fail(FS_EINVAL, "term_id[%lld] = %d out of range", (long long)ref, t);
    if (t >= 0) counts[t] += 1;
  }
  for (int t = 0; t < nt; ++t)
    if (counts[t] == 0) return fail(FS_EINVAL, "terminal %d has no nodes", t);
  std::vector<int64_t> offs(nbins + 1, 0);
  for (int t = 0; t < nbins; ++t) offs[t + 1] = offs[t] + counts[t];
  std::vector<int64_t> nodes((size_t)std::max<int64_t>(offs[nbins], 1));
  std::vector<int64_t> fill(offs.begin(), offs.end() - 1);
  for (int64_t i = 0; i < d.nx; ++i)
    for (int64_t j = 0; j < d.ny; ++j)
      for (int64_t k = 0; k < d.nz; ++k) {
        const int t = term_id[(i * d.ny + j) * d.nz + k];
        if (t >= 0) nodes[fill[t]++] = (k * d.ny + j) * d.nx + i;
      }
  fs_cap* cap = new fs_cap();
  cap->sys = s;
  cap->nt = nt;
  cap->nbins = nbins;
  cap->has_shield = counts[nt] > 0;
  int64_t maxcnt = 1;
  for (int t = 0; t < nbins; ++t) maxcnt = std::max(maxcnt, counts[t]);
  cap->nbx = (int)std::min<int64_t>((maxcnt + fs::kBlock - 1) / fs::kBlock, 64);
  int rc = FS_OK;
  if ((rc = dalloc(&cap->tid, (size_t)d.n)) || (rc = dalloc(&cap->nodes, nodes.size())) ||
      (rc = dalloc(&cap->offs, offs.size())) ||
      (rc = dalloc(&cap->out, (size_t)nbins * fs::kMaxRhs))) {
    fs_cap_destroy(cap);
    return rc;
  }
  rc = upload_permuted(ctx, term_id, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz, cap->tid);
  if (rc == FS_OK &&
      (hipMemcpy(cap->nodes, nodes.data(), nodes.size() * sizeof(int64_t),
                 hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(cap->offs, offs.data(), offs.size() * sizeof(int64_t),
                 hipMemcpyHostToDevice) != hipSuccess))
    rc = fail(FS_EHIP, "node list upload failed");
  const int64_t need = (int64_t)nbins * fs::kMaxRhs * cap->nbx;
  if (rc == FS_OK && need > s->partials_cap) {
    dfree(s->partials);
    rc = dalloc(&s->partials, (size_t)need);
    s->partials_cap = rc == FS_OK ? need : 0;
  }
  if (rc != FS_OK) {
    fs_cap_destroy(cap);
    return rc;
  }
  *out = cap;
  return FS_OK;
  FS_GUARD_END
}

int fs_cap_destroy(fs_cap* cap) {
  if (!cap) return FS_OK;
  (void)hipSetDevice(cap->sys->ctx->device);
  dfree(cap->tid);
  dfree(cap->nodes);
  dfree(cap->offs);
  dfree(cap->out);
  delete cap;
  return FS_OK;
}

int fs_cap_solve(fs_cap* cap, const fs_op* L, double* C, double* q_shield, double* phi_out,
                 const fs_solve_params* params, fs_solve_stats* stats) {
  FS_GUARD_BEGIN
  if (!cap || !L || !C || !params) return fail(FS_EINVAL, "NULL argument");
  fs_sys* s = cap->sys;
  if (L->Nx != s->op->Nx || L->Ny != s->op->Ny || L->Nz != s->op->Nz)
    return fail(FS_EINVAL, "L and the system differ in grid");
  if (!(params->tol > 0.0)) return fail(FS_EINVAL, "tol must be > 0");
  if (params->precision != FS_PREC_FP64 && params->precision != FS_PREC_MIXED &&
      params->precision != FS_PREC_FCG)
    return fail(FS_EINVAL, "unknown precision %d", params->precision);
  fs_ctx* ctx = s->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  const fs::Dims& d = s->d;
  const int nt = cap->nt, nbins = cap->nbins;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->converged = 1;
  }
  kernel_widths(s, params->precision, stats);
  const int maxiter = default_maxiter(s, params);
  const bool mixed = params->precision != FS_PREC_FP64;
  int status = FS_OK;
  float ms_total = 0.0f;
  std::vector<double> hout((size_t)nbins * fs::kMaxRhs);
  for (int j0 = 0; j0 < nt; j0 += fs::kMaxRhs) {
    const int nr = std::min(fs::kMaxRhs, nt - j0);
    if (int rc = ensure_ws(s, nr)) return rc;
    prof_begin(s);
    FS_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    fs::launch_lift_cap(s, nr, L, cap->tid, j0);
    fs::launch_init(s, nr, params->tol, maxiter);
    FS_HIP(hipGetLastError());
    if (int rc = run_krylov(s, nr, *params)) return rc;
    fs::launch_residual(s, nr, s->r);
    fs::launch_expand(s, nr, nullptr, cap->tid, j0, s->q);
    fs::launch_charge(s, nr, L, s->q, cap->nodes, cap->offs, nbins, cap->nbx, cap->out);
    FS_HIP(hipGetLastError());
    FS_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    FS_HIP(hipMemcpyAsync(hout.data(), cap->out, sizeof(double) * nbins * nr,
                          hipMemcpyDeviceToHost, ctx->stream));
    FS_HIP(hipMemcpyAsync(s->sc_host, s->sc, sizeof(fs::CgScalars) * nr, hipMemcpyDeviceToHost,
                          ctx->stream));
    FS_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ms_total += ms;
    int st = FS_OK;
    collect_stats(s, nr, stats, st);
    prof_end(s, stats, params->precision);
    if (st != FS_OK) {
      status = st;
      if (stats) stats->converged = 0;
    }
    for (int r = 0; r < nr; ++r) {
      for (int i = 0; i < nt; ++i) C[(size_t)i * nt + j0 + r] = hout[(size_t)i * nr + r];
      if (q_shield) q_shield[j0 + r] = cap->has_shield ? hout[(size_t)nt * nr + r] : 0.0;
    }
    if (phi_out) {
      for (int r = 0; r < nr; ++r)
        if (int rc = download_nodes(ctx, d, s->q + (size_t)r * s->vstride,
                                    phi_out + (size_t)(j0 + r) * d.n, s->r))
          return rc;
    }
  }
  if (stats) stats->solve_time = ms_total * 1e-3;
  if (status == FS_NOT_CONVERGED)
    fail(FS_NOT_CONVERGED, "Krylov iteration did not reach tol %.3g within %d iterations",
         params->tol, maxiter);
  return status;
  FS_GUARD_END
}

int fs_capacitance(fs_sys* s, const fs_op* L, int nt, const int32_t* term_id, double* C,
                   double* q_shield, double* phi_out, const fs_solve_params* params,
                   fs_solve_stats* stats) {
  fs_cap* cap = nullptr;
  if (int rc = fs_cap_prepare(s, nt, term_id, &cap)) return rc;
  const int rc = fs_cap_solve(cap, L, C, q_shield, phi_out, params, stats);
  const std::string keep = fs::g_err;
  fs_cap_destroy(cap);
  fs::g_err = keep;
  return rc;
}


int fs_op_edge_power(fs_op* op, const double* x, double* q) {
  FS_GUARD_BEGIN
  if (!op || !x || !q) return fail(FS_EINVAL, "NULL argument");
  if (op->d.per) return fail(FS_EINVAL, "periodic operators run fs_sys_solve and fs_capacitance only (FS_PREC_FP64 or FS_PREC_MIXED)");
  fs_ctx* ctx = op->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  const fs::Dims& d = op->d;
  double *dx = nullptr, *dq = nullptr;
  int rc = dalloc(&dx, (size_t)d.n);
  if (rc == FS_OK) rc = dalloc(&dq, (size_t)d.n);
  if (rc == FS_OK) rc = upload_permuted(ctx, x, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz, dx);
  if (rc == FS_OK) {
    fs::launch_edge_power(ctx->stream, op, dx, dq, ctx->n_cu);
    rc = download_nodes(ctx, d, dq, q, dx);
  }
  dfree(dx);
  dfree(dq);
  return rc;
  FS_GUARD_END
}


int fs_transient_run(fs_sys* s, const fs_op* Aexp, const fs_transient_desc* D, double* x,
                     double* stored, double* probe_i, double* probe_v, double* step_stats,
                     const fs_solve_params* params, fs_solve_stats* stats) {
  FS_GUARD_BEGIN
  if (!s || !D || !x || !params) return fail(FS_EINVAL, "NULL argument");
  if (s->d.per) return fail(FS_EINVAL, "periodic operators run fs_sys_solve and fs_capacitance only (FS_PREC_FP64 or FS_PREC_MIXED)");
  if (D->nsteps < 1) return fail(FS_EINVAL, "nsteps must be >= 1");
  if (D->ngroups < 0 || (D->ngroups > 0 && (!D->group || !D->values)))
    return fail(FS_EINVAL, "groups need group ids and values");
  if (D->nsrc < 0 || (D->nsrc > 0 && (!D->src || !D->src_scale)))
    return fail(FS_EINVAL, "sources need vectors and scales");
  if (D->nprobe < 0 || (D->nprobe > 0 && (!D->probe || !D->probe_w)))
    return fail(FS_EINVAL, "probes need ids and weights");
  if (D->warm_start < 0 || D->warm_start > 3)
    return fail(FS_EINVAL, "warm_start must be 0, 1, 2 or 3 (got %d)", D->warm_start);
  if (D->warm_start == 3 && (D->proj_dim < 0 || D->proj_dim == 1 || D->proj_dim > fs::kProjMax))
    return fail(FS_EINVAL, "proj_dim must be 0 (default 8) or 2..%d (got %d)", fs::kProjMax,
                D->proj_dim);
  if (params->precision != FS_PREC_FP64 && params->precision != FS_PREC_MIXED &&
      params->precision != FS_PREC_FCG)
    return fail(FS_EINVAL, "unknown precision %d", params->precision);
  if (!(params->tol > 0.0)) return fail(FS_EINVAL, "tol must be > 0");
  const fs_op* ops[5] = {Aexp, D->cur_a, D->cur_b, D->quad_a, D->quad_b};
  for (const fs_op* o : ops)
    if (o && (o->Nx != s->op->Nx || o->Ny != s->op->Ny || o->Nz != s->op->Nz))
      return fail(FS_EINVAL, "operators differ in grid");
  if (D->nprobe > 0 && !D->cur_a) return fail(FS_EINVAL, "probe currents need cur_a");
  fs_ctx* ctx = s->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  const fs::Dims& d = s->d;
  const int nsteps = D->nsteps, ng = D->ngroups, np = D->nprobe;
  if (int rc = ensure_ws(s, 1)) return rc;
  DevBufs B;
  double *xa = nullptr, *xb = nullptr, *xc = nullptr, *vals = nullptr, *src = nullptr;
  double* scale = nullptr;
  double *pw = nullptr, *oi = nullptr, *ov = nullptr, *ost = nullptr;
  int32_t *grp = nullptr;
  int64_t *pnodes = nullptr, *poffs = nullptr;
  int rc = FS_OK;
  if ((rc = B.get(&xa, (size_t)s->vstride)) || (rc = B.get(&xb, (size_t)s->vstride)) ||
      (rc = B.get(&grp, (size_t)d.n)) ||
      (rc = B.get(&vals, (size_t)std::max(1, (nsteps + 1) * ng))) ||
      (rc = B.get(&ost, (size_t)(nsteps + 1) * 4)))
    return rc;
  // the field two steps back, for the extrapolated initial guess
  const bool rot = D->warm_start >= 2;
  if (rot && (rc = B.get(&xc, (size_t)s->vstride))) return rc;
  // projection guesses: A_s-orthonormal basis X, AX = A_s X, the guess x0 of
  // the step, the previous step's solution, coefficients and a norm slot
  const int pm = D->warm_start == 3 ? (D->proj_dim > 0 ? D->proj_dim : 8) : 0;
  double *PX = nullptr, *PAX = nullptr, *px0 = nullptr, *pxs = nullptr, *pcoef = nullptr;
  if (pm > 0) {
    if ((rc = B.get(&PX, (size_t)pm * s->vstride)) || (rc = B.get(&PAX, (size_t)pm * s->vstride)) ||
        (rc = B.get(&px0, (size_t)s->vstride)) || (rc = B.get(&pxs, (size_t)s->vstride)) ||
        (rc = B.get(&pcoef, (size_t)fs::kProjMax + 2)))
      return rc;
    const int64_t need = (int64_t)fs::kProjMax * s->vec_blocks;
    if (need > s->partials_cap) {
      dfree(s->partials);
      if ((rc = dalloc(&s->partials, (size_t)need))) {
        s->partials_cap = 0;
        return rc;
      }
      s->partials_cap = need;
    }
  }
  int pnb = 0;            // basis vectors held
  bool have_pxs = false;  // pxs holds the previous step's x_s
  if ((rc = upload_permuted(ctx, x, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz, xa))) return rc;
  const bool have_before = rot && D->x_before;
  if (have_before &&
      (rc = upload_permuted(ctx, D->x_before, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz, xc)))
    return rc;
  if (ng > 0) {
    // group ids must agree with the system's Dirichlet rows
    for (int64_t i = 0; i < d.n; ++i)
      if (D->group[i] < -1 || D->group[i] >= ng)
        return fail(FS_EINVAL, "group[%lld] = %d out of range", (long long)i, D->group[i]);
    if ((rc = upload_permuted(ctx, D->group, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz, grp))) return rc;
    FS_HIP(hipMemcpy(vals, D->values, sizeof(double) * (nsteps + 1) * ng, hipMemcpyHostToDevice));
  } else {
    FS_HIP(hipMemset(grp, 0xff, sizeof(int32_t) * d.n));  // all -1
  }
  if (D->nsrc > 0) {
    if ((rc = B.get(&src, (size_t)D->nsrc * s->vstride)) ||
        (rc = B.get(&scale, (size_t)(nsteps + 1) * D->nsrc)))
      return rc;
    for (int k = 0; k < D->nsrc; ++k)
      if ((rc = upload_permuted(ctx, D->src + (size_t)k * d.n, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz,
                                src + (size_t)k * s->vstride)))
        return rc;
    FS_HIP(hipMemcpy(scale, D->src_scale, sizeof(double) * (nsteps + 1) * D->nsrc,
                     hipMemcpyHostToDevice));
  }
  int nbx = 1;
  if (np > 0) {
    std::vector<int64_t> nodes, offs;
    if ((rc = build_lists(d, D->probe, np, nodes, offs))) return rc;
    int64_t maxc = 1;
    for (int t = 0; t < np; ++t) maxc = std::max(maxc, offs[t + 1] - offs[t]);
    nbx = (int)std::min<int64_t>((maxc + fs::kBlock - 1) / fs::kBlock, 64);
    if ((rc = B.get(&pnodes, nodes.size())) || (rc = B.get(&poffs, offs.size())) ||
        (rc = B.get(&pw, (size_t)d.n)) || (rc = B.get(&oi, (size_t)(nsteps + 1) * np)) ||
        (rc = B.get(&ov, (size_t)(nsteps + 1) * np)))
      return rc;
    FS_HIP(hipMemcpy(pnodes, nodes.data(), nodes.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    FS_HIP(hipMemcpy(poffs, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    if ((rc = upload_permuted(ctx, D->probe_w, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz, pw))) return rc;
    const int64_t need = (int64_t)np * 2 * nbx;
    if (need > s->partials_cap) {
      dfree(s->partials);
      if ((rc = dalloc(&s->partials, (size_t)need))) {
        s->partials_cap = 0;
        return rc;
      }
      s->partials_cap = need;
    }
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->converged = 1;
  }
  kernel_widths(s, params->precision, stats);
  const int maxiter = default_maxiter(s, params);
  const int store_every = D->store_every;
  s->acc_fallbacks = 0;
  s->acc_pc_applies = 0;
  s->acc_replacements = 0;
  int worst_done = 1;
  int64_t nstored = 0;
  auto store = [&](const double* dev) -> int {
    if (!stored) return FS_OK;
    return download_nodes(ctx, d, dev, stored + (size_t)nstored++ * d.n, s->r);
  };
  // step 0 records
  fs::launch_step_stats(s, xa, D->quad_a, D->quad_b, ost);
  if (np > 0)
    fs::launch_probe(s, D->cur_a, D->cur_b, xa, xa, pw, pnodes, poffs, np, nbx, nullptr, ov);
  FS_HIP(hipGetLastError());
  if (store_every > 0 && (rc = store(xa))) return rc;
  int status = FS_OK;
  double worst_res = 0.0, worst_req = 0.0;
  int64_t total_it = 0;
  FS_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  for (int k = 1; k <= nsteps; ++k) {
    const double* vk = vals + (size_t)k * ng;
    fs::launch_explicit(s, Aexp, xa, src, scale ? scale + (size_t)k * D->nsrc : nullptr, D->nsrc,
                        s->vstride, s->node_tmp);
    fs::launch_lift_groups(s, s->node_tmp, grp, vk);
    // initial guess: 0, the previous field, the linear extrapolation
    // 2 x_{k-1} - x_{k-2} (from the second step on unless x_{-1} was given),
    // or the projection onto the basis of earlier solutions
    const bool extrap = rot && (k >= 2 || have_before);
    const bool from_basis = pm > 0 && pnb > 0;
    if (from_basis)
      fs::launch_proj_guess(s, PX, PAX, s->vstride, pnb, pcoef, px0, params->tol, maxiter);
    else if (D->warm_start)
      fs::launch_init_warm(s, xa, extrap ? xc : nullptr, params->tol, maxiter);
    else fs::launch_init(s, 1, params->tol, maxiter);
    FS_HIP(hipGetLastError());
    if ((rc = run_krylov(s, 1, *params))) return rc;
    fs::launch_residual(s, 1, s->r);
    fs::launch_expand_groups(s, grp, vk, xb);
    fs::launch_step_stats(s, xb, D->quad_a, D->quad_b, ost + (size_t)k * 4);
    if (np > 0)
      fs::launch_probe(s, D->cur_a, D->cur_b, xb, xa, pw, pnodes, poffs, np, nbx,
                       oi + (size_t)k * np, ov + (size_t)k * np);
    FS_HIP(hipGetLastError());
    FS_HIP(hipMemcpyAsync(s->sc_host, s->sc, sizeof(fs::CgScalars), hipMemcpyDeviceToHost,
                          ctx->stream));
    FS_HIP(hipStreamSynchronize(ctx->stream));
    const fs::CgScalars& c = s->sc_host[0];
    total_it += c.iter;
    double res = c.aux2 > 0 ? std::sqrt(c.aux1 / c.aux2) : 0.0;
    double req = c.bnorm2 > 0 ? std::sqrt(c.aux3 / c.bnorm2) : 0.0;
    if (!(res >= 0.0)) res = INFINITY;
    if (!(req >= 0.0)) req = INFINITY;
    worst_res = std::max(worst_res, res);
    worst_req = std::max(worst_req, req);
    if (stats) stats->refreshes += c.refreshes;
    worst_done = std::max(worst_done, c.done);
    if (c.done != 1) {
      status = FS_NOT_CONVERGED;
      if (stats) stats->converged = 0;
    }
    if (pm > 0) {
      // extend the basis with the new direction x_s - x0 (x_s itself when
      // the guess was not from the basis); a full basis restarts from x_s
      // and x_s - x_s(previous step), keeping the extrapolation direction
      const bool restart = pnb == pm;
      if (restart) pnb = 0;
      const int ncand = restart && have_pxs ? 2 : 1;
      for (int c = 0; c < ncand; ++c) {
        double* slot = PX + (int64_t)pnb * s->vstride;
        fs::launch_diff(s, s->x, c == 1 ? pxs : (from_basis && !restart ? px0 : nullptr), slot);
        fs::launch_proj_append(s, PX, PAX, s->vstride, pnb, pcoef, pcoef + fs::kProjMax);
        double q = 0.0;
        FS_HIP(hipMemcpyAsync(&q, pcoef + fs::kProjMax, sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
        FS_HIP(hipStreamSynchronize(ctx->stream));
        if (q > 0.0 && q < INFINITY) ++pnb;
      }
      fs::launch_diff(s, s->x, nullptr, pxs);
      have_pxs = true;
    }
    if (rot) std::swap(xc, xa);  // xc <- x_{k-1}
    std::swap(xa, xb);           // xa <- x_k
    if (store_every > 0 && k % store_every == 0 && (rc = store(xa))) return rc;
  }
  FS_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  FS_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  if ((rc = download_nodes(ctx, d, xa, x, s->r))) return rc;
  // x_{nsteps-1}: the last field before xa (x_0 itself after one step)
  if (D->x_before_out) {
    const double* xl = nullptr;
    if (rot) xl = xc;                       // rotated to hold x_{nsteps-1}
    else if (nsteps >= 1) xl = xb;          // swapped out of xa
    if ((rc = download_nodes(ctx, d, xl, D->x_before_out, s->r))) return rc;
  }
  if (step_stats)
    FS_HIP(hipMemcpy(step_stats, ost, sizeof(double) * (nsteps + 1) * 4, hipMemcpyDeviceToHost));
  if (np > 0) {
    if (probe_i) {
      FS_HIP(hipMemcpy(probe_i, oi, sizeof(double) * (nsteps + 1) * np, hipMemcpyDeviceToHost));
      for (int m = 0; m < np; ++m) probe_i[m] = nsteps >= 1 ? probe_i[np + m] : 0.0;
    }
    if (probe_v)
      FS_HIP(hipMemcpy(probe_v, ov, sizeof(double) * (nsteps + 1) * np, hipMemcpyDeviceToHost));
  }
  if (stats) {
    stats->iterations = (int)std::min<int64_t>(total_it, INT32_MAX);
    stats->residual = worst_res;
    stats->residual_eq = worst_req;
    stats->solve_time = ms * 1e-3;
    stats->fallbacks = s->acc_fallbacks;
    stats->pc_applies = s->acc_pc_applies;
    stats->replacements = s->acc_replacements;
    stats->stop_reason = worst_done;
    stats->maxiter = maxiter;
  }
  s->acc_fallbacks = 0;
  s->acc_pc_applies = 0;
  s->acc_replacements = 0;
  if (status == FS_NOT_CONVERGED)
    fail(FS_NOT_CONVERGED, "a time step did not reach tol %.3g within %d iterations", params->tol,
         maxiter);
  return status;
  FS_GUARD_END
}

int fs_sys_set_profiling(fs_sys* s, int on) {
  FS_GUARD_BEGIN
  if (!s) return fail(FS_EINVAL, "NULL argument");
  FS_HIP(hipSetDevice(s->ctx->device));
  if (on && s->prof_slots == nullptr) {
    // one slot per iteration of a chunk, and one for the start of a cycle
    s->prof_slots = new fs::ProfSlot[kMaxChunk + 1];
    for (int c = 0; c <= kMaxChunk; ++c) {
      for (int k = 0; k < FS_PROF_KINDS; ++k) {
        FS_HIP(hipEventCreate(&s->prof_slots[c].a[k]));
        FS_HIP(hipEventCreate(&s->prof_slots[c].b[k]));
      }
      s->ev_cap = c + 1;
    }
    if (int rc = dalloc(&s->prof_dev, 2 * FS_PROF_KINDS)) return rc;
    FS_HIP(hipMemset(s->prof_dev, 0, 2 * FS_PROF_KINDS * sizeof(unsigned long long)));
  }
  s->profiling = on ? 1 : 0;
  return FS_OK;
  FS_GUARD_END
}


// --------------------------------------------------------------------------
// stencil tile autotuning
// --------------------------------------------------------------------------
}  // extern "C"

namespace {

// tiles of the FP32 stencil with w nodes per lane (1: the 64-wide tiles of
// the one-node-per-thread kernel)
fs::Tiling tiling_for(const fs::Dims& d, int kz, int w = 1) {
  fs::Tiling tl;
  tl.tiles_x = w > 1 ? (d.nx / w + fs::kWave - 1) / fs::kWave : (d.nx + fs::kTileX - 1) / fs::kTileX;
  tl.tiles_y = (d.ny + fs::kTileY - 1) / fs::kTileY;
  tl.kz = std::max(1, std::min(kz, d.nz));
  tl.tiles_z = (d.nz + tl.kz - 1) / tl.kz;
  tl.ntiles = tl.tiles_x * tl.tiles_y * tl.tiles_z;
  tl.nblocks = (tl.ntiles + 7) / 8 * 8;
  return tl;
}

struct TuneResult {
  int kz = 0;    // z-march depth of the one-node-per-thread kernel
  int kzw = 0;   // ... and of the wide kernel at width xw
  int xw = 0;    // fastest width (0: the one-node kernel)
};

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, TuneResult> g_tune_cache;  // (nx, ny, nz, nr)

}  // namespace

extern "C" {

int fs_sys_autotune(fs_sys* s, int nr, int reps, int* kz_out) {
  FS_GUARD_BEGIN
  if (!s) return fail(FS_EINVAL, "NULL argument");
  if (s->d.per) return fail(FS_EINVAL, "periodic operators run fs_sys_solve and fs_capacitance only (FS_PREC_FP64 or FS_PREC_MIXED)");
  if (nr < 1 || nr > fs::kMaxRhs) return fail(FS_EINVAL, "nr must be in 1..%d", fs::kMaxRhs);
  if (reps < 1) reps = 5;
  fs_ctx* ctx = s->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  const fs::Dims& d = s->d;
  const auto key = std::make_tuple(d.nx, d.ny, d.nz, nr);
  TuneResult best;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) best = it->second;
  }
  auto grow_partials = [&](int64_t nblocks) -> int {
    const int64_t need = nblocks * fs::kMaxRhs * 2 + 64 * (fs::kMaxRhs + 1) * fs::kMaxRhs;
    if (need <= s->partials_cap) return FS_OK;
    double* np = nullptr;
    if (int rc = dalloc(&np, (size_t)need)) return rc;
    dfree(s->partials);
    s->partials = np;
    s->partials_cap = need;
    return FS_OK;
  };
  const bool wide_ok = d.per == 0 && d.n + d.sxy < (int64_t)INT32_MAX;
  if (best.kz == 0) {
    if (int rc = ensure_ws(s, nr)) return rc;
    // candidates: z-march depths from one plane to 128, for the one-node
    // kernel and every width that divides the rows
    std::vector<int> cand;
    for (int kz = 1; kz <= 128; kz *= 2)
      if (kz <= d.nz) cand.push_back(kz);
    std::vector<int> widths{1};
    for (int w : {2, 3, 4})
      if (wide_ok && d.nx % w == 0) widths.push_back(w);
    int64_t need = 0;
    for (int w : widths)
      for (int kz : cand) need = std::max<int64_t>(need, tiling_for(d, kz, w).nblocks);
    if (int rc = grow_partials(need)) return rc;
    // every right-hand side live for the timed launches; the next solve
    // re-initialises the scalars
    for (int r = 0; r < fs::kMaxRhs; ++r) s->sc_host[r] = fs::CgScalars{};
    FS_HIP(hipMemcpy(s->sc, s->sc_host, sizeof(fs::CgScalars) * fs::kMaxRhs,
                     hipMemcpyHostToDevice));
    FS_HIP(hipMemsetAsync(s->p32, 0, sizeof(float) * nr * s->vstride, s->stream));
    float t_one = 1e30f, t_wide = 1e30f;
    for (int w : widths) {
      for (int kz : cand) {
        const fs::Tiling tl = tiling_for(d, kz, w);
        fs::launch_stencil_only(s, nr, tl, w);  // warm
        FS_HIP(hipEventRecord(ctx->ev0, s->stream));
        for (int k = 0; k < reps; ++k) fs::launch_stencil_only(s, nr, tl, w);
        FS_HIP(hipEventRecord(ctx->ev1, s->stream));
        FS_HIP(hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        FS_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        if (w == 1 && ms < t_one) {
          t_one = ms;
          best.kz = kz;
        } else if (w > 1 && ms < t_wide) {
          t_wide = ms;
          best.kzw = kz;
          best.xw = w;
        }
      }
    }
    if (!(t_wide < t_one)) best.xw = 0;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
  }
  const fs::Tiling tl = tiling_for(d, best.kz);
  if (int rc = grow_partials(tl.nblocks)) return rc;
  s->tl = tl;
  s->xw = best.xw;
  if (best.xw) {
    const fs::Tiling tw = tiling_for(d, best.kzw, best.xw);
    if (int rc = grow_partials(tw.nblocks)) return rc;
    s->tlw = tw;
  }
  // the 4-wide layout of the preconditioner and the deflated stencil follows
  // the tuned depth when the 4-wide stencil won
  s->use_x4 = best.xw == 4 ? 1 : 0;
  if (best.xw == 4) s->tl4 = s->tlw;
  const int best_kz = best.xw ? best.kzw : best.kz;
  if (kz_out) *kz_out = best_kz;
  return FS_OK;
  FS_GUARD_END
}

}  // extern "C"

// --------------------------------------------------------------------------
// batched extraction: many prepared systems in groups, one captured graph
// per group, each replayed by its own host thread
// --------------------------------------------------------------------------
struct fs_batch {
  fs_ctx* ctx = nullptr;
  std::vector<hipStream_t> streams;  // one per group
  hipEvent_t fork = nullptr;         // the systems' initial state is on the main stream
  std::vector<hipEvent_t> joins;     // a group's last replay
  std::vector<hipEvent_t> polled;    // [2 per group]: replays in flight
  struct Inst {
    fs_cap* cap;
    const fs_op* L;
  };
  std::vector<Inst> inst;
  fs::CgScalars* sc_all = nullptr;       // [ninst][kMaxRhs] device
  fs::CgScalars* sc_all_host = nullptr;  // pinned mirror
  fs::BatchCtl* ctl = nullptr;           // [ninst] device control
  fs::BatchCtl* ctl_host = nullptr;      // pinned mirror
  int* ndone = nullptr;                  // device [group]: finished systems
  unsigned long long* clock0 = nullptr;  // device wall clock at the start of a solve
  int* ndone_host = nullptr;             // pinned [2 per group]: two replays in flight
  int sc_cap = 0;
};

namespace {

void free_batch(fs_batch* b) {
  if (!b) return;
  for (auto st : b->streams) (void)hipStreamDestroy(st);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (auto e : b->joins) (void)hipEventDestroy(e);
  for (auto e : b->polled) (void)hipEventDestroy(e);
  dfree(b->sc_all);
  dfree(b->ctl);
  dfree(b->ndone);
  dfree(b->clock0);
  if (b->sc_all_host) (void)hipHostFree(b->sc_all_host);
  if (b->ctl_host) (void)hipHostFree(b->ctl_host);
  if (b->ndone_host) (void)hipHostFree(b->ndone_host);
  delete b;
}


// A group of a batch's systems: one captured graph on one stream, replayed
// by one host thread.
struct BatchGroup {
  std::vector<int> sys;
  double work = 0.0;
  int slot0 = 0;  // the group's systems take control slots slot0, slot0 + 1, ...
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipError_t err = hipSuccess;
  bool finished = false;
};

// ng groups of about equal work (nodes x right-hand sides, longest first to
// the least loaded group); slot[i]: system i's control slot, consecutive
// within a group so that one kernel counts the group's finished systems
std::vector<BatchGroup> batch_groups(const fs_batch* b, int ng, std::vector<int>& slot) {
  const int ni = (int)b->inst.size();
  std::vector<BatchGroup> grp(ng);
  std::vector<int> order(ni);
  for (int i = 0; i < ni; ++i) order[i] = i;
  auto work = [&](int i) {
    return (double)b->inst[i].cap->sys->d.n * (double)b->inst[i].cap->nt;
  };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
  for (int i : order) {
    BatchGroup* best = &grp[0];
    for (auto& g : grp)
      if (g.work < best->work) best = &g;
    best->sys.push_back(i);
    best->work += work(i);
  }
  slot.assign(ni, 0);
  for (int g = 0, k = 0; g < ng; ++g) {
    grp[g].slot0 = k;
    for (int i : grp[g].sys) slot[i] = k++;
  }
  return grp;
}

// the start of an FP32 cycle (and, for flexible CG, its first
// preconditioner application)
int batch_start_cycle(fs_sys* s, int nt, int mode, double inner_tol2) {
  fs::launch_cycle_start(s, nt, inner_tol2);
  if (mode == 2)
    if (int r = fs::launch_fcg_start(s, nt)) return r;
  return FS_OK;
}

// One group's graph on stream st: per system, `chunk` iterations, then the
// system's device control: when every running right-hand side has ended its
// FP32 cycle, the FP64 refresh and, if one is still running, the next
// cycle's start.  The gated kernels exit at once on the replays that do not
// need them, and finished systems skip everything, so the host only replays
// the graph until the device counts the group's systems finished.
int capture_group(fs_batch* b, BatchGroup& G, int g, const std::vector<int>& slot, int chunk,
                  int mode, double inner_tol2) {
  const hipStream_t st = b->streams[g];
  const hipStream_t main = b->ctx->stream;
  const bool cyc = mode != 0;
  FS_HIP(hipStreamWaitEvent(st, b->fork, 0));
  FS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int crc = FS_OK;
  for (int i : G.sys) {
    if (crc != FS_OK) break;
    fs_sys* s = b->inst[i].cap->sys;
    const int nt = b->inst[i].cap->nt;
    fs::BatchCtl* c = b->ctl + slot[i];
    s->stream = st;
    s->pc_count = &c->pc_applies;
    for (int k = 0; k < chunk && crc == FS_OK; ++k)
      crc = fs::launch_iteration(s, nt, mode, nullptr);
    if (crc == FS_OK && cyc) {
      fs::launch_batch_refresh_gate(s, nt, c);
      s->gate = &c->refresh_skip;
      fs::launch_refresh(s, nt);
    }
    if (crc == FS_OK) fs::launch_batch_start_gate(s, nt, cyc ? 1 : 0, c);
    if (crc == FS_OK && cyc) {
      s->gate = &c->start_skip;
      crc = batch_start_cycle(s, nt, mode, inner_tol2);
    }
    s->gate = nullptr;
    s->pc_count = nullptr;
    s->stream = main;
  }
  if (crc == FS_OK)
    fs::launch_batch_count(st, b->ctl + G.slot0, (int)G.sys.size(), b->ndone + g);
  const hipError_t ee = hipStreamEndCapture(st, &G.graph);
  if (crc != FS_OK) return crc;
  if (ee != hipSuccess) return fail(FS_EHIP, "graph capture failed: %s", hipGetErrorString(ee));
  FS_HIP(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
  return FS_OK;
}

// Replays group g's graph until the device counts its systems finished (at
// most max_rounds replays), two replays in flight: the finished count of
// replay k is read while replay k+1 runs; the replay queued behind the last
// one finds the group finished and exits kernel by kernel.  Runs on a host
// thread of its own (graph launches cost host time per node, so one thread
// would feed the device too slowly); errors go to G.err.
void replay_group(fs_batch* b, BatchGroup& G, int g, int max_rounds) {
  const hipStream_t st = b->streams[g];
  int* cnt = b->ndone_host + 2 * g;
  hipEvent_t* ev = b->polled.data() + 2 * g;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && G.err == hipSuccess) G.err = e;
    return e == hipSuccess;
  };
  if (!chk(hipSetDevice(b->ctx->device))) return;
  auto replay = [&](int k) {
    return chk(hipGraphLaunch(G.exec, st)) &&
           chk(hipMemcpyAsync(cnt + (k & 1), b->ndone + g, sizeof(int), hipMemcpyDeviceToHost,
                              st)) &&
           chk(hipEventRecord(ev[k & 1], st));
  };
  const int want = (int)G.sys.size();
  if (!replay(0)) return;
  for (int k = 1; k <= max_rounds && !G.finished; ++k) {
    if (k < max_rounds && !replay(k)) return;
    if (!chk(hipEventSynchronize(ev[(k - 1) & 1]))) return;
    G.finished = cnt[(k - 1) & 1] >= want;
  }
  (void)chk(hipEventRecord(b->joins[g], st));
}
}  // namespace

extern "C" {

int fs_batch_create(fs_ctx* ctx, int nstreams, fs_batch** out) {
  FS_GUARD_BEGIN
  if (!ctx || !out) return fail(FS_EINVAL, "NULL argument");
  if (nstreams < 1 || nstreams > 64) return fail(FS_EINVAL, "nstreams must be in 1..64");
  *out = nullptr;
  FS_HIP(hipSetDevice(ctx->device));
  fs_batch* b = new fs_batch();
  b->ctx = ctx;
  for (int k = 0; k < nstreams; ++k) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      free_batch(b);
      return fail(FS_EHIP, "stream creation failed");
    }
    b->streams.push_back(st);
    for (int e = 0; e < 3; ++e) {
      hipEvent_t ev = nullptr;
      if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        free_batch(b);
        return fail(FS_EHIP, "event creation failed");
      }
      (e == 0 ? b->joins : b->polled).push_back(ev);
    }
  }
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
    free_batch(b);
    return fail(FS_EHIP, "event creation failed");
  }
  *out = b;
  return FS_OK;
  FS_GUARD_END
}

int fs_batch_destroy(fs_batch* b) {
  if (b) (void)hipSetDevice(b->ctx->device);
  free_batch(b);
  return FS_OK;
}

int fs_batch_add(fs_batch* b, fs_cap* cap, const fs_op* L) {
  FS_GUARD_BEGIN
  if (!b || !cap || !L) return fail(FS_EINVAL, "NULL argument");
  fs_sys* s = cap->sys;
  if (s->ctx != b->ctx) return fail(FS_EINVAL, "system on another context");
  if (s->d.per) return fail(FS_EINVAL, "periodic operators run fs_sys_solve and fs_capacitance only (FS_PREC_FP64 or FS_PREC_MIXED)");
  if (cap->nt > fs::kMaxRhs)
    return fail(FS_EINVAL, "a batched extraction holds at most %d terminals", fs::kMaxRhs);
  if (L->Nx != s->op->Nx || L->Ny != s->op->Ny || L->Nz != s->op->Nz)
    return fail(FS_EINVAL, "L and the system differ in grid");
  for (const auto& in : b->inst)
    if (in.cap->sys == s) return fail(FS_EINVAL, "system already in the batch");
  if (s->defl_n > 0) return fail(FS_EINVAL, "systems with an island space are not batched");
  b->inst.push_back({cap, L});
  return (int)b->inst.size() - 1 >= 0 ? FS_OK : FS_EINVAL;
  FS_GUARD_END
}

int fs_batch_solve(fs_batch* b, const fs_solve_params* params, int chunk, double* C,
                   double* inst_time, int* inst_iters, fs_solve_stats* stats) {
  FS_GUARD_BEGIN
  if (!b || !params || !C) return fail(FS_EINVAL, "NULL argument");
  if (!(params->tol > 0.0)) return fail(FS_EINVAL, "tol must be > 0");
  const int P = params->precision;
  if (P != FS_PREC_FP64 && P != FS_PREC_MIXED && P != FS_PREC_FCG)
    return fail(FS_EINVAL, "unknown precision %d", P);
  if (chunk < 1 || chunk > 64) return fail(FS_EINVAL, "chunk must be in 1..64");
  const int ni = (int)b->inst.size();
  if (ni == 0) return fail(FS_EINVAL, "empty batch");
  fs_ctx* ctx = b->ctx;
  FS_HIP(hipSetDevice(ctx->device));
  const hipStream_t main = ctx->stream;
  const int mode = P == FS_PREC_FP64 ? 0 : (P == FS_PREC_MIXED ? 1 : 2);
  const bool cyc = mode != 0;
  const double inner_tol = params->inner_tol > 0 ? params->inner_tol : (mode == 2 ? kFcgInnerTol : 1e-4);
  for (const auto& in : b->inst)
    if (mode == 2 && !in.cap->sys->pc)
      return fail(FS_EINVAL, "FS_PREC_FCG needs a preconditioner on every system of the batch");
  if (b->sc_cap < ni) {
    dfree(b->sc_all);
    dfree(b->ctl);
    if (b->sc_all_host) (void)hipHostFree(b->sc_all_host);
    if (b->ctl_host) (void)hipHostFree(b->ctl_host);
    b->sc_all_host = nullptr;
    b->ctl_host = nullptr;
    b->sc_cap = 0;
    if (int rc = dalloc(&b->sc_all, (size_t)ni * fs::kMaxRhs)) return rc;
    if (int rc = dalloc(&b->ctl, (size_t)ni)) return rc;
    if (hipHostMalloc((void**)&b->sc_all_host, sizeof(fs::CgScalars) * ni * fs::kMaxRhs) !=
        hipSuccess) {
      b->sc_all_host = nullptr;
      return fail(FS_ENOMEM, "pinned allocation failed");
    }
    if (hipHostMalloc((void**)&b->ctl_host, sizeof(fs::BatchCtl) * ni) != hipSuccess) {
      b->ctl_host = nullptr;
      return fail(FS_ENOMEM, "pinned allocation failed");
    }
    b->sc_cap = ni;
  }
  const int ng = std::min((int)b->streams.size(), ni);
  if (!b->ndone_host) {  // allocated last: set only when all three are
    const int nst = (int)b->streams.size();
    dfree(b->ndone);
    dfree(b->clock0);
    if (int rc = dalloc(&b->ndone, (size_t)nst)) return rc;
    if (int rc = dalloc(&b->clock0, 1)) return rc;
    if (hipHostMalloc((void**)&b->ndone_host, 2 * sizeof(int) * nst) != hipSuccess) {
      b->ndone_host = nullptr;
      return fail(FS_ENOMEM, "pinned allocation failed");
    }
  }
  int clock_khz = 0;
  FS_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device));
  if (clock_khz <= 0) return fail(FS_EHIP, "device wall clock rate unavailable");
  FS_HIP(hipMemsetAsync(b->sc_all, 0, sizeof(fs::CgScalars) * ni * fs::kMaxRhs, main));
  FS_HIP(hipMemsetAsync(b->ctl, 0, sizeof(fs::BatchCtl) * ni, main));
  // workspaces first (an allocation failure returns before any system is
  // changed), then lend every system the batch's scalars and the main stream
  std::vector<fs::CgScalars*> own_sc(ni);
  std::vector<int> maxit(ni);
  for (int i = 0; i < ni; ++i) {
    fs_sys* s = b->inst[i].cap->sys;
    if (int rc = ensure_ws(s, b->inst[i].cap->nt)) return rc;
    maxit[i] = default_maxiter(s, params);
  }
  for (int i = 0; i < ni; ++i) {
    fs_sys* s = b->inst[i].cap->sys;
    own_sc[i] = s->sc;
    s->sc = b->sc_all + (size_t)i * fs::kMaxRhs;
    s->stream = main;
    s->acc_fallbacks = 0;
    s->acc_pc_applies = 0;
    s->acc_replacements = 0;
  }
  auto restore = [&]() {
    for (int i = 0; i < ni; ++i) {
      fs_sys* s = b->inst[i].cap->sys;
      s->sc = own_sc[i];
      s->stream = ctx->stream;
      s->gate = nullptr;
      s->pc_count = nullptr;
    }
  };
  std::vector<std::vector<double>> hout(ni);
  std::vector<double> fallback_time(ni, -1.0);
  std::vector<int> slot;
  std::vector<BatchGroup> grp = batch_groups(b, ng, slot);
  int rc = FS_OK;
  auto t0 = std::chrono::steady_clock::now();
  auto destroy_graphs = [&]() {
    for (int g = 0; g < ng; ++g)
      if (grp[g].exec) (void)hipStreamSynchronize(b->streams[g]);  // replays still queued
    for (auto& g : grp) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      if (g.graph) (void)hipGraphDestroy(g.graph);
      g.exec = nullptr;
      g.graph = nullptr;
    }
  };
  auto body = [&]() -> int {
    // initial state of every system (lift, zero guess, first cycle)
    fs::launch_batch_clock(main, b->clock0);
    for (int i = 0; i < ni; ++i) {
      fs_cap* cap = b->inst[i].cap;
      fs_sys* s = cap->sys;
      fs::launch_lift_cap(s, cap->nt, b->inst[i].L, cap->tid, 0);
      fs::launch_init(s, cap->nt, params->tol, maxit[i]);
      if (cyc) {
        if (int r = batch_start_cycle(s, cap->nt, mode, inner_tol * inner_tol)) return r;
        s->acc_pc_applies += mode == 2 ? 1 : 0;
      }
    }
    FS_HIP(hipGetLastError());
    FS_HIP(hipEventRecord(b->fork, main));
    for (int g = 0; g < ng; ++g)
      if (int r = capture_group(b, grp[g], g, slot, chunk, mode, inner_tol * inner_tol)) return r;
    int max_rounds = 0;
    for (int i = 0; i < ni; ++i) max_rounds = std::max(max_rounds, maxit[i]);
    max_rounds = 4 * (max_rounds / chunk + 2) + 64;
    bool spawned = true;
    if (ng == 1) {
      replay_group(b, grp[0], 0, max_rounds);  // one group: the calling thread
    } else {
      std::vector<std::thread> workers;
      workers.reserve(ng);
      try {
        for (int g = 0; g < ng; ++g)
          workers.emplace_back(replay_group, b, std::ref(grp[g]), g, max_rounds);
      } catch (const std::exception&) {
        spawned = false;  // the threads already running finish their groups first
      }
      for (auto& w : workers) w.join();
    }
    if (!spawned) return fail(FS_EHIP, "could not start the batch's replay threads");
    for (int g = 0; g < ng; ++g) {
      if (grp[g].err != hipSuccess)
        return fail(FS_EHIP, "batch replay failed: %s", hipGetErrorString(grp[g].err));
      if (!grp[g].finished) return fail(FS_EINVAL, "batch driver exceeded its round guard");
      FS_HIP(hipStreamWaitEvent(main, b->joins[g], 0));
    }
    FS_HIP(hipMemcpyAsync(b->sc_all_host, b->sc_all, sizeof(fs::CgScalars) * ni * fs::kMaxRhs,
                          hipMemcpyDeviceToHost, main));
    FS_HIP(hipMemcpyAsync(b->ctl_host, b->ctl, sizeof(fs::BatchCtl) * ni,
                          hipMemcpyDeviceToHost, main));
    FS_HIP(hipStreamSynchronize(main));
    // per system: the FP64 redo of a flexible-CG solve that failed, the last
    // FP64 step and true-residual check, then the charges
    for (int i = 0; i < ni; ++i) {
      fs_cap* cap = b->inst[i].cap;
      fs_sys* s = cap->sys;
      const int nt = cap->nt;
      const fs::CgScalars* c = b->sc_all_host + (size_t)i * fs::kMaxRhs;
      s->acc_pc_applies += b->ctl_host[slot[i]].pc_applies;
      bool ok = true;
      for (int r = 0; r < nt; ++r) ok = ok && c[r].done == 1;
      if (mode == 2 && !ok) {
        // plain FP64 CG redoes this system (blocking)
        fs_solve_params Q = *params;
        Q.precision = FS_PREC_FP64;
        Q.maxiter = 0;
        s->acc_fallbacks += 1;
        fs::launch_init(s, nt, params->tol, default_maxiter(s, &Q));
        if (int r = run_krylov_mode(s, nt, Q, 0)) return r;
        FS_HIP(hipStreamSynchronize(main));
        fallback_time[i] =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      } else if (mode == 0) {
        // the pending x += alpha p, and the true-residual check (blocking)
        if (int r = fp64_finish(s, nt, *params)) return r;
      }
      fs::launch_residual(s, nt, s->r);
      fs::launch_expand(s, nt, nullptr, cap->tid, 0, s->q);
      fs::launch_charge(s, nt, b->inst[i].L, s->q, cap->nodes, cap->offs, cap->nbins,
                        cap->nbx, cap->out);
      hout[i].assign((size_t)cap->nbins * fs::kMaxRhs, 0.0);
      FS_HIP(hipMemcpyAsync(hout[i].data(), cap->out, sizeof(double) * cap->nbins * nt,
                            hipMemcpyDeviceToHost, main));
      FS_HIP(hipGetLastError());
    }
    return FS_OK;
  };
  rc = body();
  destroy_graphs();
  unsigned long long clock0 = 0;
  if (rc == FS_OK) {
    // final scalars (residuals of the finishing pass)
    if (hipMemcpy(b->sc_all_host, b->sc_all, sizeof(fs::CgScalars) * ni * fs::kMaxRhs,
                  hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&clock0, b->clock0, sizeof(clock0), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(FS_EHIP, "scalar download failed");
  }
  restore();
  if (rc != FS_OK) return rc;
  const double total =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (inst_time)
    for (int i = 0; i < ni; ++i)
      inst_time[i] = fallback_time[i] >= 0.0
                         ? fallback_time[i]
                         : (double)(b->ctl_host[slot[i]].stamp - clock0) / (1e3 * clock_khz);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->converged = 1;
    stats->solve_time = total;
    stats->stencil_width = 4;
    stats->cycle_width = 4;
  }
  int status = FS_OK;
  size_t off = 0;
  for (int i = 0; i < ni; ++i) {
    fs_cap* cap = b->inst[i].cap;
    fs_sys* s = cap->sys;
    const int nt = cap->nt;
    for (int r = 0; r < nt; ++r)
      for (int q = 0; q < nt; ++q) C[off + (size_t)q * nt + r] = hout[i][(size_t)q * nt + r];
    off += (size_t)nt * nt;
    // per-system verdict from the batch's scalars
    std::memcpy(s->sc_host, b->sc_all_host + (size_t)i * fs::kMaxRhs, sizeof(fs::CgScalars) * nt);
    int st = FS_OK;
    fs_solve_stats one{};
    collect_stats(s, nt, &one, st);
    if (inst_iters) inst_iters[i] = one.iterations;
    if (stats) {
      stats->iterations = std::max(stats->iterations, one.iterations);
      stats->refreshes = std::max(stats->refreshes, one.refreshes);
      stats->residual = std::max(stats->residual, one.residual);
      stats->residual_eq = std::max(stats->residual_eq, one.residual_eq);
      stats->fallbacks += one.fallbacks;
      stats->pc_applies += one.pc_applies;
      stats->replacements += one.replacements;
      stats->stop_reason = std::max(stats->stop_reason, one.stop_reason);
      stats->maxiter = std::max(stats->maxiter, one.maxiter);
      // the narrowest kernels any system of the batch ran
      kernel_widths(s, P, &one);
      stats->stencil_width = std::min(stats->stencil_width, one.stencil_width);
      stats->cycle_width = std::min(stats->cycle_width, one.cycle_width);
      if (st != FS_OK) stats->converged = 0;
    }
    if (st != FS_OK) status = st;
  }
  if (status == FS_NOT_CONVERGED)
    fail(FS_NOT_CONVERGED, "a batched extraction did not reach tol %.3g", params->tol);
  return status;
  FS_GUARD_END
}

}  // extern "C"
