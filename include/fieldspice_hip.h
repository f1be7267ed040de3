/*
 * fieldspice_hip.h -- C ABI of the MI355X (gfx950) elliptic field solve.
 *
 * The library solves the node-Laplacian systems fieldspice builds as
 *     A = G^T diag(w) G + diag(extra)          (w >= 0 per edge)
 * with Dirichlet rows eliminated symmetrically, entirely on one GPU and
 * matrix-free: the operator is stored as three per-edge weight arrays, never
 * as a CSR matrix.
 *
 * Conventions (all entry points):
 *   - Host arrays use fieldspice's layout: C order over (i, j, k) with k (z)
 *     fastest.  Node arrays have (Nx+1)*(Ny+1)*(Nz+1) entries; x-edge arrays
 *     (Nx, Ny+1, Nz+1), y-edge (Nx+1, Ny, Nz+1), z-edge (Nx+1, Ny+1, Nz).
 *     The library permutes to its own x-fastest device layout on upload and
 *     back on download.
 *   - Host pointers are caller-owned and only read (or written, for outputs)
 *     during the call.  Inputs are never modified.
 *   - Return value: FS_OK (0), FS_NOT_CONVERGED (1), or a negative error code;
 *     fs_last_error() describes the most recent failure of the calling thread.
 *   - A context and the objects created from it may be used by one host
 *     thread at a time.  Work runs on the context's own stream (default
 *     priority).  No C++ exception crosses this boundary.
 */
#ifndef FIELDSPICE_HIP_H
#define FIELDSPICE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS_OK 0
#define FS_NOT_CONVERGED 1
#define FS_EINVAL (-1)
#define FS_EHIP (-2)
#define FS_ENOMEM (-3)

/* cell -> edge averaging rules of fieldspice/operators.py:144-237 */
#define FS_AVG_PARALLEL 0
#define FS_AVG_HARMONIC 1
#define FS_AVG_MIN 2
#define FS_AVG_MAX 3

/* node mask values (fs_sys_create) */
#define FS_NODE_FREE 0
#define FS_NODE_DIRICHLET 1

/* precision of the Krylov iteration */
#define FS_PREC_FP64 0          /* FP64 stencil and vectors throughout */
#define FS_PREC_MIXED 1         /* FP32 stencil/vectors, FP64 residual refresh and dots */
#define FS_PREC_FCG 2           /* FS_PREC_MIXED cycles of flexible CG preconditioned by the
                                   system's attached fs_pc (fs_sys_set_precond); a run that
                                   breaks down or hits maxiter is redone by FP64 CG */

typedef struct fs_ctx fs_ctx;
typedef struct fs_op fs_op;
typedef struct fs_sys fs_sys;
typedef struct fs_cap fs_cap;
typedef struct fs_net fs_net;
typedef struct fs_pc fs_pc;

typedef struct {
    double tol;          /* relative residual ||r_s|| < tol*||b_s|| in the equilibrated space */
    int maxiter;         /* <= 0: max(500, floor(10*sqrt(n_free))) as the reference */
    int precision;       /* FS_PREC_FP64, FS_PREC_MIXED or FS_PREC_FCG */
    double inner_tol;    /* MIXED/FCG: residual drop per FP32 cycle before an FP64 refresh
                            (<= 0: 1e-4 MIXED, 1e-3 FCG) */
    int check_every;     /* iterations between host convergence checks (<= 0: adaptive) */
} fs_solve_params;

typedef struct {
    int iterations;      /* Krylov iterations (maximum over right-hand sides) */
    int refreshes;       /* FP64 residual recomputations (mixed precision) */
    int converged;       /* 1 when every right-hand side met tol */
    double residual;     /* fieldspice's info["residual"]: ||D^1/2 (b_s - A_s x_s)|| / ||b|| */
    double residual_eq;  /* FP64 recomputed ||b_s - A_s x_s|| / ||b_s|| (max over RHS) */
    double solve_time;   /* seconds, device work of this call */
    /* with fs_sys_set_profiling(sys, 1): the stencil kernel of the Krylov loop */
    double stencil_time;      /* seconds summed over its launches (HIP events) */
    int64_t stencil_launches; /* launches, including ones that found no live RHS */
    double stencil_bytes;     /* algorithmic bytes: couplings once + p read + q write per live RHS */
    int fallbacks;       /* FS_PREC_FCG runs redone by FP64 CG */
    int pc_applies;      /* preconditioner applications (FS_PREC_FCG) */
    int stop_reason;     /* worst over right-hand sides: 1 converged, 2 iteration cap,
                            3 breakdown (non-finite or non-positive curvature) */
    int maxiter;         /* iteration cap the solve ran with */
    int replacements;    /* FS_PREC_FP64: restarts from the recomputed FP64 residual after
                            the recursive residual met tol but the true one did not */
    int stencil_width;   /* nodes per lane of the Krylov stencil kernel that ran:
                            FP64 1 or 2 (even rows); FP32 1, or 2, 3 or 4 (a width that
                            divides the rows: the widest by default, the fastest after
                            fs_sys_autotune) */
    int cycle_width;     /* FS_PREC_FCG: nodes per lane of the level-0 cycle kernels (1 or 4);
                            0 without a preconditioner */
    /* with fs_sys_set_profiling(sys, 1): every full-grid kernel of a Krylov
     * iteration, indexed by FS_PROF_*: seconds summed over its launches (HIP
     * events at the kernel's own start and end), launches, and algorithmic
     * bytes (each array once per launch: couplings and masks once, the
     * vectors of every live right-hand side).  Kinds a method does not run
     * stay 0.  Index FS_PROF_STENCIL repeats stencil_time/launches/bytes. */
    double kernel_time[5];
    int64_t kernel_launches[5];
    double kernel_bytes[5];
} fs_solve_stats;

/* fs_solve_stats.kernel_* index */
#define FS_PROF_STENCIL   0  /* q = A_s p with the p.q partials */
#define FS_PROF_DIRECTION 1  /* p = z + beta p, carrying the deferred x += alpha p */
#define FS_PROF_RESIDUAL  2  /* r -= alpha q with the r.r partials */
#define FS_PROF_RESTRICT  3  /* FS_PREC_FCG: level-0 pre-smoothing and restriction */
#define FS_PROF_PROLONG   4  /* FS_PREC_FCG: level-0 prolongation, post-smoothing, <z,r>, <z,q> */
#define FS_PROF_KINDS     5

typedef struct {
    int64_t n_nodes;
    int64_t n_free;      /* unknowns carried into the Krylov solve */
    int64_t n_fixed;     /* Dirichlet rows */
    int64_t n_empty;     /* rows coupled to nothing, pinned to 0 */
    int64_t gauge;       /* node pinned to fix the gauge of a floating system, or -1 */
    int singular;        /* 1 if the operator had the constant null vector */
} fs_sys_info;

/* ---- context ------------------------------------------------------------ */
int fs_ctx_create(int device, fs_ctx** out);
int fs_ctx_destroy(fs_ctx* ctx);
const char* fs_last_error(void);
/* device name and compute-unit count, for reports */
int fs_ctx_device_info(fs_ctx* ctx, char* name, int name_len, int* n_cu);

/* Measurement aid, no counterpart in the reference: seconds per launch of a
 * kernel that streams nread + nwrite FP32 arrays of n elements once (16-byte
 * accesses, non-temporal stores, grid-stride), averaged over reps launches
 * after one warm-up.  It prices a kernel's read/write mix at the rate this
 * device streams it (nread = 3 + k, nwrite = k is the FP32 stencil's mix for
 * k right-hand sides, k = 1..8; 1/1 is a copy) and calibrates hardware
 * traffic counters against exactly n*4*(nread + nwrite) bytes.  Allocates and
 * frees its own buffers; n a positive multiple of 4. */
int fs_stream_probe(fs_ctx* ctx, int64_t n, int nread, int nwrite, int reps, double* seconds);

/* ---- operators: A = G^T diag(w) G + diag(extra) ------------------------- */
/* Replaces edge_mass + nodal_laplacian (fieldspice/operators.py:302-315,
 * 340-351) for given edge weights.  extra may be NULL. */
int fs_op_create_edges(fs_ctx* ctx, int Nx, int Ny, int Nz,
                       const double* w_x, const double* w_y, const double* w_z,
                       const double* extra, fs_op** out);

/* Replaces cell_to_edge + edge_mass (fieldspice/operators.py:144-237,302-315):
 * w_e = avg(prop over the <=4 cells around e) * A_dual(e) / L(e), built on the
 * device from per-cell prop (Nx,Ny,Nz) and the primal cell widths hx,hy,hz. */
int fs_op_create_cells(fs_ctx* ctx, int Nx, int Ny, int Nz,
                       const double* hx, const double* hy, const double* hz,
                       const double* prop_cell, int avg_mode, fs_op** out);

/* Operator of a grid whose opposite walls are identified along the axes in
 * `periodic` (bit 0 x, 1 y, 2 z): the merged node planes of fieldspice's
 * periodic reduction (fieldspice/solvers/poisson.py:665-719,942-956:
 * P^T L P with each wall pair one unknown), given on the reduced grid of
 * nx*ny*nz nodes (a periodic axis keeps one of its two wall planes, so nx is
 * the cell count along a periodic x).  Edge arrays in C order: along a
 * periodic axis there are as many edges as nodes, the last one joining the
 * last node to the first; along the others one fewer.  A periodic axis needs
 * >= 2 nodes.  Such operators run fs_op_apply, fs_sys_create, fs_sys_solve
 * and fs_capacitance (FS_PREC_FP64, FS_PREC_MIXED); the other entry points
 * reject them. */
int fs_op_create_periodic(fs_ctx* ctx, int nx, int ny, int nz, const double* w_x,
                          const double* w_y, const double* w_z, int periodic, fs_op** out);

/* out = a*A + b*B + diag(c*extra): the EQS/thermal step matrices
 * (fieldspice/solvers/eqs.py:279-280, thermal.py:259-263).  B and extra may
 * be NULL. */
int fs_op_combine(double a, const fs_op* A, double b, const fs_op* B,
                  double c, const double* extra, fs_op** out);
int fs_op_destroy(fs_op* op);

/* y = A x (node vectors, host layout): L @ phi (fieldspice/solvers/poisson.py:1219-1221) */
int fs_op_apply(fs_op* op, const double* x, double* y);
/* edge weights back to host (edge layouts above), for inspection */
int fs_op_get_edges(fs_op* op, double* w_x, double* w_y, double* w_z);

/* ---- prepared solves: LinearSystem (fieldspice/solvers/poisson.py:260-436) */
/* mask: per node FS_NODE_FREE / FS_NODE_DIRICHLET (host layout), may be NULL.
 * Dirichlet rows become identity rows (apply_dirichlet, operators.py:362-407).
 * Empty rows are pinned to 0; a floating system (no Dirichlet row, zero row
 * sums) gets its gauge node pinned when pin != 0, else FS_EINVAL. */
int fs_sys_create(fs_op* op, const uint8_t* mask, int pin, fs_sys** out,
                  fs_sys_info* info);
int fs_sys_destroy(fs_sys* sys);

/* LinearSystem.solve (fieldspice/solvers/poisson.py:487-589) for nrhs
 * right-hand sides at once: b and x are nrhs node vectors back to back.
 * b holds the Dirichlet values on Dirichlet rows (the output of
 * apply_dirichlet); x0 (may be NULL) holds nrhs initial guesses, of which
 * the free entries start the iteration (poisson.py:543-547).  Returns
 * FS_NOT_CONVERGED when maxiter is reached. */
int fs_sys_solve(fs_sys* sys, int nrhs, const double* b, const double* x0,
                 double* x, const fs_solve_params* params, fs_solve_stats* stats);

/* fs_sys_solve that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and status codes as fs_sys_solve; the Krylov workspace,
 * the initial guesses and the upload staging live in a static 256 MiB array
 * per GPU, which needs 68 bytes per node and right-hand side (76 with x0),
 * for up to 8 right-hand sides at a time (nrhs above 8 run in batches of 8):
 * larger requests fail with FS_EINVAL.  Every call on one GPU shares that
 * GPU's array, so calls that overlap on one GPU (several contexts or threads)
 * race; serialise them.  The system itself (and its preconditioner or island
 * space) must have been created beforehand. */
int fs_sys_solve_noalloc(fs_sys* sys, int nrhs, const double* b, const double* x0,
                         double* x, const fs_solve_params* params, fs_solve_stats* stats);

/* PoissonSolver.capacitance_matrix inner loop
 * (fieldspice/solvers/poisson.py:1208-1223): sys was created from op L with
 * every terminal and shield node marked Dirichlet.  term_id per node:
 * -1 free, 0..nt-1 terminal, nt = grounded shield.  For each j, terminal j at
 * 1 V and all other fixed nodes at 0 V are solved (all nt at once) and
 * C[i*nt + j] = sum over terminal i of (L phi_j), q_shield[j] the shield
 * charge.  C is returned unsymmetrised. phi_out (nt node vectors) may be
 * NULL. */
int fs_capacitance(fs_sys* sys, const fs_op* L, int nt, const int32_t* term_id,
                   double* C, double* q_shield, double* phi_out,
                   const fs_solve_params* params, fs_solve_stats* stats);

/* The same extraction split into a setup (terminal map and node lists
 * resident on the device) and repeated solves, so that repeated extractions
 * of one geometry pay the setup once. */
int fs_cap_prepare(fs_sys* sys, int nt, const int32_t* term_id, fs_cap** out);
int fs_cap_solve(fs_cap* cap, const fs_op* L, double* C, double* q_shield, double* phi_out,
                 const fs_solve_params* params, fs_solve_stats* stats);
int fs_cap_destroy(fs_cap* cap);

/* ---- implicit time loops: EQS and thermal steps ------------------------- */
/* One field x advanced by
 *     K x_k = Aexp x_{k-1} + sum_s src_scale[k][s] * src[s]      (free rows)
 *     x_k[n] = values[k][group[n]]                              (Dirichlet rows)
 * for k = 1..nsteps, where K is the matrix of sys.  This is the per-step loop
 * of EQSSolver.solve (fieldspice/solvers/eqs.py:291-307) and
 * ThermalSolver.transient (fieldspice/solvers/thermal.py:271-285), with the
 * solve, the explicit product, the Dirichlet lifting and the recorded
 * quantities on the GPU.  Arrays are host layout. */
typedef struct {
    int nsteps;
    int ngroups;               /* Dirichlet groups (terminals, walls) */
    const int32_t* group;      /* per node: -1 not Dirichlet, else group id; must equal sys's mask */
    const double* values;      /* [(nsteps+1) * ngroups] group values per time point */
    int nsrc;                  /* source vectors */
    const double* src;         /* [nsrc * n_nodes] */
    const double* src_scale;   /* [(nsteps+1) * nsrc] */
    int warm_start;            /* initial guess of each solve: 0 zero (the reference's),
                                  1 the previous field, 2 from the second step on the
                                  linear extrapolation 2 x_{k-1} - x_{k-2}, 3 the
                                  A-norm projection onto a basis of earlier solutions
                                  (proj_dim vectors; the first step as 2) */
    int nprobe;                /* recorded terminals */
    const int32_t* probe;      /* per node: -1 or probe id */
    const double* probe_w;     /* per node weight of the probe voltage (sum over probe = 1) */
    const fs_op* cur_a;        /* I_m(k) = sum_{n in m} [cur_a (x_k - x_{k-1}) + cur_b x_k]_n */
    const fs_op* cur_b;        /* may be NULL */
    const fs_op* quad_a;       /* per-step x^T quad_a x (may be NULL) */
    const fs_op* quad_b;       /* per-step x^T quad_b x (may be NULL) */
    int store_every;           /* copy x_k to `stored` every store_every steps (0: never) */
    const double* x_before;    /* x_{-1}, the field one step before x_0 (host layout, may
                                  be NULL): with warm_start 2 the first step extrapolates
                                  from it too */
    double* x_before_out;      /* out: x_{nsteps-1} (host layout, may be NULL) */
    int proj_dim;              /* warm_start 3: basis vectors, 2..16 (0: 8) */
} fs_transient_desc;

/* x: in x_0, out x_nsteps.  Outputs (NULL to skip):
 *   stored      [nstore * n_nodes], nstore = nsteps/store_every + 1 (x_0 first)
 *   probe_i     [(nsteps+1) * nprobe] currents (row 0 copies row 1)
 *   probe_v     [(nsteps+1) * nprobe] weighted probe voltages
 *   step_stats  [(nsteps+1) * 4]: sum(x), max(x), x^T quad_a x, x^T quad_b x */
int fs_transient_run(fs_sys* sys, const fs_op* Aexp, const fs_transient_desc* desc,
                     double* x, double* stored, double* probe_i, double* probe_v,
                     double* step_stats, const fs_solve_params* params,
                     fs_solve_stats* stats);

/* Joule heat per node: q_n = 1/2 sum_{e at n} w_e (x_n - x_m)^2
 * (joule_heating_nodes, fieldspice/solvers/thermal.py:66-87). */
int fs_op_edge_power(fs_op* op, const double* x, double* q);

/* Record HIP events around every stencil launch of the Krylov loop of this
 * system and report their summed duration in fs_solve_stats (off by default). */
int fs_sys_set_profiling(fs_sys* sys, int on);

/* ---- learned multilevel preconditioner ----------------------------------
 * No counterpart in the reference, whose CG takes ILU or Jacobi
 * (fieldspice/solvers/poisson.py:466-484); it stands beside them for
 * FS_PREC_FCG.  fs_pc holds an aggregation hierarchy of one prepared system
 * (per-axis groupings of consecutive nodes, Galerkin 7-point coarse
 * operators, damped-Jacobi smoothing) whose bottom level is solved by an
 * fs_net: a 3-level convolutional network (16/32/64 channels, 3x3x3, BF16
 * weights and activations, FP32 accumulation on MFMA) that maps the bottom
 * residual to an approximate error.  Without a network the bottom takes one
 * Jacobi step.  Vectors of fs_pc_apply live in the system's equilibrated
 * space (r_s = D^-1/2 r, x_s = D^1/2 x), C order. */
int fs_net_create(fs_ctx* ctx, uint64_t seed, fs_net** out);
int fs_net_destroy(fs_net* net);
int fs_net_num_params(const fs_net* net, int64_t* count);
/* flat FP32 parameters: per layer W[27][cin][cout] (tap = (dz+1)*9 + (dy+1)*3
 * + (dx+1), x fastest), then b[cout] */
int fs_net_get_params(fs_net* net, float* params);
int fs_net_set_params(fs_net* net, const float* params);
/* The network alone on an nx*ny*nz voxel grid (x fastest), batch nb:
 * X [nb][n][16] -> Y [nb][n][16] (BF16 values returned as float).  With dY
 * non-NULL also grad[num_params] = d(sum dY*Y)/d(params). */
int fs_net_eval(fs_net* net, int nx, int ny, int nz, int nb, const float* X, float* Y,
                const float* dY, float* grad);

/* groups: for each level l < nlevels, three arrays (x, y, z) mapping the
 * level-l index along that axis to its level-(l+1) index (0, then steps of 0
 * or 1), concatenated.  net may be NULL.  The fs_pc must not outlive sys. */
int fs_pc_create(fs_sys* sys, fs_net* net, int nlevels, const int32_t* groups, double omega,
                 fs_pc** out);
int fs_pc_destroy(fs_pc* pc);
/* level dims (l = 0 .. nlevels) */
int fs_pc_level_dims(fs_pc* pc, int level, int* nx, int* ny, int* nz);
/* attach (or detach with NULL) the preconditioner used by FS_PREC_FCG */
int fs_sys_set_precond(fs_sys* sys, fs_pc* pc);
/* z = M(r) for nrhs equilibrated vectors */
int fs_pc_apply(fs_pc* pc, int nrhs, const double* r, double* z);
/* Training pairs from exact errors e (node space, C order, as differences of
 * solutions): the library forms e_s = D^1/2 e, r_s = A_s e_s in FP64 and keeps
 * both on the device in FP32. */
int fs_pc_add_pairs(fs_pc* pc, int npairs, const double* e);
int fs_pc_num_pairs(fs_pc* pc, int* npairs);
int fs_pc_clear_pairs(fs_pc* pc);
/* Training losses of one pair (e_s, r_s = A_s e_s):
 *   FS_LOSS_CYCLE   L = ||e_s - M(r_s)||_A^2 / ||e_s||_A^2 (the whole cycle)
 *   FS_LOSS_BOTTOM  L = (y^T A_K y - 2 y^T r_K) / (r_K^T D_K^-1 r_K), with r_K
 *                   the residual the down-sweep hands to the bottom and y the
 *                   network's bottom error (minimal at y = A_K^-1 r_K) */
#define FS_LOSS_CYCLE 0
#define FS_LOSS_BOTTOM 1
/* loss[b] of nb errors (node space) and the gradient of mean_b loss[b] with
 * respect to the network parameters */
int fs_pc_loss_grad(fs_pc* pc, int kind, int nb, const double* e, double* loss, float* grad);
/* Adam steps on the resident pairs (batches of `batch` <= 8 drawn with the
 * seed, random signs); losses[step] = mean batch loss before the update. */
int fs_pc_train(fs_pc* pc, int kind, int steps, int batch, double lr, uint64_t seed,
                double* losses);

/* ---- tiles and batches ----------------------------------------------------
 * No counterpart in the reference, which solves one system at a time on the
 * CPU (fieldspice/solvers/poisson.py:1100-1249 per extraction). */
typedef struct fs_batch fs_batch;
/* Time the FP32 multi-RHS stencil of this system for every z-march depth
 * (1, 2, 4, ... <= nz planes per tile), with one node per thread and with
 * every width of 2, 3 or 4 nodes per lane that divides the rows, and keep
 * the fastest (cached per grid shape and nr).  Run before fs_pc_create: the
 * preconditioner copies the system's 4-wide tiling. */
int fs_sys_autotune(fs_sys* sys, int nr, int reps, int* kz);
/* Many capacitance extractions (each <= 8 terminals) solved together.  The
 * systems are split into min(nstreams, systems) groups of about equal work,
 * each captured as one hipGraph on a stream of its own.  Per system a graph
 * holds `chunk` iterations and then the system's device-side control: when
 * every running right-hand side has ended its FP32 cycle, the FP64 refresh
 * and the next cycle's start run; on other replays those kernels, and every
 * kernel of a finished system, exit at once.  One host thread per group
 * replays its graph, two replays in flight, until a device counter reports
 * the group's systems finished; then the charges are integrated (and any
 * flexible-CG system that failed is redone in FP64). */
int fs_batch_create(fs_ctx* ctx, int nstreams, fs_batch** out);
int fs_batch_destroy(fs_batch* batch);
int fs_batch_add(fs_batch* batch, fs_cap* cap, const fs_op* L);
/* C: the systems' nt_i x nt_i matrices concatenated (Maxwell form, not
 * symmetrised); inst_time[i]: seconds from the start of the iterations until
 * the device found system i finished (device wall clock; a system redone in
 * FP64: until that redo ended, host clock); inst_iters[i]: its Krylov
 * iterations (NULL to skip). */
int fs_batch_solve(fs_batch* batch, const fs_solve_params* params, int chunk, double* C,
                   double* inst_time, int* inst_iters, fs_solve_stats* stats);

/* ---- island coarse space (high-contrast systems) -------------------------
 * No counterpart in the reference, whose Jacobi-CG (fieldspice/solvers/
 * poisson.py:487-589) hits its max(500, 10 sqrt(n)) cap when a conductor
 * floats in an insulator (an EQS step, eqs.py:279-300, on copper in SiO2).
 * Each floating conductor ("island") is a near-null mode; the island space
 * adds one coarse vector per island, Z_I = D^1/2 1_I (equilibrated space), to
 * the preconditioner: M = I + Z E^-1 Z^T with E = Z^T A_s Z = 1_I^T L_ff 1_J.
 * FS_PREC_FP64 runs PCG with M; FS_PREC_MIXED runs its FP32 cycles on the
 * unscaled operator in difference form (y_n = e_n u_n + sum_e w_e (u_n - u_m))
 * with M.  The answer is the same linear system's; only the iteration
 * changes.  Not with FS_PREC_FCG or fs_batch. */
/* Islands of the system's operator: components of its strong edges (w_e >
 * sqrt(w_min w_max) over the positive weights) with >= 2 nodes and no
 * Dirichlet or pinned row.  None when w_max < contrast * w_min.  island: C
 * order, -1 or the island id; at most max_islands (<= 1024) are returned,
 * the largest ones. */
int fs_sys_find_islands(fs_sys* sys, double contrast, int max_islands, int32_t* island,
                        int* n_islands);
/* Attach (n_islands >= 1) or remove (0) the island space; island: C order,
 * -1 or an id in [0, n_islands), every island non-empty and free of
 * Dirichlet/pinned rows.  E must be positive definite (each island coupled
 * to a Dirichlet row through the operator). */
int fs_sys_set_deflation(fs_sys* sys, int n_islands, const int32_t* island);
/* number of islands attached; einv (nullable): E^-1, row major */
int fs_sys_get_deflation(fs_sys* sys, int* n_islands, double* einv);

#ifdef __cplusplus
}
#endif
#endif /* FIELDSPICE_HIP_H */
