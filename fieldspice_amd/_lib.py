"""ctypes binding of ``libfieldspice_hip.so`` (C ABI in include/fieldspice_hip.h).

The library is built in-tree by ``fieldspice_amd.build``.  There is no CPU
fallback: if the library or a GPU is missing, every solve raises.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfieldspice_hip.so")

FS_OK = 0
FS_NOT_CONVERGED = 1
FS_EINVAL = -1
FS_EHIP = -2
FS_ENOMEM = -3

FS_AVG = {"parallel": 0, "arithmetic": 0, "harmonic": 1, "series": 1, "min": 2, "max": 3}
FS_PREC_FP64 = 0
FS_PREC_MIXED = 1
FS_PREC_FCG = 2
FS_LOSS_CYCLE = 0
FS_LOSS_BOTTOM = 1

EXPORTS = (
    "fs_ctx_create", "fs_ctx_destroy", "fs_last_error", "fs_ctx_device_info", "fs_stream_probe",
    "fs_op_create_edges", "fs_op_create_cells", "fs_op_combine", "fs_op_destroy",
    "fs_op_apply", "fs_op_get_edges", "fs_sys_create", "fs_sys_destroy",
    "fs_sys_solve", "fs_capacitance", "fs_cap_prepare", "fs_cap_solve", "fs_cap_destroy",
    "fs_sys_set_profiling", "fs_transient_run", "fs_op_edge_power",
    "fs_net_create", "fs_net_destroy", "fs_net_num_params", "fs_net_get_params",
    "fs_net_set_params", "fs_net_eval", "fs_pc_create", "fs_pc_destroy", "fs_pc_level_dims",
    "fs_sys_set_precond", "fs_pc_apply", "fs_pc_add_pairs", "fs_pc_num_pairs",
    "fs_pc_clear_pairs", "fs_pc_loss_grad", "fs_pc_train", "fs_sys_autotune",
    "fs_batch_create", "fs_batch_destroy", "fs_batch_add", "fs_batch_solve",
    "fs_sys_find_islands", "fs_sys_set_deflation", "fs_sys_get_deflation",
    "fs_sys_solve_noalloc", "fs_op_create_periodic",
)


# fs_solve_stats.kernel_* (FS_PROF_* in include/fieldspice_hip.h)
PROF_KINDS = 5
PROF_NAMES = ("stencil", "direction", "residual", "restrict", "prolong")


class SolveParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxiter", C.c_int), ("precision", C.c_int),
                ("inner_tol", C.c_double), ("check_every", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("refreshes", C.c_int), ("converged", C.c_int),
                ("residual", C.c_double), ("residual_eq", C.c_double),
                ("solve_time", C.c_double), ("stencil_time", C.c_double),
                ("stencil_launches", C.c_int64), ("stencil_bytes", C.c_double),
                ("fallbacks", C.c_int), ("pc_applies", C.c_int), ("stop_reason", C.c_int),
                ("maxiter", C.c_int), ("replacements", C.c_int), ("stencil_width", C.c_int),
                ("cycle_width", C.c_int), ("kernel_time", C.c_double * PROF_KINDS),
                ("kernel_launches", C.c_int64 * PROF_KINDS),
                ("kernel_bytes", C.c_double * PROF_KINDS)]

    def as_dict(self) -> dict:
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v) if isinstance(v, C.Array) else v
        return out


class SysInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_free", C.c_int64), ("n_fixed", C.c_int64),
                ("n_empty", C.c_int64), ("gauge", C.c_int64), ("singular", C.c_int)]


# fs_transient_desc.warm_start: the initial guess of each time step's solve
GUESSES = {"zero": 0, "previous": 1, "extrapolate": 2, "project": 3}


class TransientDesc(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("ngroups", C.c_int), ("group", C.c_void_p),
                ("values", C.c_void_p), ("nsrc", C.c_int), ("src", C.c_void_p),
                ("src_scale", C.c_void_p), ("warm_start", C.c_int), ("nprobe", C.c_int),
                ("probe", C.c_void_p), ("probe_w", C.c_void_p), ("cur_a", C.c_void_p),
                ("cur_b", C.c_void_p), ("quad_a", C.c_void_p), ("quad_b", C.c_void_p),
                ("store_every", C.c_int), ("x_before", C.c_void_p),
                ("x_before_out", C.c_void_p), ("proj_dim", C.c_int)]


class LibraryError(RuntimeError):
    """A failure reported by the HIP library (or the library is unavailable)."""


_lock = threading.Lock()
_lib = None
_ctx = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def _declare(lib):
    lib.fs_last_error.restype = C.c_char_p
    lib.fs_last_error.argtypes = []
    lib.fs_ctx_create.argtypes = [C.c_int, C.POINTER(_vp)]
    lib.fs_ctx_destroy.argtypes = [_vp]
    lib.fs_ctx_device_info.argtypes = [_vp, C.c_char_p, C.c_int, C.POINTER(C.c_int)]
    lib.fs_stream_probe.argtypes = [_vp, C.c_int64, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_double)]
    lib.fs_op_create_edges.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                       _vp, C.POINTER(_vp)]
    lib.fs_op_create_cells.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                       C.c_int, C.POINTER(_vp)]
    lib.fs_op_create_periodic.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int,
                                          C.POINTER(_vp)]
    lib.fs_op_combine.argtypes = [C.c_double, _vp, C.c_double, _vp, C.c_double, _vp,
                                  C.POINTER(_vp)]
    lib.fs_op_destroy.argtypes = [_vp]
    lib.fs_op_apply.argtypes = [_vp, _dp, _dp]
    lib.fs_op_get_edges.argtypes = [_vp, _dp, _dp, _dp]
    lib.fs_sys_create.argtypes = [_vp, _vp, C.c_int, C.POINTER(_vp), C.POINTER(SysInfo)]
    lib.fs_sys_destroy.argtypes = [_vp]
    lib.fs_sys_solve.argtypes = [_vp, C.c_int, _dp, _vp, _dp, C.POINTER(SolveParams),
                                 C.POINTER(SolveStats)]
    lib.fs_sys_solve_noalloc.argtypes = [_vp, C.c_int, _dp, _vp, _dp, C.POINTER(SolveParams),
                                         C.POINTER(SolveStats)]
    lib.fs_capacitance.argtypes = [_vp, _vp, C.c_int, _i32p, _dp, _vp, _vp,
                                   C.POINTER(SolveParams), C.POINTER(SolveStats)]
    lib.fs_cap_prepare.argtypes = [_vp, C.c_int, _i32p, C.POINTER(_vp)]
    lib.fs_cap_solve.argtypes = [_vp, _vp, _dp, _vp, _vp, C.POINTER(SolveParams),
                                 C.POINTER(SolveStats)]
    lib.fs_cap_destroy.argtypes = [_vp]
    lib.fs_sys_set_profiling.argtypes = [_vp, C.c_int]
    lib.fs_transient_run.argtypes = [_vp, _vp, C.POINTER(TransientDesc), _dp, _vp, _vp, _vp,
                                     _vp, C.POINTER(SolveParams), C.POINTER(SolveStats)]
    lib.fs_op_edge_power.argtypes = [_vp, _dp, _dp]
    lib.fs_net_create.argtypes = [_vp, C.c_uint64, C.POINTER(_vp)]
    lib.fs_net_destroy.argtypes = [_vp]
    lib.fs_net_num_params.argtypes = [_vp, C.POINTER(C.c_int64)]
    lib.fs_net_get_params.argtypes = [_vp, _fp]
    lib.fs_net_set_params.argtypes = [_vp, _fp]
    lib.fs_net_eval.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _fp, _fp, _vp, _vp]
    lib.fs_pc_create.argtypes = [_vp, _vp, C.c_int, _i32p, C.c_double, C.POINTER(_vp)]
    lib.fs_pc_destroy.argtypes = [_vp]
    lib.fs_pc_level_dims.argtypes = [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
    lib.fs_sys_set_precond.argtypes = [_vp, _vp]
    lib.fs_pc_apply.argtypes = [_vp, C.c_int, _dp, _dp]
    lib.fs_pc_add_pairs.argtypes = [_vp, C.c_int, _dp]
    lib.fs_pc_num_pairs.argtypes = [_vp, C.POINTER(C.c_int)]
    lib.fs_pc_clear_pairs.argtypes = [_vp]
    lib.fs_pc_loss_grad.argtypes = [_vp, C.c_int, C.c_int, _dp, _dp, _vp]
    lib.fs_pc_train.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, _dp]
    lib.fs_sys_autotune.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_int)]
    lib.fs_batch_create.argtypes = [_vp, C.c_int, C.POINTER(_vp)]
    lib.fs_batch_destroy.argtypes = [_vp]
    lib.fs_batch_add.argtypes = [_vp, _vp, _vp]
    lib.fs_batch_solve.argtypes = [_vp, C.POINTER(SolveParams), C.c_int, _dp, _vp, _vp,
                                   C.POINTER(SolveStats)]
    lib.fs_sys_find_islands.argtypes = [_vp, C.c_double, C.c_int, _i32p, C.POINTER(C.c_int)]
    lib.fs_sys_set_deflation.argtypes = [_vp, C.c_int, _vp]
    lib.fs_sys_get_deflation.argtypes = [_vp, C.POINTER(C.c_int), _vp]
    for name in EXPORTS:
        if name != "fs_last_error":
            getattr(lib, name).restype = C.c_int


def load():
    """Load the library (no GPU needed to load it)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise LibraryError(
                    f"{LIB_PATH} is missing: build it with `python -m fieldspice_amd.build` "
                    "(there is no CPU fallback)")
            lib = C.CDLL(LIB_PATH)
            _declare(lib)
            _lib = lib
        return _lib


def last_error() -> str:
    msg = load().fs_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise LibraryError(f"{what}: {last_error()} (status {rc})")
    return rc


def context():
    """Process-wide context on device 0 (created on first use)."""
    global _ctx
    lib = load()
    with _lock:
        if _ctx is None:
            h = _vp()
            rc = lib.fs_ctx_create(int(os.environ.get("FIELDSPICE_DEVICE", "0")), C.byref(h))
            if rc != FS_OK:
                raise LibraryError(f"fs_ctx_create: {last_error()} (status {rc})")
            _ctx = h
        return _ctx


def device_info() -> tuple[str, int]:
    lib = load()
    buf = C.create_string_buffer(256)
    ncu = C.c_int(0)
    check(lib.fs_ctx_device_info(context(), buf, 256, C.byref(ncu)), "fs_ctx_device_info")
    return buf.value.decode(), ncu.value


def stream_probe(n: int, nread: int, nwrite: int, reps: int = 20) -> float:
    """Seconds per launch of a kernel streaming nread + nwrite FP32 arrays of
    n elements once (fs_stream_probe)."""
    sec = C.c_double(0.0)
    check(load().fs_stream_probe(context(), int(n), int(nread), int(nwrite), int(reps),
                                 C.byref(sec)), "fs_stream_probe")
    return sec.value


def as_f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class DeviceOperator:
    """``A = G^T diag(w) G + diag(extra)`` resident on the GPU (an ``fs_op``)."""

    def __init__(self, handle, shape_cells, periodic: int = 0):
        self._h = handle
        self.shape_cells = tuple(shape_cells)
        self.periodic = int(periodic)

    @classmethod
    def from_periodic_edges(cls, shape_nodes, wx, wy, wz, periodic: int) -> "DeviceOperator":
        """Operator of a grid whose walls are identified along the axes in
        ``periodic`` (bit a for axis a), given on the reduced grid: along a
        periodic axis as many edges as nodes, the last one wrapping around."""
        nx, ny, nz = (int(v) for v in shape_nodes)
        h = _vp()
        check(load().fs_op_create_periodic(context(), nx, ny, nz, as_f64(wx).ravel(),
                                           as_f64(wy).ravel(), as_f64(wz).ravel(),
                                           int(periodic), C.byref(h)), "fs_op_create_periodic")
        return cls(h, (nx - 1, ny - 1, nz - 1), periodic)

    @classmethod
    def from_cells(cls, grid, prop_cell, mode: str = "parallel") -> "DeviceOperator":
        lib = load()
        if mode not in FS_AVG:
            raise ValueError(f"unknown averaging mode {mode!r}")
        Nx, Ny, Nz = grid.shape_cells
        h = _vp()
        check(lib.fs_op_create_cells(context(), Nx, Ny, Nz, as_f64(grid.hx), as_f64(grid.hy),
                                     as_f64(grid.hz), as_f64(prop_cell), FS_AVG[mode],
                                     C.byref(h)), "fs_op_create_cells")
        return cls(h, grid.shape_cells)

    @classmethod
    def from_edges(cls, grid, w_edges, extra=None) -> "DeviceOperator":
        """``w_edges``: concatenated [x, y, z] edge weights (fieldspice order)."""
        lib = load()
        w = as_f64(w_edges).ravel()
        nex, ney, nez = grid.n_edges_each
        if w.size != nex + ney + nez:
            raise ValueError(f"expected {nex + ney + nez} edge weights, got {w.size}")
        wx = np.ascontiguousarray(w[:nex])
        wy = np.ascontiguousarray(w[nex:nex + ney])
        wz = np.ascontiguousarray(w[nex + ney:])
        ex = None if extra is None else as_f64(extra).ravel()
        if ex is not None and ex.size != grid.n_nodes:
            raise ValueError("extra must have one value per node")
        Nx, Ny, Nz = grid.shape_cells
        h = _vp()
        check(lib.fs_op_create_edges(context(), Nx, Ny, Nz, wx, wy, wz,
                                     None if ex is None else ex.ctypes.data_as(_vp),
                                     C.byref(h)), "fs_op_create_edges")
        return cls(h, grid.shape_cells)

    def combine(self, a: float, b: float = 0.0, other: "DeviceOperator | None" = None,
                c: float = 0.0, extra=None) -> "DeviceOperator":
        """``a*self + b*other + diag(c*extra)``."""
        lib = load()
        ex = None if extra is None else as_f64(extra).ravel()
        h = _vp()
        check(lib.fs_op_combine(float(a), self._h, float(b), None if other is None else other._h,
                                float(c), None if ex is None else ex.ctypes.data_as(_vp),
                                C.byref(h)), "fs_op_combine")
        return DeviceOperator(h, self.shape_cells, self.periodic)

    @property
    def n_nodes(self) -> int:
        Nx, Ny, Nz = self.shape_cells
        return (Nx + 1) * (Ny + 1) * (Nz + 1)

    def apply(self, x) -> np.ndarray:
        x = as_f64(x).ravel()
        if x.size != self.n_nodes:
            raise ValueError(f"vector must have {self.n_nodes} entries")
        y = np.empty_like(x)
        check(load().fs_op_apply(self._h, x, y), "fs_op_apply")
        return y

    def edge_power(self, x) -> np.ndarray:
        """``q_n = 1/2 sum_e w_e (x_n - x_m)^2`` (Joule heat per node)."""
        x = as_f64(x).ravel()
        q = np.empty_like(x)
        check(load().fs_op_edge_power(self._h, x, q), "fs_op_edge_power")
        return q

    def edges(self) -> np.ndarray:
        return np.concatenate([w.ravel() for w in self.edges_split()])

    def edges_split(self):
        """Edge weights per family, shaped (Nx, Ny+1, Nz+1), (Nx+1, Ny, Nz+1),
        (Nx+1, Ny+1, Nz)."""
        nodes = [c + 1 for c in self.shape_cells]
        shapes = []
        for a in range(3):
            shp = list(nodes)
            shp[a] = nodes[a] if (self.periodic >> a) & 1 else nodes[a] - 1
            shapes.append(tuple(shp))
        w = [np.empty(int(np.prod(shp))) for shp in shapes]
        check(load().fs_op_get_edges(self._h, w[0], w[1], w[2]), "fs_op_get_edges")
        return tuple(a.reshape(shp) for a, shp in zip(w, shapes))

    def close(self):
        if getattr(self, "_h", None):
            load().fs_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceSystem:
    """A prepared solve of ``op`` with a Dirichlet mask (an ``fs_sys``)."""

    def __init__(self, op: DeviceOperator, mask=None, pin: bool = True):
        lib = load()
        self.op = op  # keeps the operator alive
        m = None
        if mask is not None:
            m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
            if m.size != op.n_nodes:
                raise ValueError("mask must have one entry per node")
        h = _vp()
        info = SysInfo()
        rc = lib.fs_sys_create(op._h, None if m is None else m.ctypes.data_as(_vp),
                               1 if pin else 0, C.byref(h), C.byref(info))
        if rc == FS_EINVAL and "singular" in last_error():
            raise ValueError(last_error())
        check(rc, "fs_sys_create")
        self._h = h
        self.info = {k: getattr(info, k) for k, _ in info._fields_}

    @staticmethod
    def params(tol, maxiter=None, precision=FS_PREC_FP64, inner_tol=0.0,
               check_every=0) -> SolveParams:
        """``inner_tol`` 0: the library's default per method (1e-4 mixed CG,
        1e-3 flexible CG)."""
        return SolveParams(float(tol), int(maxiter or 0), int(precision), float(inner_tol),
                           int(check_every))

    def solve(self, b, params: SolveParams, x0=None, noalloc: bool = False):
        """Solve for one (n,) or several (k, n) right-hand sides; ``x0`` (same
        shape as ``b``) is an initial guess.  ``noalloc``: through
        ``fs_sys_solve_noalloc`` (the workspace in the library's static
        scratch, no device allocation)."""
        b = as_f64(b)
        single = b.ndim == 1
        b2 = b.reshape(1, -1) if single else b
        n = self.op.n_nodes
        if b2.shape[1] != n:
            raise ValueError(f"right-hand side must have shape ({n},), got {b.shape}")
        g = None
        if x0 is not None:
            g = as_f64(x0)
            if g.shape != b.shape:
                raise ValueError(f"x0 must have the shape of b {b.shape}, got {g.shape}")
            g = np.ascontiguousarray(g.reshape(b2.shape))
        x = np.empty_like(b2)
        st = SolveStats()
        name = "fs_sys_solve_noalloc" if noalloc else "fs_sys_solve"
        rc = getattr(load(), name)(self._h, b2.shape[0], np.ascontiguousarray(b2),
                                   None if g is None else g.ctypes.data_as(_vp), x,
                                   C.byref(params), C.byref(st))
        check(rc, name)
        return (x[0] if single else x), st.as_dict(), rc

    def capacitance(self, L: DeviceOperator, term_id, nt: int, params: SolveParams,
                    want_phi: bool = False):
        tid = np.ascontiguousarray(term_id, dtype=np.int32).ravel()
        Cm = np.zeros((nt, nt))
        qsh = np.zeros(nt)
        phi = np.empty((nt, self.op.n_nodes)) if want_phi else None
        st = SolveStats()
        rc = load().fs_capacitance(self._h, L._h, nt, tid, Cm, qsh.ctypes.data_as(_vp),
                                   None if phi is None else phi.ctypes.data_as(_vp),
                                   C.byref(params), C.byref(st))
        check(rc, "fs_capacitance")
        return Cm, qsh, phi, st.as_dict(), rc

    def transient(self, Aexp, x0, nsteps, group=None, values=None, src=None, src_scale=None,
                  probe=None, probe_w=None, nprobe=0, cur_a=None, cur_b=None, quad_a=None,
                  quad_b=None, store_every=0, initial_guess="project", x_before=None,
                  proj_dim=0, params=None):
        """Device time loop (``fs_transient_run``); returns a dict of outputs.

        ``initial_guess`` of every step's solve: "zero" (as the reference's
        solves start), "previous" (the last field) or "extrapolate" (from the
        second step on, 2 x_{k-1} - x_{k-2}; from the first when ``x_before``, the
        field one step before ``x0``, is given) or "project" (the A-norm
        projection of each step's right-hand side onto a basis of ``proj_dim``
        earlier solutions, 0 for 8)."""
        if initial_guess not in GUESSES:
            raise ValueError(f"initial_guess must be one of {sorted(GUESSES)}, "
                             f"got {initial_guess!r}")
        n = self.op.n_nodes
        keep = []

        def ptr(a, dtype):
            if a is None:
                return None
            arr = np.ascontiguousarray(a, dtype=dtype)
            keep.append(arr)
            return arr.ctypes.data_as(_vp)
        ng = 0 if values is None else int(np.asarray(values).reshape(nsteps + 1, -1).shape[1])
        nsrc = 0 if src is None else int(np.asarray(src).reshape(-1, n).shape[0])
        d = TransientDesc(int(nsteps), ng, ptr(group, np.int32), ptr(values, np.float64), nsrc,
                          ptr(src, np.float64), ptr(src_scale, np.float64),
                          GUESSES[initial_guess], int(nprobe), ptr(probe, np.int32),
                          ptr(probe_w, np.float64),
                          None if cur_a is None else cur_a._h, None if cur_b is None else cur_b._h,
                          None if quad_a is None else quad_a._h,
                          None if quad_b is None else quad_b._h, int(store_every),
                          ptr(None if x_before is None else as_f64(x_before).ravel(), np.float64),
                          None, int(proj_dim))
        x_last_before = np.empty(n)
        d.x_before_out = x_last_before.ctypes.data_as(_vp)
        x = as_f64(x0).ravel().copy()
        nstore = (nsteps // store_every + 1) if store_every > 0 else 0
        stored = np.empty((nstore, n)) if nstore else None
        pi = np.empty((nsteps + 1, nprobe)) if nprobe else None
        pv = np.empty((nsteps + 1, nprobe)) if nprobe else None
        ss = np.empty((nsteps + 1, 4))
        st = SolveStats()
        rc = load().fs_transient_run(self._h, None if Aexp is None else Aexp._h, C.byref(d), x,
                                     None if stored is None else stored.ctypes.data_as(_vp),
                                     None if pi is None else pi.ctypes.data_as(_vp),
                                     None if pv is None else pv.ctypes.data_as(_vp),
                                     ss.ctypes.data_as(_vp), C.byref(params), C.byref(st))
        check(rc, "fs_transient_run")
        return {"x": x, "x_before": x_last_before, "stored": stored, "probe_i": pi,
                "probe_v": pv, "step_stats": ss, "stats": st.as_dict(), "rc": rc}

    def prepare_capacitance(self, term_id, nt: int) -> "DeviceCapacitance":
        return DeviceCapacitance(self, term_id, nt)

    def set_profiling(self, on: bool) -> None:
        check(load().fs_sys_set_profiling(self._h, 1 if on else 0), "fs_sys_set_profiling")

    def autotune(self, nr: int, reps: int = 5) -> int:
        """Pick the stencil's z-march depth for ``nr`` right-hand sides by
        timing; returns the chosen planes per tile."""
        kz = C.c_int(0)
        check(load().fs_sys_autotune(self._h, int(nr), int(reps), C.byref(kz)), "fs_sys_autotune")
        return kz.value

    def set_precond(self, pc: "DevicePreconditioner | None") -> None:
        """Preconditioner of FS_PREC_FCG solves (the preconditioner keeps this
        system alive; the system keeps only the handle)."""
        if pc is not None and pc.system is not self:
            raise ValueError("the preconditioner was built for another system")
        check(load().fs_sys_set_precond(self._h, None if pc is None else pc._h),
              "fs_sys_set_precond")

    def find_islands(self, contrast: float = 1e3, max_islands: int = 1024):
        """Floating high-conductance components of the operator (node ids in
        C order, -1 outside); returns (ids, count)."""
        ids = np.empty(self.op.n_nodes, np.int32)
        n = C.c_int(0)
        check(load().fs_sys_find_islands(self._h, float(contrast), int(max_islands), ids,
                                         C.byref(n)), "fs_sys_find_islands")
        return ids, n.value

    def set_deflation(self, ids=None, n_islands: int = 0) -> None:
        """Attach the island coarse space (ids from find_islands), or remove it
        (no ids)."""
        if ids is None or n_islands == 0:
            check(load().fs_sys_set_deflation(self._h, 0, None), "fs_sys_set_deflation")
            return
        a = np.ascontiguousarray(ids, dtype=np.int32).ravel()
        if a.size != self.op.n_nodes:
            raise ValueError("island ids must have one entry per node")
        check(load().fs_sys_set_deflation(self._h, int(n_islands), a.ctypes.data_as(_vp)),
              "fs_sys_set_deflation")

    def deflation(self):
        """(number of islands, E^-1 or None)."""
        n = C.c_int(0)
        check(load().fs_sys_get_deflation(self._h, C.byref(n), None), "fs_sys_get_deflation")
        if n.value == 0:
            return 0, None
        einv = np.empty((n.value, n.value))
        check(load().fs_sys_get_deflation(self._h, C.byref(n), einv.ctypes.data_as(_vp)),
              "fs_sys_get_deflation")
        return n.value, einv

    def close(self):
        if getattr(self, "_h", None):
            load().fs_sys_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceCapacitance:
    """Terminal map and node lists of one extraction resident on the GPU
    (an ``fs_cap``), solved repeatedly with ``solve``."""

    def __init__(self, system: DeviceSystem, term_id, nt: int):
        self.system = system
        self.nt = int(nt)
        tid = np.ascontiguousarray(term_id, dtype=np.int32).ravel()
        if tid.size != system.op.n_nodes:
            raise ValueError("term_id must have one entry per node")
        h = _vp()
        check(load().fs_cap_prepare(system._h, self.nt, tid, C.byref(h)), "fs_cap_prepare")
        self._h = h

    def solve(self, L: DeviceOperator, params: SolveParams, want_phi: bool = False):
        nt = self.nt
        Cm = np.zeros((nt, nt))
        qsh = np.zeros(nt)
        phi = np.empty((nt, self.system.op.n_nodes)) if want_phi else None
        st = SolveStats()
        rc = load().fs_cap_solve(self._h, L._h, Cm, qsh.ctypes.data_as(_vp),
                                 None if phi is None else phi.ctypes.data_as(_vp),
                                 C.byref(params), C.byref(st))
        check(rc, "fs_cap_solve")
        if want_phi:
            return Cm, qsh, phi, st.as_dict(), rc
        return Cm, qsh, st.as_dict(), rc

    def close(self):
        if getattr(self, "_h", None):
            load().fs_cap_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceNet:
    """The convolutional network of the learned preconditioner (an ``fs_net``):
    11 3x3x3 convolutions on three levels (16/32/64 channels), BF16 weights
    and activations with FP32 accumulation, trained with Adam on the GPU."""

    def __init__(self, seed: int = 0):
        lib = load()
        h = _vp()
        check(lib.fs_net_create(context(), int(seed) & (2 ** 64 - 1), C.byref(h)),
              "fs_net_create")
        self._h = h
        cnt = C.c_int64(0)
        check(lib.fs_net_num_params(h, C.byref(cnt)), "fs_net_num_params")
        self.num_params = int(cnt.value)

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        check(load().fs_net_get_params(self._h, out), "fs_net_get_params")
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32).ravel()
        if p.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {p.size}")
        check(load().fs_net_set_params(self._h, p), "fs_net_set_params")

    def eval(self, X, dY=None):
        """``X``: (nb, nz, ny, nx, 16) channels-last.  Returns Y (same shape)
        and, with ``dY``, the parameter gradient of ``sum(dY * Y)``."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 5 or X.shape[-1] != 16:
            raise ValueError("X must have shape (nb, nz, ny, nx, 16)")
        nb, nz, ny, nx, _ = X.shape
        Y = np.empty_like(X)
        grad = None
        dYa = None
        if dY is not None:
            dYa = np.ascontiguousarray(dY, dtype=np.float32)
            if dYa.shape != X.shape:
                raise ValueError("dY must have the shape of X")
            grad = np.empty(self.num_params, np.float32)
        check(load().fs_net_eval(self._h, nx, ny, nz, nb, X, Y,
                                 None if dYa is None else dYa.ctypes.data_as(_vp),
                                 None if grad is None else grad.ctypes.data_as(_vp)),
              "fs_net_eval")
        return (Y, grad) if dY is not None else Y

    def close(self):
        if getattr(self, "_h", None):
            load().fs_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DevicePreconditioner:
    """Aggregation hierarchy of one prepared system with the network (or one
    Jacobi step) at its bottom (an ``fs_pc``)."""

    def __init__(self, system: DeviceSystem, net: DeviceNet | None, groups, omega: float = 0.95):
        self.system = system
        self.net = net
        flat = np.ascontiguousarray(np.concatenate([np.asarray(g, np.int32).ravel()
                                                    for lv in groups for g in lv]),
                                    dtype=np.int32)
        self.nlevels = len(groups)
        h = _vp()
        check(load().fs_pc_create(system._h, None if net is None else net._h, self.nlevels,
                                  flat, float(omega), C.byref(h)), "fs_pc_create")
        self._h = h

    def level_dims(self, level: int) -> tuple[int, int, int]:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(load().fs_pc_level_dims(self._h, int(level), C.byref(a), C.byref(b), C.byref(c)),
              "fs_pc_level_dims")
        return a.value, b.value, c.value

    def apply(self, r) -> np.ndarray:
        """z = M(r) for equilibrated vectors, (n,) or (k, n) in C order."""
        r = as_f64(r)
        single = r.ndim == 1
        r2 = np.ascontiguousarray(r.reshape(1, -1) if single else r)
        z = np.empty_like(r2)
        check(load().fs_pc_apply(self._h, r2.shape[0], r2, z), "fs_pc_apply")
        return z[0] if single else z

    def add_pairs(self, e) -> None:
        """Exact errors (node space, (k, n)) as training pairs."""
        e2 = np.ascontiguousarray(as_f64(e).reshape(-1, self.system.op.n_nodes))
        check(load().fs_pc_add_pairs(self._h, e2.shape[0], e2), "fs_pc_add_pairs")

    @property
    def num_pairs(self) -> int:
        k = C.c_int()
        check(load().fs_pc_num_pairs(self._h, C.byref(k)), "fs_pc_num_pairs")
        return k.value

    def clear_pairs(self) -> None:
        check(load().fs_pc_clear_pairs(self._h), "fs_pc_clear_pairs")

    def loss_grad(self, e, kind: int = FS_LOSS_BOTTOM):
        e2 = np.ascontiguousarray(as_f64(e).reshape(-1, self.system.op.n_nodes))
        loss = np.empty(e2.shape[0])
        grad = np.empty(self.net.num_params, np.float32)
        check(load().fs_pc_loss_grad(self._h, int(kind), e2.shape[0], e2, loss,
                                     grad.ctypes.data_as(_vp)), "fs_pc_loss_grad")
        return loss, grad

    def train(self, steps: int, batch: int = 4, lr: float = 1e-3, seed: int = 0,
              kind: int = FS_LOSS_BOTTOM) -> np.ndarray:
        losses = np.empty(int(steps))
        check(load().fs_pc_train(self._h, int(kind), int(steps), int(batch), float(lr),
                                 int(seed) & (2 ** 64 - 1), losses), "fs_pc_train")
        return losses

    def close(self):
        if getattr(self, "_h", None):
            load().fs_pc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBatch:
    """Many capacitance extractions solved together (an ``fs_batch``): the
    systems in ``nstreams`` groups, each group's iterations and device-side
    control one captured hipGraph on its own stream, replayed by its own host
    thread until the device counts the group finished."""

    def __init__(self, nstreams: int = 32):
        h = _vp()
        check(load().fs_batch_create(context(), int(nstreams), C.byref(h)), "fs_batch_create")
        self._h = h
        self.items = []   # keeps the extractions and operators alive

    def add(self, cap: DeviceCapacitance, L: DeviceOperator) -> int:
        check(load().fs_batch_add(self._h, cap._h, L._h), "fs_batch_add")
        self.items.append((cap, L))
        return len(self.items) - 1

    def solve(self, params: SolveParams, chunk: int = 4):
        """Returns (list of C matrices, per-system finish times [s],
        per-system iterations, stats dict, status)."""
        sizes = [cap.nt for cap, _ in self.items]
        Cflat = np.zeros(sum(n * n for n in sizes))
        t = np.zeros(len(sizes))
        it = np.zeros(len(sizes), dtype=np.int32)
        st = SolveStats()
        rc = load().fs_batch_solve(self._h, C.byref(params), int(chunk), Cflat,
                                   t.ctypes.data_as(_vp), it.ctypes.data_as(_vp), C.byref(st))
        check(rc, "fs_batch_solve")
        out, off = [], 0
        for n in sizes:
            out.append(Cflat[off:off + n * n].reshape(n, n).copy())
            off += n * n
        return out, t, it, st.as_dict(), rc

    def close(self):
        if getattr(self, "_h", None):
            load().fs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
