"""ctypes view of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, where it is the checker (and the timed CPU port),
never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

TT2006, ORD2011, MS2003, CRN1998 = 0, 1, 2, 3
MODEL_IDS = {"tt2006": TT2006, "ord2011": ORD2011, "ms2003": MS2003, "crn1998": CRN1998}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = C.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True,
                   cwd=os.path.dirname(HERE))


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_model_neq.restype = C.c_int
        L.orc_model_neq.argtypes = [C.c_int]
        L.orc_model_initial_v.restype = C.c_double
        L.orc_model_initial_v.argtypes = [C.c_int]
        L.orc_model_init.argtypes = [C.c_int, _dp, C.c_int64]
        L.orc_model_solve.argtypes = [C.c_int, _dp, C.c_int64, C.c_void_p, C.c_double, C.c_int]
        L.orc_model_solve_adaptive.restype = C.c_int64
        L.orc_model_solve_adaptive.argtypes = [C.c_int, _dp, C.c_int64, C.c_void_p, C.c_double, _dp] + [C.c_double] * 4
        L.orc_model_solve_scaled.argtypes = [C.c_int, _dp, C.c_int64, C.c_void_p, C.c_void_p, C.c_double, C.c_int]
        L.orc_alpha.restype = C.c_double
        L.orc_alpha.argtypes = [C.c_double] * 6
        L.orc_assemble_iso.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [_dp, _u8p, _dp]
        L.orc_assemble_aniso.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [_dp, _u8p, _dp]
        L.orc_fibre_tensor.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.orc_normalise.argtypes = [_dp]
        L.orc_stencil_apply.argtypes = [C.c_int] * 3 + [_dp, _u8p, _dp, _dp]
        L.orc_cg_stencil.restype = C.c_int
        L.orc_cg_stencil.argtypes = [C.c_int] * 3 + [_dp, _u8p, _dp, _dp, C.c_double, C.c_int,
                                                     C.c_int, C.POINTER(C.c_double)]
        L.orc_cg_csr.restype = C.c_int
        L.orc_cg_csr.argtypes = [C.c_int, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS"),
                                 np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"), _dp, _dp, _dp,
                                 C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.orc_tissue_init.argtypes = [C.c_void_p]
        L.orc_tissue_step.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _mask_ptr(active):
    if active is None:
        return None
    a = np.ascontiguousarray(active, dtype=np.uint8)
    return a, a.ctypes.data


def neq(model: int) -> int:
    return lib().orc_model_neq(model)


def initial_v(model: int) -> float:
    return lib().orc_model_initial_v(model)


def model_init(model: int, n: int) -> np.ndarray:
    sv = np.zeros(n * neq(model), dtype=np.float64)
    lib().orc_model_init(model, sv, n)
    return sv.reshape(n, neq(model))


def model_solve(model: int, sv: np.ndarray, stim: np.ndarray | None, dt: float, num_steps: int = 1) -> None:
    """In place on an AoS (n, neq) float64 array."""
    assert sv.dtype == np.float64 and sv.flags.c_contiguous
    n = sv.shape[0]
    st = None if stim is None else np.ascontiguousarray(stim, dtype=np.float64)
    lib().orc_model_solve(model, sv.reshape(-1), n, None if st is None else st.ctypes.data, dt, num_steps)


def model_solve_scaled(model: int, sv: np.ndarray, stim: np.ndarray | None, scale: np.ndarray, dt: float,
                       num_steps: int = 1) -> None:
    """model_solve with per-cell scalings scale (n, 4) float32: TT2006 GKr, GCaL,
    GKs; MS2003 tau_close (no reference counterpart; all 1 is model_solve)."""
    assert sv.dtype == np.float64 and sv.flags.c_contiguous
    n = sv.shape[0]
    sc = np.ascontiguousarray(scale, dtype=np.float32).reshape(n, 4)
    st = None if stim is None else np.ascontiguousarray(stim, dtype=np.float64)
    lib().orc_model_solve_scaled(model, sv.reshape(-1), n, None if st is None else st.ctypes.data, sc.ctypes.data, dt,
                                 num_steps)


def model_solve_adaptive(model: int, sv: np.ndarray, stim, final_time: float, aux: np.ndarray, min_dt: float,
                         max_dt: float, abs_tol: float, rel_tol: float) -> int:
    """Adaptive Euler to final_time, in place on AoS states (n, neq) and the
    per-cell (dt, time_new, previous_dt) rows aux (n, 3)."""
    assert sv.dtype == np.float64 and sv.flags.c_contiguous and aux.dtype == np.float64 and aux.flags.c_contiguous
    n = sv.shape[0]
    st = None if stim is None else np.ascontiguousarray(stim, dtype=np.float64)
    return lib().orc_model_solve_adaptive(model, sv.reshape(-1), n, None if st is None else st.ctypes.data, final_time,
                                          aux.reshape(-1), min_dt, max_dt, abs_tol, rel_tol)


def alpha(beta, cm, dt, dx, dy, dz) -> float:
    return lib().orc_alpha(beta, cm, dt, dx, dy, dz)


def assemble_iso(shape, h, alpha_v, sigma3, active=None) -> np.ndarray:
    nx, ny, nz = shape
    dx, dy, dz = h
    n = nx * ny * nz
    out = np.zeros(n * 27, dtype=np.float64)
    keep = _mask_ptr(active)
    lib().orc_assemble_iso(nx, ny, nz, dx, dy, dz, alpha_v, np.ascontiguousarray(sigma3, np.float64).reshape(-1),
                           None if keep is None else keep[1], out)
    return out.reshape(n, 27)


def assemble_aniso(shape, h, alpha_v, tensor6, active=None) -> np.ndarray:
    nx, ny, nz = shape
    dx, dy, dz = h
    n = nx * ny * nz
    out = np.zeros(n * 27, dtype=np.float64)
    keep = _mask_ptr(active)
    lib().orc_assemble_aniso(nx, ny, nz, dx, dy, dz, alpha_v, np.ascontiguousarray(tensor6, np.float64).reshape(-1),
                             None if keep is None else keep[1], out)
    return out.reshape(n, 27)


def normalise(f) -> np.ndarray:
    v = np.array(f, dtype=np.float64)
    lib().orc_normalise(v)
    return v


def fibre_tensor(f, sigma_l, sigma_t) -> np.ndarray:
    out = np.zeros(6, dtype=np.float64)
    lib().orc_fibre_tensor(np.ascontiguousarray(f, np.float64), sigma_l, sigma_t, out)
    return out


def stencil_apply(shape, coef27, x, active=None) -> np.ndarray:
    nx, ny, nz = shape
    y = np.zeros(nx * ny * nz, dtype=np.float64)
    keep = _mask_ptr(active)
    lib().orc_stencil_apply(nx, ny, nz, np.ascontiguousarray(coef27, np.float64).reshape(-1),
                            None if keep is None else keep[1], np.ascontiguousarray(x, np.float64), y)
    return y


def cg_stencil(shape, coef27, b, x0, tol, max_its=200, jacobi=False, active=None):
    nx, ny, nz = shape
    x = np.array(x0, dtype=np.float64, copy=True)
    err = C.c_double(0)
    keep = _mask_ptr(active)
    its = lib().orc_cg_stencil(nx, ny, nz, np.ascontiguousarray(coef27, np.float64).reshape(-1),
                               None if keep is None else keep[1], np.ascontiguousarray(b, np.float64), x, tol,
                               max_its, int(jacobi), C.byref(err))
    return x, its, err.value


def cg_csr(row_ptr, col, val, b, x0, tol, max_its=200, jacobi=False):
    x = np.array(x0, dtype=np.float64, copy=True)
    err = C.c_double(0)
    its = lib().orc_cg_csr(len(row_ptr) - 1, np.ascontiguousarray(row_ptr, np.int64),
                           np.ascontiguousarray(col, np.int32), np.ascontiguousarray(val, np.float64),
                           np.ascontiguousarray(b, np.float64), x, tol, max_its, int(jacobi), C.byref(err))
    return x, its, err.value


class _Tissue(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("dz", C.c_double), ("model", C.c_int), ("dt", C.c_double), ("beta", C.c_double),
                ("cm", C.c_double), ("cg_tol", C.c_double), ("cg_max_its", C.c_int), ("cg_jacobi", C.c_int),
                ("stim_start", C.c_double), ("stim_duration", C.c_double), ("stim_current", C.c_double),
                ("stim_x_limit", C.c_double), ("stim_period", C.c_double), ("coef27", C.c_void_p),
                ("active", C.c_void_p), ("sv", C.c_void_p), ("v", C.c_void_p), ("b", C.c_void_p),
                ("cur_time", C.c_double), ("count", C.c_int64), ("total_cg_its", C.c_int64),
                ("last_cg_its", C.c_int), ("last_cg_err", C.c_double), ("stim_mask", C.c_void_p),
                ("ionic_scale", C.c_void_p), ("adaptive", C.c_int), ("ode_min_dt", C.c_double),
                ("ode_abs_tol", C.c_double), ("ode_rel_tol", C.c_double), ("aux", C.c_void_p)]


class Tissue:
    """The oracle's operator-split run on a structured box (AoS state)."""

    def __init__(self, shape, h, model, coef27, dt=0.02, beta=0.14, cm=1.0, cg_tol=1e-16, cg_max_its=200,
                 jacobi=False, stim=(0.0, 2.0, -50.0, 500.0, 0.0), active=None, stim_mask=None, ionic_scale=None,
                 adaptive=None):
        """adaptive: None, or (min_dt, abs_tol, rel_tol) for the adaptive Euler
        ODE solver (max dt = dt)."""
        nx, ny, nz = shape
        n = nx * ny * nz
        self.shape, self.n, self.model = shape, n, model
        self.neq = neq(model)
        self.coef27 = np.ascontiguousarray(coef27, np.float64).reshape(-1)
        self.active = None if active is None else np.ascontiguousarray(active, np.uint8)
        self.sv = np.zeros(n * self.neq, np.float64)
        self.v = np.zeros(n, np.float64)
        self.b = np.zeros(n, np.float64)
        t = _Tissue()
        t.nx, t.ny, t.nz = nx, ny, nz
        t.dx, t.dy, t.dz = h
        t.model, t.dt, t.beta, t.cm = model, dt, beta, cm
        t.cg_tol, t.cg_max_its, t.cg_jacobi = cg_tol, cg_max_its, int(jacobi)
        t.stim_start, t.stim_duration, t.stim_current, t.stim_x_limit, t.stim_period = stim
        t.coef27 = self.coef27.ctypes.data
        t.active = None if self.active is None else self.active.ctypes.data
        t.sv, t.v, t.b = self.sv.ctypes.data, self.v.ctypes.data, self.b.ctypes.data
        self.stim_mask = None if stim_mask is None else np.ascontiguousarray(stim_mask, np.uint8)
        self.ionic_scale = None if ionic_scale is None else np.ascontiguousarray(ionic_scale, np.float32).reshape(n, 4)
        t.stim_mask = None if self.stim_mask is None else self.stim_mask.ctypes.data
        t.ionic_scale = None if self.ionic_scale is None else self.ionic_scale.ctypes.data
        self.aux = np.zeros((n, 3), np.float64)
        if adaptive is not None:
            t.adaptive = 1
            t.ode_min_dt, t.ode_abs_tol, t.ode_rel_tol = adaptive
            t.aux = self.aux.ctypes.data
        self._t = t
        lib().orc_tissue_init(C.byref(t))

    def step(self, k: int = 1) -> None:
        for _ in range(k):
            lib().orc_tissue_step(C.byref(self._t))

    def restore(self, ck: dict) -> None:
        """Continue from a checkpoint read by read_reference_checkpoint, as
        restore_simulation_state leaves the solver (restore_state.c:45-256):
        potentials and states of the active cells, current time and iteration;
        the stimulus start stays the configured one."""
        have = ~np.isnan(ck["v"])
        self.v[have] = ck["v"][have]
        sv = self.sv.reshape(self.n, self.neq)
        sv[have] = ck["sv"][have]
        if ck["aux"] is not None and self._t.adaptive:
            self.aux[have] = ck["aux"][have]
        self._t.cur_time = ck["time"]
        self._t.count = ck["iteration"]

    @property
    def time(self) -> float:
        return self._t.cur_time

    @property
    def last_cg(self):
        return self._t.last_cg_its, self._t.last_cg_err

    @property
    def total_cg_its(self) -> int:
        return self._t.total_cg_its


def read_reference_checkpoint(grid: bytes, solver: bytes, ode: bytes, shape, h, neq_: int) -> dict:
    """The three files of save_simulation_state (save_state.c:16-155) on a
    structured box: v (n,) and sv (n, neq) in lexicographic cell order (NaN
    for cells the checkpoint does not hold), aux (n, 3) adaptive rows or None,
    time and iteration.  Fields are written one by one without padding; the
    k-th active grid record owns state column k (sv_position = grid_position,
    monodomain_solver.c:1116); gpu = 1 files are state-major [rows][N], gpu = 0
    files cell-major [N][neq]."""
    import struct
    nx, ny, nz = shape
    n = nx * ny * nz
    # grid: mesh header (6 doubles, 2 uint32), then per cell record
    nc, na = struct.unpack_from("<II", grid, 48)
    off, lex, vals = 56, [], []
    for _ in range(nc):
        cx, cy, cz, v = struct.unpack_from("<4d", grid, off)
        active = grid[off + 8 * 11 + 1]
        (extra,) = struct.unpack_from("<Q", grid, off + 8 * 11 + 2)
        off += 98 + extra
        if active:
            i, j, k = (int(np.floor(c / d)) for c, d in ((cx, h[0]), (cy, h[1]), (cz, h[2])))
            assert (i + 0.5) * h[0] == cx and (j + 0.5) * h[1] == cy and (k + 0.5) * h[2] == cz
            lex.append(i + nx * (j + ny * k))
            vals.append(v)
    assert off == len(grid) and len(lex) == na
    lex = np.array(lex)
    # monodomain solver: struct (88 bytes) + time_info {current_t, final_t, dt, iteration}
    assert len(solver) == 120
    time = struct.unpack_from("<d", solver, 88)[0]
    (iteration,) = struct.unpack_from("<i", solver, 112)
    # ODE solver header
    adaptive = ode[0]
    (ncts,) = struct.unpack_from("<Q", ode, 33)
    flag, gpu = ode[41], ode[42]
    assert flag == 0
    (N,) = struct.unpack_from("<I", ode, 55)
    assert N == na
    rows = neq_ + (3 if adaptive and gpu else 0)
    body = np.frombuffer(ode[59:], np.float64)
    assert body.size == rows * N
    cols = body.reshape(rows, N) if gpu else body.reshape(N, neq_).T
    v = np.full(n, np.nan)
    v[lex] = vals
    sv = np.full((n, neq_), np.nan)
    sv[lex] = cols[:neq_].T
    aux = None
    if rows > neq_:
        aux = np.full((n, 3), np.nan)
        aux[lex] = cols[neq_:].T
    return {"v": v, "sv": sv, "aux": aux, "time": time, "iteration": iteration, "adaptive": bool(adaptive), "gpu": bool(gpu)}


class Observer:
    """Activation times and APDs as save_with_activation_times keeps them
    (save_mesh.c:754-913), vectorised over cells: Vm, counters and times in
    float32 as in the reference's hash values and arrays."""

    def __init__(self, n, act_thr=-30.0, apd_thr=-83.0, time_thr=10.0):
        self.act_thr, self.apd_thr, self.time_thr = np.float32(act_thr), np.float32(apd_thr), np.float32(time_thr)
        self.last_v = np.full(n, -100.0, np.float32)   # hmdefault(last_time_v, -100.0)
        self.was_active = np.zeros(n, np.float32)       # cell_was_active
        self.acts = [[] for _ in range(n)]              # activation_times (float arrays)
        self.apds = [[] for _ in range(n)]

    def observe(self, v, t, active=None):
        """One save call at time t (double) on Vm v (double)."""
        v = np.asarray(v, np.float64)
        cells = np.arange(len(v)) if active is None else np.nonzero(active)[0]
        if t == 0.0:                                     # :836-837
            self.last_v[cells] = v[cells].astype(np.float32)
            return
        thr = np.float64(self.act_thr)
        cross = (self.last_v[cells] < self.act_thr) & (v[cells] >= thr)   # :839
        for c in cells[cross]:
            a = self.acts[c]
            if not a or t - np.float64(a[-1]) > np.float64(self.time_thr):  # :840-857
                a.append(np.float32(t))
                self.was_active[c] += np.float32(1.0)
        for c in cells[self.was_active[cells] != 0.0]:   # :861-874
            if v[c] <= np.float64(self.apd_thr) or self.was_active[c] == 2.0:
                tmp = int(self.was_active[c])
                last = np.float64(self.acts[c][len(self.acts[c]) - tmp])
                self.apds[c].append(np.float32(t - last))
                self.was_active[c] = np.float32(tmp - 1)
        self.last_v[cells] = v[cells].astype(np.float32)   # :875
