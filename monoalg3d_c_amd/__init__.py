"""MI355X-native operator-split monodomain timestep (MonoAlg3D drop-in).

Host-side mirror of the reference's plugin interface for this path:
  - Tissue: the solve_monodomain loop body (ionic step + RHS update + diffusion
    solve) on one structured box, device-resident;
  - ionic_run / assemble / cg_stencil / cg_csr: the SOLVE_MODEL_ODES,
    ASSEMBLY_MATRIX and SOLVE_LINEAR_SYSTEM plugins on their own;
  - workloads: the cuboid, layered-fibre and stimulus set-ups of the benchmark
    configurations.
All compute runs in libmonoalg_hip.so (HIP, gfx950).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import (FP32_REFINE, FP64, MODEL_CRN1998, MODEL_MS2003, MODEL_ORD2011, MODEL_TT2006,  # noqa: F401
                   TENSOR_PER_CELL,
                   TENSOR_PER_LAYER, TENSOR_UNIFORM, MonoError, Stats, TissueDesc, check, lib, ptr)

MODELS = {"tt2006": MODEL_TT2006, "ord2011": MODEL_ORD2011, "ms2003": MODEL_MS2003, "crn1998": MODEL_CRN1998}


def model_neq(model: int) -> int:
    return lib().mono_model_neq(model)


def model_initial_state(model: int) -> np.ndarray:
    out = np.zeros(model_neq(model), np.float64)
    check(lib().mono_model_initial_state(model, ptr(out)), "mono_model_initial_state")
    return out


def model_initial_v(model: int) -> float:
    return lib().mono_model_initial_v(model)


def alpha(beta=0.14, cm=1.0, dt=0.02, dx=100.0, dy=100.0, dz=100.0) -> float:
    """ALPHA of monodomain/constants.h:10-11 (float-rounded 1e-8 factor)."""
    return float(((beta * cm) / dt) * float(np.float32(1e-8)) * ((dx * dy) * dz))


def iso_tensor(sx: float, sy: float | None = None, sz: float | None = None) -> np.ndarray:
    sy = sx if sy is None else sy
    sz = sx if sz is None else sz
    return np.array([sx, sy, sz, 0.0, 0.0, 0.0], np.float64)


def normalise(f) -> np.ndarray:
    """Unit fibre as the reference's fibre reader makes it (assembly_common.c:1141-1146)."""
    f = np.asarray(f, np.float64)
    m = np.sqrt((f[0] * f[0]) + (f[1] * f[1]) + (f[2] * f[2]))
    return np.array([f[0] / m, f[1] / m, f[2] / m])


def fibre_tensor(f, sigma_l: float, sigma_t: float) -> np.ndarray:
    """calc_tensor2 (assembly_common.c:1121-1139): D = (sl - st) f f^T + st I."""
    f = np.asarray(f, np.float64)
    D = np.empty((3, 3))
    for i in range(3):
        for j in range(3):
            D[i, j] = (f[i] * f[j]) * (sigma_l - sigma_t) + (sigma_t if i == j else 0.0)
    return np.array([D[0, 0], D[1, 1], D[2, 2], D[0, 1], D[0, 2], D[1, 2]])


def layered_fibre_tensors(nz: int, theta0_deg: float, theta1_deg: float, sigma_l: float, sigma_t: float) -> np.ndarray:
    """Per-layer tensors for fibres in the xy plane rotating linearly with depth:
    theta_k = theta0 + (theta1 - theta0) k / (nz - 1), f = (cos, sin, 0) normalised."""
    out = np.zeros((nz, 6))
    for k in range(nz):
        th = np.deg2rad(theta0_deg + (theta1_deg - theta0_deg) * (k / (nz - 1) if nz > 1 else 0.0))
        out[k] = fibre_tensor(normalise([np.cos(th), np.sin(th), 0.0]), sigma_l, sigma_t)
    return out


def splitmix64_uniform(seed: int, index: np.ndarray) -> np.ndarray:
    """U[0, 1) draws keyed by (seed, index): SplitMix64 (Steele, Lea & Flood,
    "Fast splittable pseudorandom number generators", OOPSLA 2014) of
    seed + (index + 1) * 0x9E3779B97F4A7C15, top 53 bits."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(index, np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def fibrosis_mask(shape, fraction: float, seed: int, chunk: int = 1 << 22) -> np.ndarray:
    """Conducting cells (1) of a box with Bernoulli(fraction) fibrotic cells (0),
    cell c fibrotic when its SplitMix64 draw keyed by (seed, c) is below
    fraction (c lexicographic, x fastest)."""
    n = int(np.prod(shape))
    out = np.empty(n, np.uint8)
    for c0 in range(0, n, chunk):
        idx = np.arange(c0, min(n, c0 + chunk), dtype=np.uint64)
        out[c0:c0 + len(idx)] = splitmix64_uniform(seed, idx) >= fraction
    return out


MAX_INSTANCES = 4096  # rows of a per-instance scaling table (kMaxInstances, ionic.h)


def make_desc(shape, h=(100.0, 100.0, 100.0), model=MODEL_TT2006, precision=FP64, dt=0.02, dt_ode=None, beta=0.14,
              cm=1.0, aniso=False, tensor=None, tensor_source=TENSOR_UNIFORM, g_cell=None, active=None,
              cg_tol=1e-16, cg_rel_tol=0.0, cg_max_its=200, jacobi=False, inner_rel=1e-4,
              stim=(0.0, 2.0, -50.0, 500.0, 0.0), stim_mask=None, ionic_scale=None, adaptive=None, alpha_dt=0.0,
              instance=None, instance_scale=None):
    """Build a TissueDesc; returns (desc, keepalive) where keepalive holds the arrays.

    stim_mask (n,) 0/1 replaces the x < stim_x_limit rule; ionic_scale (n, 4)
    float32 scales TT2006 GKr/GCaL/GKs or MS2003 tau_close per cell; instead,
    instance (n,) uint16 names each cell's row of instance_scale (k, 4),
    k <= 4096 (the form a batch uses: the table sits in LDS).
    adaptive: None, or (abs_tol, rel_tol) for the adaptive Euler ODE solver
    (min dt = dt_ode, max dt = dt; CRN 1998).  alpha_dt > 0: the dt of ALPHA
    when it differs from the loop's dt (a dt_ode above dt_pde)."""
    d = TissueDesc()
    d.nx, d.ny, d.nz = shape
    d.dx, d.dy, d.dz = h
    d.model, d.precision = model, precision
    d.dt_pde, d.dt_ode = dt, (dt if dt_ode is None else dt_ode)
    d.beta, d.cm = beta, cm
    d.aniso = int(aniso)
    if tensor is None:
        tensor = iso_tensor(0.00013)
    tensor = np.ascontiguousarray(tensor, np.float64)
    d.tensor_source = tensor_source
    d.tensor = tensor.ctypes.data
    keep = [tensor]
    if g_cell is not None:
        g_cell = np.ascontiguousarray(g_cell, np.float64)
        keep.append(g_cell)
        d.g_cell = g_cell.ctypes.data
    if active is not None:
        active = np.ascontiguousarray(active, np.uint8)
        keep.append(active)
        d.active = active.ctypes.data
    n = int(d.nx) * int(d.ny) * int(d.nz)
    if stim_mask is not None:
        stim_mask = np.ascontiguousarray(stim_mask, np.uint8).reshape(-1)
        if stim_mask.size != n:
            raise ValueError(f"stim_mask has {stim_mask.size} entries for {n} cells")
        keep.append(stim_mask)
        d.stim_mask = stim_mask.ctypes.data
    if ionic_scale is not None:
        ionic_scale = np.ascontiguousarray(ionic_scale, np.float32).reshape(-1)
        if ionic_scale.size != 4 * n:
            raise ValueError(f"ionic_scale has {ionic_scale.size} entries for 4 x {n} cells")
        keep.append(ionic_scale)
        d.ionic_scale = ionic_scale.ctypes.data
    if (instance is None) != (instance_scale is None):
        raise ValueError("instance and instance_scale go together")
    if instance is not None:
        if ionic_scale is not None:
            raise ValueError("give ionic_scale or instance/instance_scale, not both")
        instance = np.ascontiguousarray(instance, np.uint16).reshape(-1)
        instance_scale = np.ascontiguousarray(instance_scale, np.float32).reshape(-1, 4)
        k = instance_scale.shape[0]
        if instance.size != n:
            raise ValueError(f"instance has {instance.size} entries for {n} cells")
        if not 1 <= k <= MAX_INSTANCES:
            raise ValueError(f"instance_scale has {k} rows; 1 to {MAX_INSTANCES} allowed")
        if int(instance.max()) >= k:
            raise ValueError(f"instance names row {int(instance.max())} of a {k}-row table")
        keep += [instance, instance_scale]
        d.instance, d.instance_scale, d.n_instances = instance.ctypes.data, instance_scale.ctypes.data, k
    if adaptive is not None:
        d.ode_adaptive = 1
        d.ode_abs_tol, d.ode_rel_tol = adaptive
    d.cg_tol, d.cg_rel_tol, d.cg_max_its, d.cg_jacobi = cg_tol, cg_rel_tol, cg_max_its, int(jacobi)
    d.inner_rel = inner_rel
    d.alpha_dt = alpha_dt
    d.stim_start, d.stim_duration, d.stim_current, d.stim_x_limit, d.stim_period = stim
    return d, keep


class Tissue:
    """Device-resident monodomain run on a structured box (see make_desc for arguments)."""

    def __init__(self, shape, **kw):
        self.desc, self._keep = make_desc(shape, **kw)
        self._h = C.c_void_p()
        check(lib().mono_tissue_create(C.byref(self.desc), C.byref(self._h)), "mono_tissue_create")
        self.n = int(lib().mono_tissue_num_cells(self._h))
        self.neq = int(lib().mono_tissue_neq(self._h))
        self.shape = tuple(shape)

    def close(self) -> None:
        if self._h:
            check(lib().mono_tissue_destroy(self._h), "mono_tissue_destroy")
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, k: int = 1) -> int:
        """Advance k loop iterations; returns 1 if a NaN stopped the run."""
        return check(lib().mono_tissue_step(self._h, k), "mono_tissue_step")

    def sync(self) -> None:
        check(lib().mono_tissue_sync(self._h), "mono_tissue_sync")

    def set_timing(self, on: bool) -> None:
        check(lib().mono_tissue_set_timing(self._h, int(on)), "mono_tissue_set_timing")

    def graph_counts(self) -> tuple:
        """(graph launches, graph recordings) of the solver's pre-poll launch sequences."""
        launches, captures = C.c_int64(), C.c_int64()
        check(lib().mono_tissue_graph_counts(self._h, C.byref(launches), C.byref(captures)), "mono_tissue_graph_counts")
        return launches.value, captures.value

    @property
    def v(self) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        check(lib().mono_tissue_get_v(self._h, ptr(out)), "mono_tissue_get_v")
        return out

    @v.setter
    def v(self, val) -> None:
        a = np.ascontiguousarray(val, np.float64)
        assert a.size == self.n
        check(lib().mono_tissue_set_v(self._h, ptr(a)), "mono_tissue_set_v")

    @property
    def state(self) -> np.ndarray:
        """(neq, n) float64; row 0 is not maintained on the fast path (V lives in .v)."""
        out = np.empty((self.neq, self.n), np.float64)
        check(lib().mono_tissue_get_state(self._h, ptr(out)), "mono_tissue_get_state")
        return out

    @state.setter
    def state(self, val) -> None:
        a = np.ascontiguousarray(val, np.float64)
        assert a.shape == (self.neq, self.n)
        check(lib().mono_tissue_set_state(self._h, ptr(a)), "mono_tissue_set_state")

    def save_state(self, directory: str) -> None:
        """Checkpoint into directory (created if missing): the reference's three
        checkpoint files plus this library's own (save_state.c:16-155)."""
        os.makedirs(directory, exist_ok=True)
        check(lib().mono_tissue_save_state(self._h, os.fsencode(directory)), "mono_tissue_save_state")

    def restore_state(self, directory: str) -> None:
        """Restart from a checkpoint written here or by the reference
        (restore_state.c:45-256)."""
        check(lib().mono_tissue_restore_state(self._h, os.fsencode(directory)), "mono_tissue_restore_state")

    @property
    def stats(self) -> dict:
        s = Stats()
        check(lib().mono_tissue_get_stats(self._h, C.byref(s)), "mono_tissue_get_stats")
        return s.as_dict()

    def device_ptrs(self):
        v, sv, b = C.c_void_p(), C.c_void_p(), C.c_void_p()
        pitch = C.c_size_t()
        check(lib().mono_tissue_device_ptrs(self._h, C.byref(v), C.byref(sv), C.byref(pitch), C.byref(b)),
              "mono_tissue_device_ptrs")
        return v.value, sv.value, pitch.value, b.value

    # activation times and APDs (save_with_activation_times, save_mesh.c:730-913)
    def observe_setup(self, activation_threshold=-30.0, apd_threshold=-83.0, time_threshold=10.0,
                      max_events=8, vm_double=False) -> None:
        check(lib().mono_tissue_observe_setup(self._h, activation_threshold, apd_threshold, time_threshold,
                                              max_events, int(vm_double)), "mono_tissue_observe_setup")

    def observe(self) -> None:
        """One save call at the current time (every print_rate steps, before stepping)."""
        check(lib().mono_tissue_observe(self._h), "mono_tissue_observe")

    # first-action-potential tracker (activation time, peak, APD90)
    def apd_setup(self, activation_threshold=-30.0) -> None:
        """Record the current Vm as each cell's resting potential and start
        tracking its first action potential (sample after every step)."""
        check(lib().mono_tissue_apd_setup(self._h, activation_threshold), "mono_tissue_apd_setup")

    def apd_sample(self) -> None:
        check(lib().mono_tissue_apd_sample(self._h), "mono_tissue_apd_sample")

    def apd(self) -> dict:
        """{'lat', 'vmax', 'apd90'}: (n,) float64, NaN where absent."""
        out = {k: np.empty(self.n, np.float64) for k in ("lat", "vmax", "apd90")}
        check(lib().mono_tissue_apd_get(self._h, ptr(out["lat"]), ptr(out["vmax"]), ptr(out["apd90"])),
              "mono_tissue_apd_get")
        return out

    def activations(self) -> dict:
        n, k = self.n, lib().mono_tissue_observe_events(self._h)
        n_act = np.zeros(n, np.int32)
        n_apd = np.zeros(n, np.int32)
        act = np.zeros((k, n), np.float32)
        apd = np.zeros((k, n), np.float32)
        check(lib().mono_tissue_get_activations(self._h, n_act.ctypes.data, act.ctypes.data, n_apd.ctypes.data,
                                                apd.ctypes.data), "mono_tissue_get_activations")
        return {"n_act": n_act, "act": act.T.copy(), "n_apd": n_apd, "apd": apd.T.copy()}


def write_acm(path, shape, h, n_act, act, apd, last_call=True, active=None) -> None:
    """The reference's activation_info .acm text (save_mesh.c:789-905): a flag
    line, then per active cell 'cx,cy,cz,hx,hy,hz n [ t.. ] [ apd.. ]' (cells in
    lexicographic order here; readers key cells by their centre)."""
    nx, ny, nz = shape
    dx, dy, dz = h
    with open(path, "w") as fh:
        fh.write(f"{int(bool(last_call))}\n")
        for c in range(nx * ny * nz):
            if active is not None and not active[c]:
                continue
            x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
            cx, cy, cz = (x + 0.5) * dx, (y + 0.5) * dy, (z + 0.5) * dz
            na = int(n_act[c])
            ts = " ".join("%lf" % v for v in act[c][:na] if np.isfinite(v))
            ap = " ".join("%lf" % v for v in apd[c] if np.isfinite(v))
            fh.write("%g,%g,%g,%g,%g,%g %d [ %s%s] [ %s%s]\n" % (cx, cy, cz, dx / 2.0, dy / 2.0, dz / 2.0, na, ts,
                                                                  " " if ts else "", ap, " " if ap else ""))


def ionic_run(model: int, states: np.ndarray, stim: np.ndarray | None, dt: float, nsteps: int = 1,
              precision: int = FP64) -> np.ndarray:
    """SOLVE_MODEL_ODES on AoS states (n, neq); returns the new states."""
    s = np.array(states, np.float64, copy=True, order="C")
    neq = model_neq(model)
    if s.ndim != 2 or s.shape[1] != neq:
        raise ValueError(f"states must have shape (n, {neq}) for model {model}, got {s.shape}")
    n = s.shape[0]
    st = None if stim is None else np.ascontiguousarray(stim, np.float64).reshape(-1)
    if st is not None and st.size != n:
        raise ValueError(f"stim has {st.size} entries for {n} cells")
    check(lib().mono_ionic_run(model, precision, ptr(s), n, ptr(st), dt, nsteps), "mono_ionic_run")
    return s


def ionic_run_adaptive(model: int, states: np.ndarray, aux: np.ndarray, stim, current_t: float, min_dt: float,
                       max_dt: float, abs_tol: float = 1e-5, rel_tol: float = 1e-5, precision: int = FP64):
    """One adaptive-Euler call (to current_t + min_dt) on AoS states (n, neq) and
    per-cell aux (n, 3) = (dt, time_new, previous_dt); returns new (states, aux)."""
    s = np.array(states, np.float64, copy=True, order="C")
    neq = model_neq(model)
    if s.ndim != 2 or s.shape[1] != neq:
        raise ValueError(f"states must have shape (n, {neq}) for model {model}, got {s.shape}")
    n = s.shape[0]
    a = np.array(aux, np.float64, copy=True)
    if a.size != 3 * n:
        raise ValueError(f"aux must have shape ({n}, 3), got {a.shape}")
    a = a.reshape(n, 3)
    dt, tn, pdt = (np.ascontiguousarray(a[:, k]) for k in range(3))
    st = None if stim is None else np.ascontiguousarray(stim, np.float64).reshape(-1)
    if st is not None and st.size != n:
        raise ValueError(f"stim has {st.size} entries for {n} cells")
    check(lib().mono_ionic_run_adaptive(model, precision, ptr(s), n, ptr(st), current_t, ptr(dt), ptr(tn), ptr(pdt), min_dt,
                                        max_dt, abs_tol, rel_tol), "mono_ionic_run_adaptive")
    return s, np.stack([dt, tn, pdt], axis=1)


def assemble(shape, **kw):
    """ASSEMBLY_MATRIX rows as (n, 27) FP64 stencils; returns (coef, alpha)."""
    d, keep = make_desc(shape, **kw)
    n = shape[0] * shape[1] * shape[2]
    out = np.zeros(27 * n, np.float64)
    a = C.c_double()
    check(lib().mono_assemble(C.byref(d), ptr(out), C.byref(a)), "mono_assemble")
    return out.reshape(27, n).T.copy(), a.value


def cg_stencil(shape, coef27, b, x0, tol=1e-16, rel_tol=0.0, max_its=200, jacobi=False, precision=FP64, active=None):
    """CG on an (n, 27) stencil operator; returns (x, iterations, error)."""
    nx, ny, nz = shape
    n = nx * ny * nz
    coef27 = np.asarray(coef27, np.float64)
    if coef27.size != 27 * n:
        raise ValueError(f"coef27 must have shape ({n}, 27), got {coef27.shape}")
    c = np.ascontiguousarray(coef27.reshape(n, 27).T)
    bb = np.ascontiguousarray(b, np.float64).reshape(-1)
    x = np.array(x0, np.float64, copy=True).reshape(-1)
    act = None if active is None else np.ascontiguousarray(active, np.uint8).reshape(-1)
    if bb.size != n or x.size != n or (act is not None and act.size != n):
        raise ValueError(f"b, x0 and active must have {n} entries")
    its, err = C.c_int(), C.c_double()
    check(lib().mono_cg_stencil(nx, ny, nz, ptr(c), ptr(act), ptr(bb), ptr(x), tol, rel_tol, max_its, int(jacobi),
                                precision, C.byref(its), C.byref(err)), "mono_cg_stencil")
    return x, its.value, err.value


def _check_csr(rp, cc, vv, n_vec):
    """Lengths of a CSR operator and its vectors (the C side checks the
    row_ptr order and column range)."""
    n = len(rp) - 1
    if n < 1:
        raise ValueError("row_ptr needs at least two entries")
    if rp[0] != 0 or np.any(np.diff(rp) < 0):
        raise ValueError("row_ptr must start at 0 and never decrease")
    if len(cc) != rp[-1] or len(vv) != rp[-1]:
        raise ValueError(f"col and val must have row_ptr[-1] = {rp[-1]} entries (got {len(cc)}, {len(vv)})")
    for k, m in n_vec.items():
        if m != n:
            raise ValueError(f"{k} has {m} entries for {n} rows")
    return n


def cg_csr(row_ptr, col, val, b, x0, tol=1e-16, max_its=200, jacobi=False):
    """SOLVE_LINEAR_SYSTEM on a CSR matrix (upload, solve, free); returns (x, iterations, error)."""
    rp = np.ascontiguousarray(row_ptr, np.int64).reshape(-1)
    cc = np.ascontiguousarray(col, np.int32).reshape(-1)
    vv = np.ascontiguousarray(val, np.float64).reshape(-1)
    bb = np.ascontiguousarray(b, np.float64).reshape(-1)
    x = np.array(x0, np.float64, copy=True).reshape(-1)
    _check_csr(rp, cc, vv, {"b": bb.size, "x0": x.size})
    its, err = C.c_int(), C.c_double()
    check(lib().mono_cg_csr(len(rp) - 1, ptr(rp), ptr(cc), ptr(vv), ptr(bb), ptr(x), tol, max_its, int(jacobi),
                            C.byref(its), C.byref(err)), "mono_cg_csr")
    return x, its.value, err.value


class Csr:
    """A CSR operator kept on the device between solves (INIT_LINEAR_SYSTEM /
    SOLVE_LINEAR_SYSTEM / END_LINEAR_SYSTEM with persistent data)."""

    def __init__(self, row_ptr, col, val):
        self._rp = np.ascontiguousarray(row_ptr, np.int64).reshape(-1)
        cc = np.ascontiguousarray(col, np.int32).reshape(-1)
        vv = np.ascontiguousarray(val, np.float64).reshape(-1)
        self.n = _check_csr(self._rp, cc, vv, {})
        self._h = C.c_void_p()
        check(lib().mono_csr_create(self.n, ptr(self._rp), ptr(cc), ptr(vv), C.byref(self._h)), "mono_csr_create")

    def set_alpha(self, alpha) -> None:
        a = np.ascontiguousarray(alpha, np.float64).reshape(-1)
        if a.size != self.n:
            raise ValueError(f"alpha has {a.size} entries for {self.n} rows")
        check(lib().mono_csr_set_alpha(self._h, ptr(a)), "mono_csr_set_alpha")

    def update_rhs(self, vm) -> None:
        """b = alpha * vm on the device (UPDATE_MONODOMAIN); the next solve(b=None) uses it."""
        v = np.ascontiguousarray(vm, np.float64).reshape(-1)
        if v.size != self.n:
            raise ValueError(f"vm has {v.size} entries for {self.n} rows")
        check(lib().mono_csr_update_rhs(self._h, ptr(v)), "mono_csr_update_rhs")

    def rhs(self) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        check(lib().mono_csr_get_rhs(self._h, ptr(out)), "mono_csr_get_rhs")
        return out

    def solve(self, b, x0, tol=1e-16, max_its=200, jacobi=False):
        """Returns (x, iterations, error); b=None uses the device right-hand side."""
        bb = None if b is None else np.ascontiguousarray(b, np.float64).reshape(-1)
        x = np.array(x0, np.float64, copy=True).reshape(-1)
        if x.size != self.n or (bb is not None and bb.size != self.n):
            raise ValueError(f"b and x0 must have {self.n} entries")
        its, err = C.c_int(), C.c_double()
        check(lib().mono_csr_solve(self._h, ptr(bb), ptr(x), tol, max_its, int(jacobi), C.byref(its), C.byref(err)),
              "mono_csr_solve")
        return x, its.value, err.value

    def close(self) -> None:
        if self._h:
            check(lib().mono_csr_destroy(self._h), "mono_csr_destroy")
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def update_rhs(alpha, vm) -> np.ndarray:
    """UPDATE_MONODOMAIN on host arrays: b = alpha * vm (computed on the device)."""
    a = np.ascontiguousarray(alpha, np.float64).reshape(-1)
    v = np.ascontiguousarray(vm, np.float64).reshape(-1)
    if a.size != v.size:
        raise ValueError("alpha and vm must have the same length")
    b = np.empty(v.size, np.float64)
    check(lib().mono_update_rhs(v.size, ptr(a), ptr(v), ptr(b)), "mono_update_rhs")
    return b


def device_count() -> int:
    return lib().mono_device_count()
