"""ctypes binding of libmonoalg_hip.so (the C ABI declared in include/monoalg_hip.h).

The compute path is the HIP library; if it is missing this module raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libmonoalg_hip.so")

MODEL_TT2006, MODEL_ORD2011, MODEL_MS2003, MODEL_CRN1998 = 0, 1, 2, 3
FP64, FP32_REFINE = 0, 1
TENSOR_UNIFORM, TENSOR_PER_LAYER, TENSOR_PER_CELL = 0, 1, 2

# the symbols include/monoalg_hip.h declares
EXPORTED = (
    "mono_last_error", "mono_device_count", "mono_tissue_create", "mono_tissue_destroy", "mono_tissue_step",
    "mono_tissue_sync", "mono_tissue_num_cells", "mono_tissue_neq", "mono_tissue_get_v", "mono_tissue_set_v",
    "mono_tissue_get_state", "mono_tissue_set_state", "mono_tissue_save_state", "mono_tissue_restore_state", "mono_tissue_get_stats", "mono_tissue_set_timing", "mono_tissue_graph_counts",
    "mono_tissue_device_ptrs", "mono_ionic_run", "mono_model_neq", "mono_model_initial_state",
    "mono_model_initial_v", "mono_assemble", "mono_cg_stencil", "mono_cg_csr",
    "mono_tissue_observe_setup", "mono_tissue_observe", "mono_tissue_get_activations", "mono_tissue_observe_events", "mono_ionic_run_adaptive",
    "mono_sensor_gradient", "mono_csr_create", "mono_csr_num_rows", "mono_csr_set_alpha", "mono_csr_update_rhs",
    "mono_csr_get_rhs", "mono_csr_solve", "mono_csr_destroy", "mono_update_rhs", "mono_ode_gpu_alloc",
    "mono_ode_gpu_solve", "mono_device_free", "mono_tissue_apd_setup", "mono_tissue_apd_sample", "mono_tissue_apd_get",
    "mono_cg_csr_device_noalloc",
)


class MonoError(RuntimeError):
    pass


class TissueDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("model", C.c_int), ("precision", C.c_int),
        ("dt_pde", C.c_double), ("dt_ode", C.c_double),
        ("beta", C.c_double), ("cm", C.c_double),
        ("aniso", C.c_int), ("tensor_source", C.c_int),
        ("tensor", C.c_void_p), ("g_cell", C.c_void_p), ("active", C.c_void_p),
        ("cg_tol", C.c_double), ("cg_rel_tol", C.c_double), ("cg_max_its", C.c_int), ("cg_jacobi", C.c_int),
        ("inner_rel", C.c_double),
        ("stim_start", C.c_double), ("stim_duration", C.c_double), ("stim_current", C.c_double),
        ("stim_x_limit", C.c_double), ("stim_period", C.c_double),
        ("stim_mask", C.c_void_p), ("ionic_scale", C.c_void_p),
        ("ode_adaptive", C.c_int), ("ode_abs_tol", C.c_double), ("ode_rel_tol", C.c_double),
        ("alpha_dt", C.c_double),
        ("instance", C.c_void_p), ("instance_scale", C.c_void_p), ("n_instances", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("time", C.c_double), ("cg_iterations_total", C.c_int64),
        ("last_cg_iterations", C.c_int), ("last_cg_error", C.c_double), ("last_outer", C.c_int),
        ("inner_iterations_total", C.c_int64), ("outer_sweeps_total", C.c_int64), ("nan_detected", C.c_int),
        ("class_operator", C.c_int), ("stencil_points", C.c_int), ("ionic_ms_total", C.c_double),
        ("solve_ms_total", C.c_double), ("ionic_block", C.c_int),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib() -> C.CDLL:
    """Load the HIP library (raises MonoError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MonoError(f"{LIB_PATH} not found: build it with `make -C monoalg3d_c_amd/csrc` "
                        "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, dp, i64 = C.c_void_p, C.POINTER(C.c_double), C.c_int64
    L.mono_last_error.restype = C.c_char_p
    L.mono_device_count.restype = C.c_int
    L.mono_tissue_create.argtypes = [C.POINTER(TissueDesc), C.POINTER(vp)]
    L.mono_tissue_destroy.argtypes = [vp]
    L.mono_tissue_step.argtypes = [vp, C.c_int]
    L.mono_tissue_sync.argtypes = [vp]
    L.mono_tissue_num_cells.argtypes = [vp]
    L.mono_tissue_num_cells.restype = i64
    L.mono_tissue_neq.argtypes = [vp]
    L.mono_tissue_get_v.argtypes = [vp, vp]
    L.mono_tissue_set_v.argtypes = [vp, vp]
    L.mono_tissue_get_state.argtypes = [vp, vp]
    L.mono_tissue_set_state.argtypes = [vp, vp]
    L.mono_tissue_save_state.argtypes = [vp, C.c_char_p]
    L.mono_tissue_restore_state.argtypes = [vp, C.c_char_p]
    L.mono_tissue_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.mono_tissue_set_timing.argtypes = [vp, C.c_int]
    L.mono_tissue_graph_counts.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.mono_tissue_device_ptrs.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_size_t), C.POINTER(vp)]
    L.mono_ionic_run.argtypes = [C.c_int, C.c_int, vp, i64, vp, C.c_double, C.c_int]
    L.mono_sensor_gradient.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, vp, C.c_int, vp, vp, vp, vp]
    L.mono_ionic_run_adaptive.argtypes = [C.c_int, C.c_int, vp, i64, vp, C.c_double, vp, vp, vp] + [C.c_double] * 4
    L.mono_model_neq.argtypes = [C.c_int]
    L.mono_model_initial_state.argtypes = [C.c_int, vp]
    L.mono_model_initial_v.argtypes = [C.c_int]
    L.mono_model_initial_v.restype = C.c_double
    L.mono_assemble.argtypes = [C.POINTER(TissueDesc), vp, dp]
    L.mono_cg_stencil.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                  C.c_int, C.POINTER(C.c_int), dp]
    L.mono_tissue_observe_setup.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int]
    L.mono_tissue_observe.argtypes = [vp]
    L.mono_tissue_get_activations.argtypes = [vp, vp, vp, vp, vp]
    L.mono_tissue_observe_events.argtypes = [vp]
    L.mono_cg_csr.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), dp]
    L.mono_csr_create.argtypes = [C.c_int, vp, vp, vp, C.POINTER(vp)]
    L.mono_csr_num_rows.argtypes = [vp]
    L.mono_csr_set_alpha.argtypes = [vp, vp]
    L.mono_csr_update_rhs.argtypes = [vp, vp]
    L.mono_csr_get_rhs.argtypes = [vp, vp]
    L.mono_csr_solve.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), dp]
    L.mono_csr_destroy.argtypes = [vp]
    L.mono_update_rhs.argtypes = [i64, vp, vp, vp]
    L.mono_ode_gpu_alloc.argtypes = [C.c_int, i64, C.POINTER(vp), C.POINTER(C.c_size_t)]
    L.mono_ode_gpu_solve.argtypes = [C.c_int, vp, C.c_size_t, i64, vp, C.c_double, C.c_int]
    L.mono_device_free.argtypes = [vp]
    L.mono_tissue_apd_setup.argtypes = [vp, C.c_double]
    L.mono_tissue_apd_sample.argtypes = [vp]
    L.mono_tissue_apd_get.argtypes = [vp, vp, vp, vp]
    _lib = L
    return L


def check(rc: int, what: str) -> int:
    if rc < 0:
        msg = lib().mono_last_error()
        raise MonoError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data
