"""ctypes binding of libsfhip.so (the C ABI declared in include/sf_hip.h).

The library is built in place (``stokesflow.jl_amd/libsfhip.so``).  There is no CPU
fallback: when the library is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libsfhip.so")

SF_OK = 0
SF_ERR_ARG = -1
SF_ERR_HIP = -2
SF_ERR_NONFINITE = -3
SF_ERR_SIZE = -4

# every exported symbol of include/sf_hip.h
EXPORTS = (
    "sf_version", "sf_last_error", "sf_ctx_create", "sf_ctx_destroy", "sf_ctx_set_stream", "sf_ctx_set_pt_tuning", "sf_ctx_set_pt_fusion", "sf_save_stats", "sf_save_markers",
    "sf_save_grid",
    "sf_solve_stokes",
    "sf_pt_iterations", "sf_marker_to_grid", "sf_move_markers_rk4", "sf_step", "sf_init_markers",
    "sf_plume_properties", "sf_rt_properties", "sf_solve_stokes_refined", "sf_solve_stokes_batched",
    "sf_marker_cell_keys", "sf_ensemble_create", "sf_ensemble_step", "sf_ensemble_destroy",
    "sf_ctx_autotune_pt", "sf_probe_dram", "sf_markers_to_grids3", "sf_ctx_time_passes", "sf_ctx_pass_times",
    "sf_solve_stokes_noalloc", "sf_lattice_markers", "sf_marker_properties", "sf_batch_autotune",
    "sf_batch_tuning_clear", "sf_graph_create", "sf_graph_step", "sf_graph_destroy", "sf_set_grid_wait_limit",
)


class SolveOut(C.Structure):
    _fields_ = [
        ("dt_new", C.c_double),
        ("err", C.c_double),
        ("t_eff", C.c_double),
        ("runtime", C.c_double),
        ("niter", C.c_int64),
        ("converged", C.c_int32),
        ("reserved", C.c_int32),
        ("dt_last", C.c_double),
        ("t_it", C.c_double),
        ("gbs_alg", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class RefineParams(C.Structure):
    _fields_ = [("rebase_interval", C.c_int64), ("iter_max", C.c_int64), ("verbose", C.c_int32),
                ("reserved", C.c_int32)]


class RefineOut(C.Structure):
    _fields_ = [("n_outer", C.c_int32), ("reserved", C.c_int32), ("n_fp32_iter", C.c_int64),
                ("n_fp64_iter", C.c_int64), ("err_fp64", C.c_double), ("seconds", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class M2GOut(C.Structure):
    _fields_ = [("n_bad_markers", C.c_int64), ("n_empty_nodes", C.c_int64)]


class RK4Out(C.Structure):
    _fields_ = [("n_clamped", C.c_int64)]


_vp = C.c_void_p
_d = C.c_double
_i = C.c_int
_i64 = C.c_int64


class StepArgs(C.Structure):
    _fields_ = [
        ("nx", _i), ("ny", _i), ("dx", _d), ("dy", _d), ("g_y", _d), ("maxdisp", _d), ("eps", _d),
        ("iter_max", _i64), ("use_fss", _i), ("dt", _d),
        ("P", _vp), ("Vx_pad", _vp), ("Vy_pad", _vp), ("rho_vy", _vp), ("mu_b", _vp), ("mu_p", _vp),
        ("dVxdtau", _vp), ("dVydtau", _vp), ("Vx_s", _vp), ("Vy_s", _vp),
        ("x_m", _vp), ("y_m", _vp), ("rho_m", _vp), ("mu_m", _vp), ("nm", _i64),
    ]


class BatchInstance(C.Structure):
    """sf_batch_instance: one problem of a batched ensemble."""
    _fields_ = [
        ("P", _vp), ("Vx_pad", _vp), ("Vy_pad", _vp), ("dVxdtau", _vp), ("dVydtau", _vp), ("Vx_s", _vp),
        ("Vy_s", _vp), ("rho_vy", _vp), ("mu_b", _vp), ("mu_p", _vp),
        ("nx", _i), ("ny", _i), ("dx", _d), ("dy", _d), ("g_y", _d), ("dt", _d), ("maxdisp", _d), ("eps", _d),
        ("iter_max", _i64), ("use_fss", _i), ("reserved", _i),
    ]


class PTTuning(C.Structure):
    _fields_ = [("rows", _i), ("depth", _i), ("max_waves", _i), ("iters_per_launch", _i), ("us_per_iter", _d)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class PassTime(C.Structure):
    _fields_ = [("precision", C.c_int32), ("iters_per_launch", C.c_int32), ("launches", C.c_int64),
                ("seconds", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchTuning(C.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("iters_per_launch", _i), ("rows", _i), ("instances", _i), ("kept", _i),
                ("ns_per_node_iter", _d)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Shape(C.Structure):
    """sf_shape: one shape of a material descriptor."""
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("a", C.c_double), ("b", C.c_double),
                ("c", C.c_double), ("d", C.c_double), ("e", C.c_double), ("rho", C.c_double), ("mu", C.c_double)]


SF_SHAPE_Y_BELOW, SF_SHAPE_Y_ABOVE, SF_SHAPE_CIRCLE, SF_SHAPE_RECT, SF_SHAPE_SINE_BELOW = 1, 2, 3, 4, 5
SF_MAX_SHAPES = 16

SF_BATCH_AUTO, SF_BATCH_PERSISTENT, SF_BATCH_WIDE, SF_BATCH_GRID = 0, 1, 2, 3


class BatchParams(C.Structure):
    _fields_ = [("rows", _i), ("depth", _i), ("threads", _i), ("max_blocks", _i), ("mode", _i), ("precision", _i),
                ("budget_mb", _d)]


_lib = None


class SFError(RuntimeError):
    """A libsfhip call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__(f"libsfhip error {code}: {msg}")
        self.code = code


def load():
    """Load libsfhip.so and declare every entry point; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libsfhip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.sf_version.restype = _i
    L.sf_last_error.restype = C.c_char_p
    L.sf_ctx_create.argtypes = [C.POINTER(_vp), _i, _i, _vp]
    L.sf_ctx_destroy.argtypes = [_vp]
    L.sf_ctx_set_stream.argtypes = [_vp, _vp]
    L.sf_ctx_set_pt_tuning.argtypes = [_vp, _i, _i, _i]
    L.sf_ctx_set_pt_fusion.argtypes = [_vp, _i, _i]
    L.sf_save_stats.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.c_int64]
    L.sf_save_markers.argtypes = [_vp, C.c_char_p, _vp, _vp, _vp, C.c_int64, _d, _d]
    L.sf_save_grid.argtypes = [_vp, C.c_char_p, _i, _i, _d, _d, _vp, _vp, _vp, _vp]
    L.sf_solve_stokes.argtypes = [_vp] + [_vp] * 19 + [_d, _d, _d, _i, _i, _d, _d, _i, _d, _i64, _i,
                                                        C.POINTER(SolveOut)]
    L.sf_solve_stokes_noalloc.argtypes = [_vp] + [_vp] * 19 + [_d, _d, _d, _i, _i, _d, _d, _i, _d, _i64, _i,
                                                                C.POINTER(SolveOut)]
    L.sf_ctx_time_passes.argtypes = [_vp, _i]
    L.sf_batch_autotune.argtypes = [_vp, C.POINTER(BatchInstance), _i, C.POINTER(BatchParams), _i,
                                    C.POINTER(BatchTuning), _i, C.POINTER(_i)]
    L.sf_batch_tuning_clear.argtypes = [_vp]
    L.sf_ctx_pass_times.argtypes = [_vp, C.POINTER(PassTime), _i, C.POINTER(_i)]
    L.sf_solve_stokes_refined.argtypes = [_vp] + [_vp] * 19 + [_d, _d, _d, _i, _i, _d, _d, _i, _d,
                                                                C.POINTER(RefineParams), C.POINTER(SolveOut),
                                                                C.POINTER(RefineOut)]
    L.sf_solve_stokes_batched.argtypes = [_vp, C.POINTER(BatchInstance), _i, C.POINTER(BatchParams),
                                          C.POINTER(SolveOut)]
    L.sf_pt_iterations.argtypes = [_vp, _i64] + [_vp] * 8 + [_d, _d, _d, _i, _i, _d]
    L.sf_marker_to_grid.argtypes = [_vp, _d, _d, _vp, _i, _i, _vp, _vp, _vp, _i64, _d, _d, _vp, _vp, _i, _i,
                                    C.POINTER(M2GOut)]
    L.sf_markers_to_grids3.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _i64, _d, _d,
                                       C.POINTER(M2GOut)]
    L.sf_probe_dram.argtypes = [_vp, _i64, C.POINTER(_d), C.POINTER(_d)]
    L.sf_ctx_autotune_pt.argtypes = [_vp, _i] + [_vp] * 8 + [_d, _d, _d, C.POINTER(PTTuning), C.POINTER(PTTuning), _i,
                                                             C.POINTER(_i)]
    L.sf_ensemble_create.argtypes = [_vp, C.POINTER(StepArgs), _i, C.POINTER(BatchParams), _i, C.POINTER(_vp)]
    L.sf_ensemble_step.argtypes = [_vp, _i, C.POINTER(SolveOut), C.POINTER(_d)]
    L.sf_ensemble_destroy.argtypes = [_vp]
    L.sf_graph_create.argtypes = [_vp, C.POINTER(StepArgs), _i, C.POINTER(_vp)]
    L.sf_graph_step.argtypes = [_vp, _i, C.POINTER(SolveOut), C.POINTER(_d)]
    L.sf_graph_destroy.argtypes = [_vp]
    L.sf_set_grid_wait_limit.argtypes = [_d]
    L.sf_marker_cell_keys.argtypes = [_vp, _d, _d, _i, _i, _vp, _vp, _i64, _d, _d, _vp, _vp, C.POINTER(M2GOut)]
    L.sf_move_markers_rk4.argtypes = [_vp, _vp, _vp, _i64, _vp, _i, _i, _vp, _i, _i, _d, _d, _d, _d, _d, _d, _d,
                                      _d, _d, C.POINTER(RK4Out)]
    L.sf_step.argtypes = [_vp, C.POINTER(StepArgs), C.POINTER(SolveOut), C.POINTER(M2GOut), C.POINTER(RK4Out)]
    L.sf_init_markers.argtypes = [_vp, _vp, _vp, _i, _i, _i, _d, _d, _d, C.c_uint64]
    L.sf_plume_properties.argtypes = [_vp, _vp, _vp, _i64, C.POINTER(_d), _vp, _vp]
    L.sf_rt_properties.argtypes = [_vp, _vp, _vp, _i64, C.POINTER(_d), _vp, _vp]
    L.sf_lattice_markers.argtypes = [_vp, _vp, _vp, _i, _i, _i, _d, _d, _d, _d, _i, C.c_uint64]
    L.sf_marker_properties.argtypes = [_vp, _vp, _vp, _i64, _d, _d, C.POINTER(Shape), _i, _d, _d, _vp, _vp]
    for name in EXPORTS:
        if name not in ("sf_version", "sf_last_error"):
            getattr(L, name).restype = _i
    _lib = L
    return L


def check(rc):
    if rc != SF_OK:
        msg = load().sf_last_error()
        raise SFError(rc, msg.decode() if msg else "")
    return rc
