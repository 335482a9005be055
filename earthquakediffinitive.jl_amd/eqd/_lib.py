"""ctypes binding of libeqd.so (the C ABI declared in include/eqd.h).

The library is built in-tree by __graft_entry__.build() (or `make -C
earthquakediffinitive.jl_amd/csrc`). Nothing here falls back to a CPU path: if the
library or a gfx950 device is missing, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libeqd.so")

EQD_OK, EQD_ERR_ARG, EQD_ERR_GRID, EQD_ERR_DEVICE, EQD_ERR_NOMEM, EQD_ERR_RANGE, EQD_ERR_NEWTON = range(7)
PREC_FP64, PREC_MIXED = 0, 1
PRECOND_NONE, PRECOND_JACOBI = 0, 1
ENGINE_BATCHED, ENGINE_PLAIN = 0, 1
STIFFNESS_EXACT, STIFFNESS_TOEPLITZ = 0, 1
KS_DIRECT, KS_FFT = 0, 1

# Every symbol include/eqd.h declares (checked by tests/test_capi_symbols.py).
EXPORTED = [
    "eqd_cg_opts_default", "eqd_operator_create", "eqd_operator_destroy", "eqd_operator_info",
    "eqd_fault_grid_axes", "eqd_apply_A", "eqd_build_rhs", "eqd_jacobi_preconditioner", "eqd_solver_create",
    "eqd_solver_destroy",
    "eqd_solver_duplicate", "eqd_solver_merge_stats", "eqd_solver_report", "eqd_split_node_solve",
    "eqd_shear_traction", "eqd_shear_traction_noalloc", "eqd_fault_stiffness", "eqd_fault_stiffness_device", "eqd_solver_last_timing", "eqd_solver_kernel_timing",
    "eqd_tables_export", "eqd_version", "eqd_last_error",
    "eqd_bp8_default_params", "eqd_bp8_create", "eqd_bp8_create_device", "eqd_bp8_create_from_solver",
    "eqd_bp8_destroy", "eqd_bp8_sizes", "eqd_bp8_stiffness", "eqd_bp8_operators", "eqd_bp8_initial_state",
    "eqd_bp8_evaluate", "eqd_bp8_rhs", "eqd_bp8_stress_report", "eqd_bp8_reset_cache", "eqd_bp8_integrate", "eqd_bp8_integrate_imex",
    "eqd_bp8_saved_device", "eqd_bp8_saved_vmax", "eqd_bp8_saved_outputs", "eqd_bp8_set_ks_method",
    "eqd_bp8_ks", "eqd_bp8_ks_timing",
    "eqd_server_create", "eqd_server_destroy", "eqd_server_add_grid", "eqd_server_submit_columns",
    "eqd_server_submit_forward", "eqd_server_step", "eqd_server_completed", "eqd_server_result",
    "eqd_server_grid_info",
    "eqd_bp8_ensemble_create", "eqd_bp8_ensemble_destroy", "eqd_bp8_ensemble_integrate",
    "eqd_bp8_ensemble_integrate_imex", "eqd_bp8_ensemble_saved", "eqd_bp8_ensemble_saved_vmax", "eqd_bp8_ensemble_gemm_timing",
]

INJECTION_GAUSSIAN, INJECTION_PEACEMAN = 0, 1

BP8_PARAM_NAMES = ["mu", "nu", "rho", "c_s", "a", "b", "D_RS", "V_star", "f_star", "sigma0", "tau_init",
                   "V_init", "V_zero", "Q0", "L_gauss", "L_fwid", "beta", "phi", "k", "viscosity", "alpha",
                   "S_well", "r_well", "l_f", "dz", "t_off", "t_f", "sigma_min"]


class BP8Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in BP8_PARAM_NAMES]


class BP8Stats(C.Structure):
    _fields_ = [("naccept", C.c_int64), ("nreject", C.c_int64), ("nrhs", C.c_int64), ("nsaved", C.c_int64),
                ("retcode", C.c_int), ("ms", C.c_double)]


class GridDesc(C.Structure):
    _fields_ = [("n1", C.c_int), ("n23", C.c_int), ("L_normal", C.c_double), ("L_fault", C.c_double),
                ("lam", C.c_double), ("mu", C.c_double), ("l_f", C.c_double)]


class OpInfo(C.Structure):
    _fields_ = [("nf", C.c_int64), ("n2f", C.c_int64), ("n3f", C.c_int64), ("ntot", C.c_int64),
                ("nreduced", C.c_int64), ("h1", C.c_double), ("h2", C.c_double)]


class CgOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("itmax", C.c_int64), ("precond", C.c_int),
                ("precision", C.c_int), ("engine", C.c_int), ("tile_k", C.c_int),
                ("max_cols_in_flight", C.c_int), ("reliable_interval", C.c_int), ("graph_iters", C.c_int), ("timing", C.c_int)]


class Report(C.Structure):
    _fields_ = [("solves", C.c_int64), ("iterations", C.c_int64), ("max_iterations", C.c_int64),
                ("unconverged", C.c_int64), ("mean_iterations", C.c_double), ("precond", C.c_int),
                ("tile_k", C.c_int)]


class EqdError(RuntimeError):
    """An eqd_status other than EQD_OK; `code` is the status."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libeqd.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        L.eqd_cg_opts_default.argtypes = [C.POINTER(CgOpts)]
        L.eqd_cg_opts_default.restype = None
        L.eqd_operator_create.argtypes = [C.POINTER(GridDesc), C.POINTER(vp)]
        L.eqd_operator_destroy.argtypes = [vp]
        L.eqd_operator_info.argtypes = [vp, C.POINTER(OpInfo)]
        L.eqd_fault_grid_axes.argtypes = [vp, dp, dp]
        L.eqd_apply_A.argtypes = [vp, dp, dp]
        L.eqd_jacobi_preconditioner.argtypes = [C.POINTER(GridDesc), dp, C.c_int64]
        L.eqd_build_rhs.argtypes = [vp, dp, dp, dp]
        L.eqd_solver_create.argtypes = [vp, C.POINTER(CgOpts), C.POINTER(vp)]
        L.eqd_solver_destroy.argtypes = [vp]
        L.eqd_solver_duplicate.argtypes = [vp, C.POINTER(vp)]
        L.eqd_solver_merge_stats.argtypes = [vp, vp]
        L.eqd_solver_report.argtypes = [vp, C.POINTER(Report)]
        L.eqd_split_node_solve.argtypes = [vp, dp, dp]
        L.eqd_shear_traction.argtypes = [vp, dp, dp, dp, dp]
        L.eqd_shear_traction_noalloc.argtypes = [vp, dp, dp, dp, dp]
        L.eqd_fault_stiffness.argtypes = [vp, C.c_int, vp, C.c_int, vp, C.c_int64]
        L.eqd_fault_stiffness_device.argtypes = [vp, C.c_int, vp, C.c_int, vp, C.c_int64]
        L.eqd_solver_last_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                             C.POINTER(C.c_int64)]
        L.eqd_solver_kernel_timing.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                               C.POINTER(C.c_int64)]
        L.eqd_tables_export.argtypes = [C.POINTER(GridDesc), C.c_int, vp, vp, vp]
        L.eqd_tables_export.restype = C.c_int64
        i64p = C.POINTER(C.c_int64)
        L.eqd_bp8_default_params.argtypes = [C.POINTER(BP8Params)]
        L.eqd_bp8_default_params.restype = None
        L.eqd_bp8_create.argtypes = [C.POINTER(BP8Params), dp, C.c_int64, C.c_int, C.c_int, C.POINTER(vp)]
        L.eqd_bp8_create_device.argtypes = [C.POINTER(BP8Params), vp, C.c_int64, C.c_int, C.c_int, C.POINTER(vp)]
        L.eqd_bp8_create_from_solver.argtypes = [vp, C.c_int, C.POINTER(BP8Params), C.c_int, C.POINTER(vp)]
        L.eqd_bp8_destroy.argtypes = [vp]
        L.eqd_bp8_sizes.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.eqd_bp8_stiffness.argtypes = [vp, dp, C.c_int64]
        L.eqd_bp8_operators.argtypes = [vp, vp, vp, vp]
        L.eqd_bp8_initial_state.argtypes = [vp, dp]
        L.eqd_bp8_evaluate.argtypes = [vp, dp, C.c_double, vp, vp, vp, vp, vp]
        L.eqd_bp8_rhs.argtypes = [vp, dp, C.c_double, dp]
        L.eqd_bp8_stress_report.argtypes = [vp, C.POINTER(C.c_double), i64p, vp]
        L.eqd_bp8_reset_cache.argtypes = [vp]
        L.eqd_bp8_integrate.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_int64,
                                        vp, vp, C.POINTER(BP8Stats)]
        L.eqd_bp8_integrate_imex.argtypes = L.eqd_bp8_integrate.argtypes
        L.eqd_bp8_saved_device.argtypes = [vp, C.POINTER(vp), i64p, i64p]
        L.eqd_bp8_saved_vmax.argtypes = [vp, dp, C.c_int64]
        L.eqd_bp8_saved_outputs.argtypes = [vp, C.c_int64, vp, C.c_int, vp, vp]
        pd = C.POINTER(C.c_double)
        L.eqd_bp8_set_ks_method.argtypes = [vp, C.c_int]
        L.eqd_bp8_ks.argtypes = [vp, dp, dp]
        L.eqd_bp8_ks_timing.argtypes = [vp, C.c_int, pd]
        L.eqd_bp8_ensemble_create.argtypes = [vp, C.c_int, vp, vp, vp, C.POINTER(vp)]
        L.eqd_bp8_ensemble_destroy.argtypes = [vp]
        L.eqd_bp8_ensemble_integrate.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, vp,
                                                 C.c_int64, C.c_int64, C.POINTER(BP8Stats), pd]
        L.eqd_bp8_ensemble_integrate_imex.argtypes = L.eqd_bp8_ensemble_integrate.argtypes
        L.eqd_bp8_ensemble_saved.argtypes = [vp, C.c_int, vp, vp, C.POINTER(C.c_int64)]
        L.eqd_bp8_ensemble_saved_vmax.argtypes = [vp, vp, C.c_int64, pd]
        L.eqd_bp8_ensemble_gemm_timing.argtypes = [vp, C.c_int, pd, pd]
        L.eqd_server_create.argtypes = [C.POINTER(CgOpts), C.c_size_t, C.POINTER(vp)]
        L.eqd_server_destroy.argtypes = [vp]
        L.eqd_server_add_grid.argtypes = [vp, C.POINTER(GridDesc), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.eqd_server_submit_columns.argtypes = [vp, C.c_int, vp, C.c_int, i64p]
        L.eqd_server_submit_forward.argtypes = [vp, C.c_int, dp, dp, i64p]
        L.eqd_server_step.argtypes = [vp, i64p, i64p]
        L.eqd_server_completed.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
        L.eqd_server_result.argtypes = [vp, C.c_int64, vp, C.c_int64, vp, C.POINTER(C.c_int)]
        L.eqd_server_grid_info.argtypes = [vp, C.c_int, C.POINTER(OpInfo), C.POINTER(C.c_int),
                                           C.POINTER(C.c_double)]
        L.eqd_version.restype = C.c_char_p
        L.eqd_last_error.argtypes = [C.c_char_p, C.c_size_t]
        L.eqd_last_error.restype = C.c_size_t
        _lib = L
    return _lib


def check(rc: int):
    if rc != EQD_OK:
        buf = C.create_string_buffer(1024)
        lib().eqd_last_error(buf, 1024)
        raise EqdError(rc, buf.value.decode(errors="replace"))


def default_opts() -> CgOpts:
    o = CgOpts()
    lib().eqd_cg_opts_default(C.byref(o))
    return o


TABLE_IDS = {"AXs": 0, "AXo": 1, "MX": 2, "Ts": 3, "To": 4, "D1": 5, "D1s": 6, "D1full": 7, "D1ext": 8,
             "Hx": 100, "Hy": 101, "Jt": 102, "Jn": 103, "Dn": 104, "Jext": 105}


def tables_export(g: GridDesc, name: str):
    """Host-built table `name` (no device needed): (start, width, coef[nrows,16]) or a vector."""
    which = TABLE_IDS[name]
    n = lib().eqd_tables_export(C.byref(g), which, None, None, None)
    if n < 0:
        check(EQD_ERR_ARG)
    if which >= 100:
        v = np.zeros(max(n, 1))
        lib().eqd_tables_export(C.byref(g), which, None, None, v.ctypes.data)
        return v[0] if which == 105 else v
    start = np.zeros(n, np.int32); width = np.zeros(n, np.int32); coef = np.zeros((n, 16))
    lib().eqd_tables_export(C.byref(g), which, start.ctypes.data, width.ctypes.data, coef.ctypes.data)
    return start, width, coef
