"""Host-side mirror of the reference's fault-stiffness API over libeqd's C ABI.

The names, arguments and error behaviour follow EarthquakeDiffinitive.jl:

  ElasticitySplitNode  CGSolver, split_node_solve, duplicate, merge_stats!, solver_report
                       (src/ElasticitySplitNode.jl:365-447)
  FaultResponse        FaultElasticity, shear_traction, fault_stiffness,
                       fault_stiffness_toeplitz, frictional_node_count, fault_grid_axes,
                       elastic_solver_report (src/FaultResponse.jl:82-393)

Everything numerical runs in libeqd.so on a gfx950 device; this module only moves
host arrays across the boundary. Julia callers use the same C ABI through `ccall`
(see julia/EqdDevice.jl).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import EqdError, check, lib

__all__ = [
    "GridPair", "bp8_grid_pair", "FaultElasticity", "CGSolver", "split_node_solve", "duplicate",
    "merge_stats", "solver_report", "shear_traction", "fault_stiffness", "fault_stiffness_toeplitz",
    "frictional_node_count", "fault_grid_axes", "elastic_solver_report", "jacobi_preconditioner", "EqdError",
]


@dataclass(frozen=True)
class GridPair:
    """The two SBP blocks g- = [-L_normal,0]x[-L_fault,L_fault]^2 and g+ = [0,L_normal]x[...]^2,
    each n1 x n23 x n23 nodes (src/BP8.jl:205-206; the test grids of test/fault_response_test.jl:16-17)."""
    n1: int
    n23: int
    L_normal: float
    L_fault: float


def bp8_grid_pair(dz: float, L_fault: float, L_normal: float, order: int = 4) -> GridPair:
    """Grid sizes as build_model computes them (src/BP8.jl:196-204), with its legality errors."""
    n1 = int(round(L_normal / dz)) + 1
    n23 = int(round(2 * L_fault / dz)) + 1
    n_min = 2 * order + 1
    if n1 < n_min:
        raise EqdError(_lib.EQD_ERR_GRID, f"L_normal/dz gives only {n1} points across the fault-normal direction; "
                                          f"SBP order {order} needs at least {n_min}.")
    if n23 < n_min:
        raise EqdError(_lib.EQD_ERR_GRID, f"2*L_fault/dz gives only {n23} points along the fault; "
                                          f"SBP order {order} needs at least {n_min}.")
    return GridPair(n1, n23, float(L_normal), float(L_fault))


def _desc(g: GridPair, lam: float, mu: float, l_f: float) -> _lib.GridDesc:
    return _lib.GridDesc(g.n1, g.n23, g.L_normal, g.L_fault, lam, mu, l_f)


_PREC = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}
_PRECOND = {"none": _lib.PRECOND_NONE, "jacobi": _lib.PRECOND_JACOBI}
_ENGINE = {"batched": _lib.ENGINE_BATCHED, "plain": _lib.ENGINE_PLAIN}


class _Operator:
    def __init__(self, g: GridPair, lam: float, mu: float, l_f: float):
        self.grids = g
        self.desc = _desc(g, lam, mu, l_f)
        h = C.c_void_p()
        check(lib().eqd_operator_create(C.byref(self.desc), C.byref(h)))
        self.handle = h
        info = _lib.OpInfo()
        check(lib().eqd_operator_info(h, C.byref(info)))
        self.nf, self.n2f, self.n3f = int(info.nf), int(info.n2f), int(info.n3f)
        self.ntot, self.nreduced = int(info.ntot), int(info.nreduced)
        self.h1, self.h2 = info.h1, info.h2

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and h.value:
            lib().eqd_operator_destroy(h)
            self.handle = None


class CGSolver:
    """CGSolver(A; rtol=1e-10, atol=0.0, itmax=0, precond=:none) (src/ElasticitySplitNode.jl:365),
    with the device options `precision` ("fp64" | "mixed"), `engine` ("batched" | "plain"),
    `tile_k`, `max_cols_in_flight`, `reliable_interval` and `graph_iters`."""

    def __init__(self, op: _Operator, rtol=1e-10, atol=0.0, itmax=0, precond="none", precision="fp64",
                 engine="batched", tile_k=0, max_cols_in_flight=0, reliable_interval=48, graph_iters=8,
                 timing=0, _handle=None):
        if precond not in _PRECOND:
            raise EqdError(_lib.EQD_ERR_ARG, f"precond must be :none or :jacobi, got {precond}")
        self.op = op
        self.precond = precond
        o = _lib.default_opts()
        o.rtol, o.atol, o.itmax = float(rtol), float(atol), int(itmax)
        o.precond, o.precision, o.engine = _PRECOND[precond], _PREC[precision], _ENGINE[engine]
        o.tile_k, o.max_cols_in_flight = int(tile_k), int(max_cols_in_flight)
        o.reliable_interval, o.graph_iters = int(reliable_interval), int(graph_iters)
        o.timing = int(timing)
        self.opts = o
        if _handle is None:
            h = C.c_void_p()
            check(lib().eqd_solver_create(op.handle, C.byref(o), C.byref(h)))
            self.handle = h
        else:
            self.handle = _handle

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and h.value:
            lib().eqd_solver_destroy(h)
            self.handle = None

    def kernel_timing(self, kernel: int):
        """(total ms, launches, active-tile launches) of one kernel in the last timed build."""
        ms = C.c_double(); la = C.c_int64(); at = C.c_int64()
        check(lib().eqd_solver_kernel_timing(self.handle, kernel, C.byref(ms), C.byref(la), C.byref(at)))
        return ms.value, la.value, at.value

    def last_timing(self):
        ms = C.c_double(); la = C.c_int64(); ti = C.c_int64()
        check(lib().eqd_solver_last_timing(self.handle, C.byref(ms), C.byref(la), C.byref(ti)))
        return ms.value, la.value, ti.value


def split_node_solve(s: CGSolver, rhs) -> np.ndarray:
    """u with A u = rhs by CG from x0 = 0 (src/ElasticitySplitNode.jl:378-396); a fresh array."""
    rhs = np.ascontiguousarray(rhs, np.float64)
    if rhs.shape != (s.op.ntot,):
        raise EqdError(_lib.EQD_ERR_ARG, f"rhs must have length {s.op.ntot}")
    u = np.zeros(s.op.ntot)
    check(lib().eqd_split_node_solve(s.handle, rhs, u))
    return u


def duplicate(s: CGSolver) -> CGSolver:
    """An independent solver sharing the operator (src/ElasticitySplitNode.jl:408-418)."""
    h = C.c_void_p()
    check(lib().eqd_solver_duplicate(s.handle, C.byref(h)))
    d = CGSolver.__new__(CGSolver)
    d.op, d.precond, d.opts, d.handle = s.op, s.precond, s.opts, h
    return d


def merge_stats(into: CGSolver, frm: CGSolver) -> CGSolver:
    """merge_stats!(into, from) (src/ElasticitySplitNode.jl:426-433)."""
    check(lib().eqd_solver_merge_stats(into.handle, frm.handle))
    return into


def solver_report(s: CGSolver) -> dict:
    """solver_report(solver) (src/ElasticitySplitNode.jl:442-447)."""
    r = _lib.Report()
    check(lib().eqd_solver_report(s.handle, C.byref(r)))
    return dict(kind="cg", precond=s.precond, solves=r.solves, iterations=r.iterations,
                mean_iterations=r.mean_iterations, max_iterations=r.max_iterations, unconverged=r.unconverged,
                tile_k=r.tile_k)


class FaultElasticity:
    """FaultElasticity(g-, g+, lambda, mu, set; l_f, solver_kwargs...) (src/FaultResponse.jl:96-146).
    `rs` is its CGSolver; Omega_f nodes are ordered x2-fastest."""

    def __init__(self, grids: GridPair, lam: float, mu: float, *, l_f: float, **solver_kwargs):
        self.op = _Operator(grids, lam, mu, l_f)
        self.rs = CGSolver(self.op, **solver_kwargs)
        self.Ntot = self.op.ntot
        x2 = np.zeros(self.op.n2f); x3 = np.zeros(self.op.n3f)
        check(lib().eqd_fault_grid_axes(self.op.handle, x2, x3))
        self.x2f, self.x3f = x2, x3

    # operator-level helpers used by the parity tests
    def apply_A(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.Ntot)
        check(lib().eqd_apply_A(self.op.handle, x, y))
        return y

    def build_rhs(self, s2, s3) -> np.ndarray:
        rhs = np.zeros(self.Ntot)
        check(lib().eqd_build_rhs(self.op.handle, np.ascontiguousarray(s2, np.float64),
                                  np.ascontiguousarray(s3, np.float64), rhs))
        return rhs


def jacobi_preconditioner(fe) -> np.ndarray:
    """jacobi_preconditioner(A) (src/ElasticitySplitNode.jl:343-357): the diagonal of
    M^{-1} = diag(A)^{-1} over the full space, A's zero diagonal entries (the far-field rows)
    floored to 1. `fe` is a FaultElasticity, a CGSolver or an operator; computed on the host
    from the operator tables."""
    op = fe.op if hasattr(fe, "op") else fe
    n = int(op.ntot)
    minv = np.zeros(n)
    check(lib().eqd_jacobi_preconditioner(C.byref(op.desc), minv, n))
    return minv


def frictional_node_count(fe: FaultElasticity) -> int:
    return fe.op.nf


def fault_grid_axes(fe: FaultElasticity):
    return fe.x2f, fe.x3f


def elastic_solver_report(fe: FaultElasticity) -> dict:
    return solver_report(fe.rs)


def shear_traction(fe: FaultElasticity, s2, s3, solver: CGSolver | None = None):
    """(dtau2, dtau3) on Omega_f from one solve (src/FaultResponse.jl:194-211)."""
    s = solver or fe.rs
    nf = fe.op.nf
    s2 = np.ascontiguousarray(s2, np.float64); s3 = np.ascontiguousarray(s3, np.float64)
    if s2.shape != (nf,) or s3.shape != (nf,):
        raise EqdError(_lib.EQD_ERR_ARG, f"slip vectors must have length nf = {nf}")
    d2 = np.zeros(nf); d3 = np.zeros(nf)
    check(lib().eqd_shear_traction(s.handle, s2, s3, d2, d3))
    return d2, d3


def fault_stiffness(fe: FaultElasticity, *, cols=None, verbose=False, solver: CGSolver | None = None):
    """Dense K (2nf x 2nf, column-major) from 2nf unit-slip solves (src/FaultResponse.jl:229-287).
    `cols` (0-based) restricts the build to those columns and returns K[:, cols]."""
    s = solver or fe.rs
    nf2 = 2 * fe.op.nf
    if cols is None:
        K = np.zeros((nf2, nf2), order="F")
        check(lib().eqd_fault_stiffness(s.handle, _lib.STIFFNESS_EXACT, None, 0, K.ctypes.data, nf2))
        return K
    cols = np.ascontiguousarray(cols, np.int32)
    K = np.zeros((nf2, len(cols)), order="F")
    check(lib().eqd_fault_stiffness(s.handle, _lib.STIFFNESS_EXACT, cols.ctypes.data, len(cols),
                                    K.ctypes.data, nf2))
    return K


def fault_stiffness_toeplitz(fe: FaultElasticity, *, verbose=False, solver: CGSolver | None = None):
    """K from 5 sources / 10 solves, priority fill (src/FaultResponse.jl:339-393)."""
    s = solver or fe.rs
    nf2 = 2 * fe.op.nf
    K = np.zeros((nf2, nf2), order="F")
    check(lib().eqd_fault_stiffness(s.handle, _lib.STIFFNESS_TOEPLITZ, None, 0, K.ctypes.data, nf2))
    return K
