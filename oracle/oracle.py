"""ctypes front end of the CPU FP64 oracle (libeqd_oracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module as the checker (or the timed CPU baseline). The
product path never imports it.

The functions restate the reference's CPU algorithm; see eqd_oracle.h for the
file:line map into the reference (src/ElasticitySplitNode.jl, src/FaultResponse.jl).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OP_A, OP_HPDSAT, OP_P, OP_T2, OP_T3, OP_DSAT, OP_H = range(7)


class _Params(C.Structure):
    _fields_ = [("n1", C.c_int), ("n23", C.c_int), ("L_normal", C.c_double), ("L_fault", C.c_double),
                ("lam", C.c_double), ("mu", C.c_double), ("l_f", C.c_double), ("precond", C.c_int),
                ("rtol", C.c_double), ("atol", C.c_double), ("itmax", C.c_int)]


class _Stats(C.Structure):
    _fields_ = [("solves", C.c_int64), ("iterations", C.c_int64), ("max_iterations", C.c_int64),
                ("unconverged", C.c_int64)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libeqd_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.eqo_create.restype = C.c_void_p
        L.eqo_create.argtypes = [C.POINTER(_Params), C.c_char_p, C.c_size_t]
        L.eqo_destroy.argtypes = [C.c_void_p]
        L.eqo_info.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.int64)]
        L.eqo_apply.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.eqo_export.restype = C.c_int64
        L.eqo_export.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.eqo_build_chi.argtypes = [C.c_void_p, dp, dp, dp]
        L.eqo_cg.argtypes = [C.c_void_p, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.eqo_shear_traction.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.eqo_fault_stiffness.argtypes = [C.c_void_p, C.c_int, dp]
        L.eqo_fault_stiffness_cols.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.int32), C.c_int, C.c_int, dp]
        L.eqo_fault_stiffness_toeplitz.argtypes = [C.c_void_p, dp]
        L.eqo_report.argtypes = [C.c_void_p, C.POINTER(_Stats)]
        L.eqo_reset_stats.argtypes = [C.c_void_p]
        L.eqo_sbp1d.argtypes = [C.c_int, C.c_double, dp, dp, dp, dp, dp]
        _LIB = L
    return _LIB


def sbp1d(n: int, h: float):
    """(H, D1, D2, dl, dr) of the 1D order-4 operators on n points with spacing h."""
    H = np.zeros(n); D1 = np.zeros((n, n)); D2 = np.zeros((n, n)); dl = np.zeros(n); dr = np.zeros(n)
    if lib().eqo_sbp1d(n, h, H, D1, D2, dl, dr) != 0:
        raise ValueError("n must be >= 9")
    return H, D1, D2, dl, dr


@dataclass
class Geometry:
    """Two blocks [-L_normal,0]x[-L_fault,L_fault]^2 and [0,L_normal]x[...]^2 (src/BP8.jl:205-206)."""
    n1: int
    n23: int
    L_normal: float
    L_fault: float
    lam: float
    mu: float
    l_f: float


class OracleSystem:
    """FaultElasticity + CGSolver restated on the CPU (src/FaultResponse.jl:96-146)."""

    def __init__(self, g: Geometry, rtol=1e-10, atol=0.0, itmax=0, precond="none"):
        pc = {"none": 0, "jacobi": 1}.get(precond, 2)
        p = _Params(g.n1, g.n23, g.L_normal, g.L_fault, g.lam, g.mu, g.l_f, pc, rtol, atol, itmax)
        err = C.create_string_buffer(512)
        self._h = lib().eqo_create(C.byref(p), err, 512)
        if not self._h:
            raise ValueError(err.value.decode())
        info = np.zeros(10, np.int64)
        lib().eqo_info(self._h, info)
        (self.nf, self.n2f, self.n3f, self.ntot, self.nnz_A, self.nnz_HPDSAT, self.nnz_P, self.N,
         self.i2lo, self.i3lo) = (int(v) for v in info)
        self.geometry = g

    def __del__(self):
        if getattr(self, "_h", None):
            lib().eqo_destroy(self._h)
            self._h = None

    def apply(self, op: int, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        rows = self.ntot if op not in (OP_T2, OP_T3) else self.geometry.n23 ** 2
        y = np.zeros(rows)
        lib().eqo_apply(self._h, op, x, y)
        return y

    def export(self, op: int):
        """(rowptr, col, val) of an assembled operator, for scipy.sparse.csr_matrix."""
        nnz = lib().eqo_export(self._h, op, None, None, None)
        nrows = self.ntot if op not in (OP_T2, OP_T3) else self.geometry.n23 ** 2
        rp = np.zeros(nrows + 1, np.int64); col = np.zeros(nnz, np.int64); val = np.zeros(nnz)
        lib().eqo_export(self._h, op, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
        return rp, col, val

    def build_chi(self, s2, s3):
        chi = np.zeros(self.ntot)
        lib().eqo_build_chi(self._h, np.ascontiguousarray(s2, np.float64), np.ascontiguousarray(s3, np.float64), chi)
        return chi

    def cg(self, b):
        x = np.zeros(self.ntot); it = C.c_int(); ok = C.c_int()
        lib().eqo_cg(self._h, np.ascontiguousarray(b, np.float64), x, C.byref(it), C.byref(ok))
        return x, it.value, bool(ok.value)

    def shear_traction(self, s2, s3):
        d2 = np.zeros(self.nf); d3 = np.zeros(self.nf)
        lib().eqo_shear_traction(self._h, np.ascontiguousarray(s2, np.float64),
                                 np.ascontiguousarray(s3, np.float64), d2, d3)
        return d2, d3

    def fault_stiffness(self, nthreads=1):
        K = np.zeros((2 * self.nf, 2 * self.nf), order="F")
        kt = np.zeros(K.size)
        lib().eqo_fault_stiffness(self._h, nthreads, kt)
        return kt.reshape((2 * self.nf, 2 * self.nf), order="F")

    def fault_stiffness_cols(self, cols, nthreads=1):
        cols = np.ascontiguousarray(cols, np.int32)
        out = np.zeros(len(cols) * 2 * self.nf)
        rc = lib().eqo_fault_stiffness_cols(self._h, cols, len(cols), nthreads, out)
        if rc != 0:
            raise IndexError("column index out of range")
        return out.reshape((2 * self.nf, len(cols)), order="F")

    def fault_stiffness_toeplitz(self):
        kt = np.zeros((2 * self.nf) ** 2)
        rc = lib().eqo_fault_stiffness_toeplitz(self._h, kt)
        if rc != 0:
            raise ValueError("the Toeplitz expansion assumes Omega_f forms a full n2 x n3 grid")
        return kt.reshape((2 * self.nf, 2 * self.nf), order="F")

    def report(self):
        s = _Stats()
        lib().eqo_report(self._h, C.byref(s))
        mean = s.iterations / s.solves if s.solves else 0.0
        return dict(solves=s.solves, iterations=s.iterations, mean_iterations=mean,
                    max_iterations=s.max_iterations, unconverged=s.unconverged)

    def reset_stats(self):
        lib().eqo_reset_stats(self._h)


def bp8_geometry(dz: float, L_fault: float, L_normal: float, l_f: float = 400.0,
                 mu: float = 32.04e9, nu: float = 0.25) -> Geometry:
    """Grid sizes and Lame constants as build_model sets them (src/BP8.jl:92,196-206)."""
    n1 = int(round(L_normal / dz)) + 1
    n23 = int(round(2 * L_fault / dz)) + 1
    lam = 2 * mu * nu / (1 - 2 * nu)
    if n1 < 9 or n23 < 9:
        raise ValueError(f"n1={n1}, n23={n23}: SBP order 4 needs at least 9 points")
    return Geometry(n1, n23, L_normal, L_fault, lam, mu, l_f)


# ---------------------------------------------------------------------------
# BP8 time loop (bp8_oracle.h): rate-state Newton, A_p, RHS and the Tsit5 driver.

BP8_PARAM_NAMES = ["mu", "nu", "rho", "c_s", "a", "b", "D_RS", "V_star", "f_star", "sigma0", "tau_init",
                   "V_init", "V_zero", "Q0", "L_gauss", "L_fwid", "beta", "phi", "k", "viscosity", "alpha",
                   "S_well", "r_well", "l_f", "dz", "t_off", "t_f", "sigma_min"]


class _BP8Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in BP8_PARAM_NAMES]


def _bp8_lib():
    L = lib()
    if not getattr(L, "_bp8_ready", False):
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.eqo_bp8_default_params.argtypes = [C.POINTER(_BP8Params)]
        L.eqo_solve_slip_rate.restype = C.c_int
        L.eqo_solve_slip_rate.argtypes = [C.c_double] * 10 + [C.c_double, C.c_int, C.c_double,
                                                               C.POINTER(C.c_double)]
        L.eqo_bp8_create.restype = C.c_void_p
        L.eqo_bp8_create.argtypes = [C.POINTER(_BP8Params), dp, C.c_int, C.c_int, C.c_char_p, C.c_size_t]
        L.eqo_bp8_destroy.argtypes = [C.c_void_p]
        L.eqo_bp8_state_length.argtypes = [C.c_void_p]
        L.eqo_bp8_initial_state.argtypes = [C.c_void_p, dp]
        L.eqo_bp8_evaluate.argtypes = [C.c_void_p, dp, C.c_double, dp, dp, dp, dp, dp]
        L.eqo_bp8_rhs.argtypes = [C.c_void_p, dp, C.c_double, dp]
        L.eqo_bp8_Ap.argtypes = [C.c_void_p, dp]
        L.eqo_bp8_source.argtypes = [C.c_void_p, dp, dp]
        L.eqo_bp8_reset_cache.argtypes = [C.c_void_p]
        i64p = C.POINTER(C.c_int64)
        L.eqo_bp8_run.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, dp,
                                  C.POINTER(C.c_int), i64p, i64p, i64p]
        L._bp8_ready = True
    return L


def bp8_default_params() -> dict:
    """BP8Params defaults (Table 1, src/BP8.jl:50-87) as a dict."""
    p = _BP8Params()
    _bp8_lib().eqo_bp8_default_params(C.byref(p))
    return {n: getattr(p, n) for n in BP8_PARAM_NAMES}


def solve_slip_rate(T, theta, sigma, eta, a, b, Dc, V_star, f_star, V0=1e-12, tol=1e-12, maxiter=50,
                    maxstep=5.0):
    """Newton in ln V (src/RateStateFriction.jl:77-96); raises RuntimeError on non-convergence."""
    V = C.c_double()
    rc = _bp8_lib().eqo_solve_slip_rate(T, theta, sigma, eta, a, b, Dc, V_star, f_star, V0, tol, maxiter,
                                        maxstep, C.byref(V))
    if rc:
        raise RuntimeError("solve_slip_rate failed to converge")
    return V.value


class OracleBP8:
    """BP8 model around a given K (build_model minus the K build, src/BP8.jl:184-258)."""

    def __init__(self, K: np.ndarray, n2f: int, params: dict | None = None, injection: str = "gaussian"):
        L = _bp8_lib()
        p = bp8_default_params()
        p.update(params or {})
        self.params = p
        cp = _BP8Params(*[p[n] for n in BP8_PARAM_NAMES])
        self.n2f = n2f
        self.nf = n2f * n2f
        Kc = np.ascontiguousarray(np.asarray(K, dtype=np.float64).ravel(order="F"))
        if Kc.size != (2 * self.nf) ** 2:
            raise ValueError("K must be 2nf x 2nf")
        err = C.create_string_buffer(512)
        inj = {"gaussian": 0, "peaceman": 1}.get(injection, -1)
        self._h = L.eqo_bp8_create(C.byref(cp), Kc, n2f, inj, err, 512)
        if not self._h:
            raise RuntimeError(err.value.decode())
        self.n = L.eqo_bp8_state_length(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            _bp8_lib().eqo_bp8_destroy(self._h)
            self._h = None

    def initial_state(self):
        u = np.zeros(self.n)
        _bp8_lib().eqo_bp8_initial_state(self._h, u)
        return u

    def evaluate(self, u, t=0.0):
        nf = self.nf
        out = [np.zeros(nf) for _ in range(5)]
        if _bp8_lib().eqo_bp8_evaluate(self._h, np.ascontiguousarray(u, dtype=np.float64), t, *out):
            raise RuntimeError("solve_slip_rate failed to converge")
        return dict(zip(["V2", "V3", "Vmag", "tau2", "tau3"], out))

    def rhs(self, u, t):
        du = np.zeros(self.n)
        if _bp8_lib().eqo_bp8_rhs(self._h, np.ascontiguousarray(u, dtype=np.float64), t, du):
            raise RuntimeError("solve_slip_rate failed to converge")
        return du

    def Ap(self):
        A = np.zeros(self.nf * self.nf)
        _bp8_lib().eqo_bp8_Ap(self._h, A)
        return A.reshape(self.nf, self.nf)

    def source(self):
        s, w = np.zeros(self.nf), np.zeros(self.nf)
        _bp8_lib().eqo_bp8_source(self._h, s, w)
        return s, w

    def reset_cache(self):
        _bp8_lib().eqo_bp8_reset_cache(self._h)

    def run(self, t0, tf, saveat, reltol=1e-8):
        nmax = int(np.floor((tf - t0) / saveat + 1e-9)) + 2
        ts, us, vmax = np.zeros(nmax), np.zeros(nmax * self.n), np.zeros(nmax)
        ns, na, nr, nrhs = C.c_int(), C.c_int64(), C.c_int64(), C.c_int64()
        rc = _bp8_lib().eqo_bp8_run(self._h, t0, tf, saveat, reltol, nmax, ts, us, vmax, C.byref(ns), C.byref(na),
                                    C.byref(nr), C.byref(nrhs))
        if rc:
            raise RuntimeError("solve_slip_rate failed to converge")
        k = ns.value
        return dict(t=ts[:k], u=us[:k * self.n].reshape(k, self.n), vmax=vmax[:k], naccept=na.value,
                    nreject=nr.value, nrhs=nrhs.value)
