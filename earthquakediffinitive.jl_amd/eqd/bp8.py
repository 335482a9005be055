"""Host-side mirror of EarthquakeDiffinitive.BP8 (src/BP8.jl) over libeqd's device time loop.

  benchmark_parameters   BP8Params / Table 1 (src/BP8.jl:50-100)
  build_model            src/BP8.jl:184-258, with K built and kept on the device
  state_length, initial_state, evaluate (evaluate!), rhs (rhs!), run_bp8,
  effective_stress_report                                  (src/BP8.jl:262-470)

The right-hand side, the Newton solves, the stiffness product and the whole Tsit5 loop
run in libeqd.so; this module only passes host arrays across the C ABI.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import FaultElasticity, bp8_grid_pair, frictional_node_count, fault_grid_axes
from . import _lib
from ._lib import EqdError, check, lib

__all__ = ["benchmark_parameters", "BP8Model", "build_model", "model_from_stiffness", "state_length",
           "initial_state", "evaluate", "rhs", "run_bp8", "Solution", "effective_stress_report"]

_INJ = {"gaussian": _lib.INJECTION_GAUSSIAN, "peaceman": _lib.INJECTION_PEACEMAN}
_STIFF = {"exact": _lib.STIFFNESS_EXACT, "toeplitz": _lib.STIFFNESS_TOEPLITZ}


def benchmark_parameters(**overrides) -> dict:
    """Table 1 of BP8 (src/BP8.jl:50-100), ASCII field names (mu, nu, sigma0, tau_init, ...)."""
    p = _lib.BP8Params()
    lib().eqd_bp8_default_params(C.byref(p))
    d = {n: getattr(p, n) for n in _lib.BP8_PARAM_NAMES}
    for k, v in overrides.items():
        if k not in d:
            raise EqdError(_lib.EQD_ERR_ARG, f"unknown BP8 parameter {k}")
        d[k] = float(v)
    return d


def lame_lambda(par: dict) -> float:
    return 2 * par["mu"] * par["nu"] / (1 - 2 * par["nu"])


def _cparams(par: dict) -> _lib.BP8Params:
    return _lib.BP8Params(*[par[n] for n in _lib.BP8_PARAM_NAMES])


class BP8Model:
    """The coupled model: K resident on the device plus A_p, source, weights, the locked ring,
    tau0 and the Peaceman constants (src/BP8.jl:123-142)."""

    def __init__(self, handle, par: dict, n2f: int, injection: str, grid_info: dict):
        self.handle = handle
        self.par = par
        self.injection = injection
        self.n2f = n2f
        self.nf = n2f * n2f
        n = C.c_int()
        nf = C.c_int()
        check(lib().eqd_bp8_sizes(handle, C.byref(nf), C.byref(n)))
        self.state_len = n.value
        h = 2 * par["l_f"] / (n2f - 1)
        self.x2 = -par["l_f"] + h * np.arange(n2f)
        self.x3 = self.x2.copy()
        self.grid_info = grid_info
        act = np.ones((n2f, n2f), bool)
        act[0, :] = act[-1, :] = act[:, 0] = act[:, -1] = False
        self.active = act.ravel(order="F")
        vn = math.hypot(par["V_init"], par["V_zero"])
        self.tau0 = np.array([par["tau_init"] * par["V_init"] / vn, par["tau_init"] * par["V_zero"] / vn])

    def __del__(self):
        if getattr(self, "handle", None):
            lib().eqd_bp8_destroy(self.handle)
            self.handle = None

    @property
    def K(self) -> np.ndarray:
        n2 = 2 * self.nf
        K = np.zeros((n2, n2), order="F")
        flat = np.zeros(n2 * n2)
        check(lib().eqd_bp8_stiffness(self.handle, flat, n2))
        K[:] = flat.reshape(n2, n2, order="F")
        return K

    def operators(self):
        """(A_p dense nf x nf, source, weights)."""
        A = np.zeros((self.nf, self.nf))
        src = np.zeros(self.nf)
        w = np.zeros(self.nf)
        check(lib().eqd_bp8_operators(self.handle, A.ctypes.data, src.ctypes.data, w.ctypes.data))
        return A, src, w

    def set_ks_method(self, method: str):
        """K s of a stiffness=:toeplitz model: "fft" (the 2D FFT convolution, the default beyond
        48 x 48 fault nodes) or "direct" (the separation-kernel sum)."""
        if method not in ("direct", "fft"):
            raise EqdError(_lib.EQD_ERR_ARG, f"K s method must be 'direct' or 'fft', got {method}")
        check(lib().eqd_bp8_set_ks_method(self.handle, _lib.KS_FFT if method == "fft" else _lib.KS_DIRECT))

    def ks(self, s) -> np.ndarray:
        """K s ([tau2; tau3], 2nf) for a slip vector [s2; s3] with the model's K s method."""
        s = np.ascontiguousarray(s, dtype=np.float64)
        if s.shape != (2 * self.nf,):
            raise EqdError(_lib.EQD_ERR_ARG, f"slip must have {2 * self.nf} entries")
        out = np.zeros(2 * self.nf)
        check(lib().eqd_bp8_ks(self.handle, s, out))
        return out

    def ks_timing(self, reps: int = 50) -> float:
        """Average device ms of one K s with the model's current method."""
        ms = C.c_double()
        check(lib().eqd_bp8_ks_timing(self.handle, int(reps), C.byref(ms)))
        return ms.value

    @property
    def Ap(self):
        return self.operators()[0]

    @property
    def source(self):
        return self.operators()[1]

    @property
    def weights(self):
        return self.operators()[2]


def _check_par(par: dict, dz: float, L_fault: float, injection: str):
    if injection not in _INJ:
        raise EqdError(_lib.EQD_ERR_ARG, f"injection must be :gaussian or :peaceman, got {injection}")
    r = par["l_f"] / dz
    if abs(r - round(r)) > 1e-9:
        raise EqdError(_lib.EQD_ERR_GRID, f"dz={dz} must divide l_f={par['l_f']} so the grid has nodes on +-l_f")
    if L_fault < par["l_f"]:
        raise EqdError(_lib.EQD_ERR_GRID, f"L_fault={L_fault} must be at least l_f={par['l_f']}")


def build_model(*, par: dict | None = None, dz: float | None = None, L_fault: float | None = None,
                L_normal: float | None = None, injection: str = "gaussian", order: int = 4, verbose: bool = False,
                stiffness: str = "toeplitz", **solver_kwargs) -> BP8Model:
    """build_model(; par, dz, L_fault, L_normal, injection, order, stiffness, solver_kwargs...)
    (src/BP8.jl:184-258). The K build runs on the device and K stays there."""
    par = dict(par or benchmark_parameters())
    dz = par["dz"] if dz is None else float(dz)
    L_fault = 3 * par["l_f"] if L_fault is None else float(L_fault)
    L_normal = 2 * par["l_f"] if L_normal is None else float(L_normal)
    par["dz"] = dz
    _check_par(par, dz, L_fault, injection)
    if order != 4:
        raise EqdError(_lib.EQD_ERR_ARG, "only SBP order 4 is provided")
    g = bp8_grid_pair(dz, L_fault, L_normal, order)
    if stiffness not in _STIFF:
        raise EqdError(_lib.EQD_ERR_ARG, f"stiffness must be :exact or :toeplitz, got {stiffness}")
    fe = FaultElasticity(g, lame_lambda(par), par["mu"], l_f=par["l_f"], **solver_kwargs)
    x2, _ = fault_grid_axes(fe)
    n2f = len(x2)
    if n2f < 2 * order + 1:
        raise EqdError(_lib.EQD_ERR_GRID, f"Omega_f spans only {n2f} nodes at dz={dz}; the pore-pressure operator "
                                          f"needs at least {2 * order + 1}.")
    h = C.c_void_p()
    check(lib().eqd_bp8_create_from_solver(fe.rs.handle, _STIFF[stiffness], C.byref(_cparams(par)), _INJ[injection],
                                           C.byref(h)))
    info = dict(dz=dz, L_fault=L_fault, L_normal=L_normal, n1=g.n1, n23=g.n23, order=order,
                elastic_dofs=6 * g.n1 * g.n23 ** 2, n2f=n2f, n3f=n2f)
    m = BP8Model(h.value, par, n2f, injection, info)
    m.fe = fe
    return m


def model_from_stiffness(K: np.ndarray, n2f: int, *, par: dict | None = None, injection: str = "gaussian") -> BP8Model:
    """A BP8 model around a given K (2nf x 2nf); the rest as build_model sets it up."""
    par = dict(par or benchmark_parameters())
    if injection not in _INJ:
        raise EqdError(_lib.EQD_ERR_ARG, f"injection must be :gaussian or :peaceman, got {injection}")
    K = np.asarray(K, dtype=np.float64)
    n2 = 2 * n2f * n2f
    if K.shape != (n2, n2):
        raise EqdError(_lib.EQD_ERR_ARG, f"K must be {n2} x {n2}")
    flat = np.ascontiguousarray(K.ravel(order="F"))
    h = C.c_void_p()
    check(lib().eqd_bp8_create(C.byref(_cparams(par)), flat, n2, n2f, _INJ[injection], C.byref(h)))
    return BP8Model(h.value, par, n2f, injection, dict(n2f=n2f, n3f=n2f))


def state_length(m: BP8Model) -> int:
    return m.state_len


def initial_state(m: BP8Model) -> np.ndarray:
    u = np.zeros(m.state_len)
    check(lib().eqd_bp8_initial_state(m.handle, u))
    return u


def evaluate(m: BP8Model, u, t: float = 0.0) -> dict:
    """evaluate!(m, u, t) (src/BP8.jl:355-396): slip rates and shear stresses on Omega_f."""
    u = np.ascontiguousarray(u, np.float64)
    if u.shape != (m.state_len,):
        raise EqdError(_lib.EQD_ERR_ARG, f"state must have length {m.state_len}")
    out = {k: np.zeros(m.nf) for k in ("V2", "V3", "Vmag", "tau2", "tau3")}
    check(lib().eqd_bp8_evaluate(m.handle, u, float(t), *[out[k].ctypes.data for k in ("V2", "V3", "Vmag", "tau2",
                                                                                         "tau3")]))
    return out


def rhs(du, u, m: BP8Model, t: float):
    """rhs!(du, u, m, t) (src/BP8.jl:403-433); fills du in place and returns it."""
    u = np.ascontiguousarray(u, np.float64)
    if u.shape != (m.state_len,) or du.shape != (m.state_len,):
        raise EqdError(_lib.EQD_ERR_ARG, f"state must have length {m.state_len}")
    out = np.zeros(m.state_len)
    check(lib().eqd_bp8_rhs(m.handle, u, float(t), out))
    du[:] = out
    return du


def reset_cache(m: BP8Model):
    check(lib().eqd_bp8_reset_cache(m.handle))


@dataclass
class Solution:
    t: np.ndarray
    u: np.ndarray          # nsaved x state_length
    naccept: int
    nreject: int
    nrhs: int
    retcode: str
    ms: float


def run_bp8(m: BP8Model, *, tspan=None, reltol: float = 1e-8, saveat: float = 3600.0, verbose: bool = False,
            maxiters: int = 100000, alg: str = "tsit5") -> Solution:
    """run_bp8(m; tspan, reltol, saveat, alg) (src/BP8.jl:442-470) on the device. alg = "tsit5" (the
    reference's default) or "imex": the implicit-local SDIRK for the stiff Peaceman-well variant
    (eqd_bp8_integrate_imex)."""
    if alg not in ("tsit5", "imex"):
        raise EqdError(_lib.EQD_ERR_ARG, f"alg must be tsit5 or imex, got {alg}")
    t0, tf = (0.0, m.par["t_f"]) if tspan is None else (float(tspan[0]), float(tspan[1]))
    nmax = int(math.floor((tf - t0) / saveat + 1e-9)) + 2
    ts = np.zeros(nmax)
    us = np.zeros(nmax * m.state_len)
    st = _lib.BP8Stats()
    fn = lib().eqd_bp8_integrate if alg == "tsit5" else lib().eqd_bp8_integrate_imex
    check(fn(m.handle, t0, tf, float(saveat), float(reltol), int(maxiters), nmax, ts.ctypes.data, us.ctypes.data,
             C.byref(st)))
    k = st.nsaved
    return Solution(ts[:k].copy(), us[:k * m.state_len].reshape(k, m.state_len).copy(), st.naccept, st.nreject,
                    st.nrhs, "Success" if st.retcode == 0 else "MaxIters", st.ms)


def saved_vmax(m: BP8Model, n: int) -> np.ndarray:
    """max |V| at the first n saved states of the last run_bp8, via evaluate! in save order."""
    v = np.zeros(n)
    check(lib().eqd_bp8_saved_vmax(m.handle, v, n))
    return v


def effective_stress_report(m: BP8Model) -> dict:
    """effective_stress_report(m) (src/BP8.jl:287-296)."""
    lo = C.c_double()
    hits = C.c_int64()
    mask = np.zeros(m.nf, np.uint8)
    check(lib().eqd_bp8_stress_report(m.handle, C.byref(lo), C.byref(hits), mask.ctypes.data))
    idx = np.nonzero(mask)[0]
    n2 = m.n2f
    radius = 0.0 if len(idx) == 0 else float(max(math.hypot(m.x2[i % n2], m.x3[i // n2]) for i in idx))
    return dict(sigma_lowest=lo.value, floor_hits=hits.value, floor=m.par["sigma_min"], bound=hits.value > 0,
                nodes=len(idx), fraction=len(idx) / m.nf, radius=radius)


def process_zone(par: dict, sigma_bar: float | None = None) -> float:
    """process_zone(par, sigma_bar) = mu D_RS / (b sigma_bar) (src/BP8.jl:297-305): the rate-and-state
    nucleation length, at the initial sigma0 by default."""
    sb = par["sigma0"] if sigma_bar is None else sigma_bar
    return par["mu"] * par["D_RS"] / (par["b"] * sb)


def resolution_report(m: BP8Model) -> dict:
    """resolution_report(m) (src/BP8.jl:307-326): cells of dz per process zone L_b (at sigma0, and
    at the lowest effective stress seen, floored at sigma_min) and per Gaussian source width;
    converged when L_b / dz >= 3."""
    par = m.par
    dz = m.grid_info.get("dz", par["dz"])
    lb0 = process_zone(par)
    lo = effective_stress_report(m)["sigma_lowest"]
    sb_lo = max(lo, par["sigma_min"]) if math.isfinite(lo) else par["sigma0"]
    return dict(dz=dz, L_b=lb0, cells_per_Lb=lb0 / dz, L_b_at_lowest_sigma=process_zone(par, sb_lo),
                cells_per_L_gauss=par["L_gauss"] / dz, converged=lb0 / dz >= 3)


from .outputs import saved_outputs, station_locations, write_outputs  # noqa: E402  (section 4 output files)
