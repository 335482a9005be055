"""Trajectory ensembles of the BP8 time loop on one resident K.

Each member is run_bp8 (src/BP8.jl:442-470) for build_model with its own injection rate Q0,
rate-state a and background normal stress sigma0 (and optionally its own reltol), sharing
the model's K, A_p, source and grid. Members step in lockstep stages, each with its own
Tsit5 controller and saves; the stage's K S over all members is one FP64 MFMA GEMM.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import EqdError, check, lib
from .bp8 import BP8Model, Solution

__all__ = ["perturbed_members", "Ensemble"]


def perturbed_members(par: dict, members: int, *, seed: int = 0, dQ0: float = 0.2, dab: float = 0.1,
                      dsigma0: float = 0.05) -> dict:
    """Seeded member parameters: Q0 (1 + U(-dQ0, dQ0)), a - b (1 + U(-dab, dab)) with b fixed,
    sigma0 (1 + U(-dsigma0, dsigma0)). Member 0 keeps the base values."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1.0, 1.0, size=(3, members))
    u[:, 0] = 0.0
    Q0 = par["Q0"] * (1.0 + dQ0 * u[0])
    a = par["b"] + (par["a"] - par["b"]) * (1.0 + dab * u[1])
    sigma0 = par["sigma0"] * (1.0 + dsigma0 * u[2])
    return dict(Q0=Q0, a=a, sigma0=sigma0)


class Ensemble:
    """`members` trajectories of one BP8Model (the model must stay alive)."""

    def __init__(self, model: BP8Model, members: int, *, Q0=None, a=None, sigma0=None):
        self.model = model
        self.members = int(members)
        arrs = []
        for v in (Q0, a, sigma0):
            if v is None:
                arrs.append(None)
            else:
                v = np.ascontiguousarray(v, dtype=np.float64)
                if v.shape != (self.members,):
                    raise EqdError(_lib.EQD_ERR_ARG, "member parameter arrays need one value per member")
                arrs.append(v)
        self._keep = arrs
        h = C.c_void_p()
        ptr = [None if v is None else v.ctypes.data for v in arrs]
        check(lib().eqd_bp8_ensemble_create(model.handle, self.members, ptr[0], ptr[1], ptr[2], C.byref(h)))
        self.handle = h
        self.stats = None
        self.loop_ms = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            lib().eqd_bp8_ensemble_destroy(h)
            self.handle = None

    def run(self, *, tspan=None, reltol: float = 1e-8, member_reltol=None, saveat: float = 3600.0,
            maxiters: int = 100000, alg: str = "tsit5"):
        """Integrates every member; returns the per-member BP8Stats list. alg = "tsit5" (run_bp8's
        integrator, src/BP8.jl:442-470) or "imex" (the implicit-local integrator of run_bp8's
        alg="imex", for the stiff Peaceman-well variant)."""
        if alg not in ("tsit5", "imex"):
            raise EqdError(_lib.EQD_ERR_ARG, f"alg must be tsit5 or imex, got {alg}")
        par = self.model.par
        t0, tf = (0.0, par["t_f"]) if tspan is None else (float(tspan[0]), float(tspan[1]))
        nmax = int(math.floor((tf - t0) / saveat + 1e-9)) + 2
        st = (_lib.BP8Stats * self.members)()
        ms = C.c_double()
        rt = None
        if member_reltol is not None:
            rt = np.ascontiguousarray(member_reltol, dtype=np.float64)
            if rt.shape != (self.members,):
                raise EqdError(_lib.EQD_ERR_ARG, "member_reltol needs one value per member")
        fn = lib().eqd_bp8_ensemble_integrate if alg == "tsit5" else lib().eqd_bp8_ensemble_integrate_imex
        check(fn(self.handle, t0, tf, float(saveat), float(reltol), None if rt is None else rt.ctypes.data,
                 int(maxiters), nmax, st, C.byref(ms)))
        self.stats = list(st)
        self.loop_ms = ms.value
        return self.stats

    def solution(self, member: int) -> Solution:
        """Member's saves as a Solution (t, u rows) like run_bp8 returns."""
        ns = self.model.state_len
        n = C.c_int64()
        check(lib().eqd_bp8_ensemble_saved(self.handle, int(member), None, None, C.byref(n)))
        k = n.value
        ts = np.zeros(max(k, 1))
        us = np.zeros(max(k, 1) * ns)
        check(lib().eqd_bp8_ensemble_saved(self.handle, int(member), ts.ctypes.data, us.ctypes.data, C.byref(n)))
        s = self.stats[member] if self.stats else None
        return Solution(ts[:k].copy(), us[:k * ns].reshape(k, ns).copy(), s.naccept if s else 0,
                        s.nreject if s else 0, s.nrhs if s else 0,
                        "Success" if (s is None or s.retcode == 0) else "MaxIters", s.ms if s else 0.0)

    def saved_vmax(self, n: int):
        """max |V| at the first n saves of every member (members x n) and the device ms."""
        v = np.zeros(self.members * n)
        ms = C.c_double()
        check(lib().eqd_bp8_ensemble_saved_vmax(self.handle, v.ctypes.data, int(n), C.byref(ms)))
        return v.reshape(self.members, n), ms.value

    def gemm_timing(self, reps: int = 20):
        """(ms per K S GEMM over all members, flops per GEMM)."""
        ms = C.c_double()
        fl = C.c_double()
        check(lib().eqd_bp8_ensemble_gemm_timing(self.handle, int(reps), C.byref(ms), C.byref(fl)))
        return ms.value, fl.value
