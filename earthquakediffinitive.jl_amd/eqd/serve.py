"""Serving loop over libeqd: stiffness-column and forward-solve requests on several grids,
admitted into resident 64-column tile pools as lanes free up (continuous batching).

A column's arithmetic depends only on its own lane, so a request's result is the same
whether it runs alone or among others, and the same as the batch build's columns
(src/FaultResponse.jl:229-287 for columns, :194-211 for forward solves).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import GridPair, _desc
from . import _lib
from ._lib import EqdError, check, lib

__all__ = ["Server"]

_PREC = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}


class Server:
    """Requests are queued with `submit_columns` / `submit_forward`; `step()` advances the
    device (retire, admit, one graph replay of CG iterations); `completed()` lists finished
    request ids and `result(id)` returns (tractions 2nf x ncols, iterations per column)."""

    def __init__(self, *, precision="mixed", rtol=1e-10, atol=0.0, hbm_budget_bytes=0, graph_iters=8,
                 reliable_interval=48):
        o = _lib.default_opts()
        o.rtol, o.atol = float(rtol), float(atol)
        o.precision = _PREC[precision]
        o.tile_k = 64
        o.graph_iters = int(graph_iters)
        o.reliable_interval = int(reliable_interval)
        h = C.c_void_p()
        check(lib().eqd_server_create(C.byref(o), int(hbm_budget_bytes), C.byref(h)))
        self.handle = h.value
        self.grids = []

    def __del__(self):
        if getattr(self, "handle", None):
            lib().eqd_server_destroy(self.handle)
            self.handle = None

    def add_grid(self, grids: GridPair, lam: float, mu: float, *, l_f: float, max_columns: int = 0) -> int:
        gid, slots = C.c_int(), C.c_int()
        check(lib().eqd_server_add_grid(self.handle, C.byref(_desc(grids, lam, mu, l_f)), int(max_columns),
                                        C.byref(gid), C.byref(slots)))
        info = _lib.OpInfo()
        hb = C.c_double()
        check(lib().eqd_server_grid_info(self.handle, gid.value, C.byref(info), None, C.byref(hb)))
        self.grids.append(dict(nf=info.nf, slots=slots.value, hbm_bytes=hb.value, n2f=info.n2f))
        return gid.value

    def submit_columns(self, grid: int, cols) -> int:
        cols = np.ascontiguousarray(cols, np.int32)
        rid = C.c_int64()
        check(lib().eqd_server_submit_columns(self.handle, grid, cols.ctypes.data, len(cols), C.byref(rid)))
        return rid.value

    def submit_forward(self, grid: int, s2, s3) -> int:
        nf = self.grids[grid]["nf"]
        s2 = np.ascontiguousarray(s2, np.float64)
        s3 = np.ascontiguousarray(s3, np.float64)
        if s2.shape != (nf,) or s3.shape != (nf,):
            raise EqdError(_lib.EQD_ERR_ARG, f"slip vectors must have length nf = {nf}")
        rid = C.c_int64()
        check(lib().eqd_server_submit_forward(self.handle, grid, s2, s3, C.byref(rid)))
        return rid.value

    def step(self):
        busy, queued = C.c_int64(), C.c_int64()
        check(lib().eqd_server_step(self.handle, C.byref(busy), C.byref(queued)))
        return busy.value, queued.value

    def completed(self, max_ids: int = 4096):
        ids = np.zeros(max_ids, np.int64)
        n = C.c_int()
        check(lib().eqd_server_completed(self.handle, ids.ctypes.data, max_ids, C.byref(n)))
        return ids[:n.value].tolist()

    def result(self, rid: int, grid: int, ncols: int):
        n2 = 2 * self.grids[grid]["nf"]
        out = np.zeros((n2, ncols), order="F")
        iters = np.zeros(ncols, np.int32)
        unc = C.c_int()
        check(lib().eqd_server_result(self.handle, rid, out.ctypes.data, n2, iters.ctypes.data, C.byref(unc)))
        return out, iters, unc.value

    def drain(self, timeout_s: float = 600.0):
        """Steps until nothing is queued or in flight; returns the completed ids. A step returns at
        once for a grid whose previous replay is still running, so the bound is wall-clock time,
        not a count of steps."""
        done = []
        t_end = time.monotonic() + timeout_s
        while True:
            busy, queued = self.step()
            done += self.completed()
            if busy == 0 and queued == 0:
                return done
            if time.monotonic() > t_end:
                raise EqdError(_lib.EQD_ERR_DEVICE, f"serving loop did not drain within {timeout_s} s")
