"""The BP8 benchmark output files (SEAS BP8-QD section 4) from device-evaluated saves.

Mirrors `write_outputs(m, sol, dir; modeler, profile_dt)` and its helpers
(src/BP8.jl:543-774): the nine on-fault station time series, `global.dat` and the ten
slip / stress / pressure profiles, in the same file layouts. The derived fields that
`write_outputs` recomputes with `evaluate!` at every saved time (slip rates, shear stresses,
maximum slip rate, moment rate, Darcy velocities) come from one device pass over the saves
(`eqd_bp8_saved_outputs`); only the text formatting runs on the host.
"""
from __future__ import annotations

import datetime
import math
import os

import numpy as np

from . import _lib
from ._lib import EqdError, check, lib

__all__ = ["station_locations", "nearest_node", "saved_outputs", "write_outputs", "TS_FIELDS", "TS_COLUMNS"]

TS_FIELDS = ["t", "slip_2", "slip_3", "slip_rate_2", "slip_rate_3", "shear_stress_2", "shear_stress_3",
             "pore_pressure", "darcy_vel_2", "darcy_vel_3", "state"]

TS_COLUMNS = ["Time (s)", "Slip_2 (m)", "Slip_3 (m)", "Slip_rate_2 (log10 m/s)", "Slip_rate_3 (log10 m/s)",
              "Shear_stress_2 (MPa)", "Shear_stress_3 (MPa)", "Pore_pressure (MPa)", "Darcy_velocity_2 (m/s)",
              "Darcy_velocity_3 (m/s)", "State (log10 s)"]


def station_locations():
    """The nine on-fault observation points (src/BP8.jl:553-561): (file stem, x2, x3)."""
    def fmt(v):
        return ("-" if v < 0 else "+") + f"{abs(int(round(v))):03d}"
    pts = [(0.0, 0.0), (-200.0, 0.0), (0.0, 200.0), (200.0, 0.0), (0.0, -200.0),
           (-200.0, -200.0), (-200.0, 200.0), (200.0, -200.0), (200.0, 200.0)]
    return [(f"fltst_strk{fmt(x2)}dp{fmt(x3)}", x2, x3) for x2, x3 in pts]


def nearest_node(axis, x) -> int:
    """0-based index of the first axis node closest to x (src/BP8.jl:564)."""
    return int(np.argmin(np.abs(np.asarray(axis) - x)))


def _safelog10(x: float) -> float:
    return math.log10(max(abs(x), 1e-300))


def _problem_name(m) -> str:
    return "BP8-QD-GS" if m.injection == "gaussian" else "BP8-QD-PW"


def _fmt_real(v: float) -> str:
    """A Float64 as Julia's string interpolation prints it for these header values."""
    return repr(float(v))


def _domain_line(m) -> str:
    gi = m.grid_info
    return (f"# elastic_domain=|x1| <= {_fmt_real(gi['L_normal'])} m, |x2|,|x3| <= {_fmt_real(gi['L_fault'])} m, "
            f"SBP order {gi['order']}, {gi['elastic_dofs']} elastic DOF")


# The code that writes these files (the format's `code=` / `version=` header keys).
CODE_NAME = "libeqd"
CODE_VERSION = "0.2.0"


def _date() -> str:
    return datetime.date.today().strftime("%Y/%m/%d")


def _ts_header(m, modeler, location, extra):
    return (["# This is the file header:", f"# problem=SEAS Benchmark {_problem_name(m)}",
             f"# code={CODE_NAME}", f"# version={CODE_VERSION}", f"# modeler={modeler}", f"# date={_date()}",
             f"# element_size={_fmt_real(m.grid_info['dz'])} m", f"# location={location}"]
            + list(extra) + [_domain_line(m)])


def saved_outputs(m, n: int, idx) -> tuple[np.ndarray, np.ndarray]:
    """Derived fields of the first n saves of the last run_bp8 (evaluate! in save order):
    fields[n, len(idx), 6] = V2, V3, tau2, tau3, darcy_vel_2, darcy_vel_3 at the listed Omega_f
    nodes (0-based, x2 fastest) and glob[n, 2] = max |V|, moment rate."""
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    fields = np.zeros(max(1, n * len(idx) * 6))
    glob = np.zeros(max(1, 2 * n))
    check(lib().eqd_bp8_saved_outputs(m.handle, int(n), idx.ctypes.data if len(idx) else None, len(idx),
                                      fields.ctypes.data, glob.ctypes.data))
    return fields[:n * len(idx) * 6].reshape(n, len(idx), 6), glob[:2 * n].reshape(n, 2)


def _profile_times(t_end: float, profile_dt: float):
    """collect(0.0:profile_dt:t_end), plus t_end when the range stops short (src/BP8.jl:705-707)."""
    k = int(math.floor(t_end / profile_dt + 1e-9)) if profile_dt > 0 else 0
    ts = [i * profile_dt for i in range(k + 1)] if t_end >= 0 else []
    if not ts:
        ts = [t_end]
    if ts[-1] < t_end:
        ts.append(t_end)
    return ts


def write_outputs(m, sol, dir: str, *, modeler: str = "", profile_dt: float = 3600.0) -> str:
    """write_outputs(m, sol, dir; modeler, profile_dt) (src/BP8.jl:605-642). `sol` is the Solution
    of the last run_bp8 on `m` (its saves are still on the device). Profile times must fall on
    saved times (the reference evaluates sol(t) there, which on the saveat grid is the save)."""
    missing = [k for k in ("dz", "L_normal", "L_fault", "order", "elastic_dofs") if k not in m.grid_info]
    if missing:
        raise EqdError(_lib.EQD_ERR_ARG, f"write_outputs needs the grid of a build_model model (missing {missing})")
    os.makedirs(dir, exist_ok=True)
    times = np.asarray(sol.t)
    ns = len(times)
    nf, n2f = m.nf, m.n2f
    if ns == 0:
        raise EqdError(_lib.EQD_ERR_ARG, "the solution has no saved states")
    stations = station_locations()
    st_idx = []
    for _, sx2, sx3 in stations:
        i, j = nearest_node(m.x2, sx2), nearest_node(m.x3, sx3)
        st_idx.append(i + n2f * j)
    j0, i0 = nearest_node(m.x3, 0.0), nearest_node(m.x2, 0.0)
    strike = [i + n2f * j0 for i in range(n2f)]
    depth = [i0 + n2f * j for j in range(n2f)]
    fields, glob = saved_outputs(m, ns, st_idx + strike + depth)
    dts = np.diff(times)
    step_lines = [] if len(dts) == 0 else [f"# minimum_time_step={dts.min():.3E}",
                                          f"# maximum_time_step={dts.max():.3E}", f"# num_time_steps={ns}"]
    _write_time_series(m, sol, dir, stations, st_idx, fields, step_lines, modeler)
    _write_global(m, dir, times, glob, step_lines, modeler)
    _write_profiles(m, sol, dir, profile_dt, strike, depth, fields[:, 9:, :], glob, modeler)
    return dir


def _write_time_series(m, sol, dir, stations, st_idx, fields, step_lines, modeler):
    nf = m.nf
    n2f = m.n2f
    for s, ((name, sx2, sx3), idx) in enumerate(zip(stations, st_idx)):
        i, j = idx % n2f, idx // n2f
        loc = "on fault, strike = %.4g m, depth = %.4g m" % (m.x2[i], m.x3[j])
        extra = []
        if not (abs(m.x2[i] - sx2) <= 1e-6 and abs(m.x3[j] - sx3) <= 1e-6):
            extra.append(f"# note=requested station ({_fmt_real(sx2)}, {_fmt_real(sx3)}) m is not a grid node; "
                         f"nearest node used")
        with open(os.path.join(dir, name + ".dat"), "w") as io:
            for line in _ts_header(m, modeler, loc, extra + step_lines):
                io.write(line + "\n")
            for c, d in enumerate(TS_COLUMNS, start=1):
                io.write(f"# Column #{c} = {d}\n")
            io.write("# The line below lists the names of the data fields\n")
            io.write(" ".join(TS_FIELDS) + "\n")
            io.write("# Here is the time-series data.\n")
            for n, t in enumerate(sol.t):
                u = sol.u[n]
                f = fields[n, s]
                io.write("%21.13E %14.6E %14.6E %14.6E %14.6E %14.6E %14.6E %14.6E %14.6E %14.6E %14.6E\n" % (
                    t, u[idx], u[nf + idx], _safelog10(f[0]), _safelog10(f[1]), f[2] / 1e6, f[3] / 1e6,
                    u[3 * nf + idx] / 1e6, f[4], f[5], u[2 * nf + idx] / math.log(10)))


def _write_global(m, dir, times, glob, step_lines, modeler):
    with open(os.path.join(dir, "global.dat"), "w") as io:
        for line in _ts_header(m, modeler, "frictional domain", step_lines):
            io.write(line + "\n")
        io.write("# Column #1 = Time (s)\n")
        io.write("# Column #2 = Max_slip_rate (log10 m/s)\n")
        io.write("# Column #3 = Moment_rate (N.m/s)\n")
        io.write("# The line below lists the names of the data fields\n")
        io.write("t max_slip_rate moment_rate\n")
        io.write("# Here is the time-series data.\n")
        for n, t in enumerate(times):
            io.write("%21.13E %14.6E %14.6E\n" % (t, _safelog10(glob[n, 0]), glob[n, 1]))


def _write_profiles(m, sol, dir, profile_dt, strike, depth, line_fields, glob, modeler):
    nf, n2f = m.nf, m.n2f
    times = np.asarray(sol.t)
    ts = _profile_times(float(times[-1]), float(profile_dt))
    rows = []
    for t in ts:
        hit = np.nonzero(np.abs(times - t) <= 1e-9 * max(1.0, abs(t)))[0]
        if len(hit) == 0:
            raise EqdError(_lib.EQD_ERR_ARG, f"profile time {t} is not a saved time; choose profile_dt as a "
                                             f"multiple of saveat")
        rows.append(int(hit[0]))
    lines = {"strike": (m.x2, strike, 0), "depth": (m.x3, depth, n2f)}
    quantities = [
        ("slip_2", lambda n, k, off, q: sol.u[n][k], "Horizontal slip (Slip_2) (m)"),
        ("slip_3", lambda n, k, off, q: sol.u[n][nf + k], "Vertical slip (Slip_3) (m)"),
        ("shear_stress_2", lambda n, k, off, q: line_fields[n, off + q, 2] / 1e6,
         "Horizontal shear stress (Shear_stress_2) (MPa)"),
        ("shear_stress_3", lambda n, k, off, q: line_fields[n, off + q, 3] / 1e6,
         "Vertical shear stress (Shear_stress_3) (MPa)"),
        ("pore_pressure", lambda n, k, off, q: sol.u[n][3 * nf + k] / 1e6, "Pore pressure (Pore_pressure) (MPa)"),
    ]
    lf = m.par["l_f"]
    for qname, getter, qdesc in quantities:
        for lname, (axis, idxs, off) in lines.items():
            coord = "x2" if lname == "strike" else "x3"
            along = "along strike" if lname == "strike" else "along depth"
            ncol = len(axis) + 2
            with open(os.path.join(dir, f"{qname}_{lname}.dat"), "w") as io:
                io.write("# This is the file header:\n")
                io.write(f"# problem=SEAS Benchmark {_problem_name(m)}\n")
                io.write(f"# author={modeler}\n")
                io.write(f"# date={_date()}\n")
                io.write(f"# code={CODE_NAME}\n")
                io.write(f"# code_version={CODE_VERSION}\n")
                io.write(f"# element_size={_fmt_real(m.grid_info['dz'])} m\n")
                io.write(f"# Row #1 = {'Strike' if lname == 'strike' else 'Depth'} (m) with two zeros first\n")
                io.write("# Column #1 = Time (s)\n")
                io.write("# Column #2 = Max slip rate (log10 m/s)\n")
                io.write(f"# Columns #3-{ncol} = {qdesc} {along}\n")
                io.write(f"# Computational domain size:  {_fmt_real(-lf)}m < x2 < {_fmt_real(lf)}m, "
                         f"{_fmt_real(-lf)}m < x3 < {_fmt_real(lf)}m\n")
                io.write(_domain_line(m) + "\n")
                io.write("# The line below lists the names of the data fields\n")
                io.write(f"{coord}\nt\nmax_slip_rate\n{qname}\n")
                io.write("# Here are the data\n")
                io.write("%21.13E %14.6E" % (0.0, 0.0))
                for x in axis:
                    io.write(" %14.6E" % x)
                io.write("\n")
                for t, n in zip(ts, rows):
                    io.write("%21.13E %14.6E" % (t, _safelog10(glob[n, 0])))
                    for q, k in enumerate(idxs):
                        io.write(" %14.6E" % getter(n, k, off, q))
                    io.write("\n")
