"""The device BP8-QD-PW loop (Peaceman well) over 100 h against the CPU oracle and the
reference's published state (dz = 50 m, L = (800, 400), l_f = 400; PROGRESS.md:1017-1044).

Bars: the save grid identical; accepted steps within 0.5% of the oracle's (the floored well
cell makes the step sequence stiff: ~31k steps, each a chance for a threshold tie in the
controller); at every hourly save the well pressure, the well-cell pressure and the pressure
field within 1e-6 relative, V_max within 1e-6 relative. The published 100 h state on the
device: well-cell p 25.85 MPa, sigma_bar_lowest -0.85 MPa (the sigma_bar floor reached at one
node), 31,249 steps within 0.5%, and the well pressure equal to eq. 25 at r = r_w to 0.2% (the
Peaceman construction; a five-point r_e = 0.198 dz would miss by 3.7%).
"""
import math

import numpy as np
import pytest

from eqd import bp8
from oracle.oracle import OracleBP8, OracleSystem, bp8_geometry

pytestmark = pytest.mark.gpu

HOURS = 100


def eq25(r, t, P):
    """BP8 eq. 25 (src/BP8.jl:532-541), with scipy's E1."""
    from scipy.special import exp1
    q0 = P["Q0"] / P["L_fwid"]
    return q0 / (4 * math.pi * P["alpha"] * P["beta"] * P["phi"]) * exp1(r * r / (4 * P["alpha"] * t))


@pytest.fixture(scope="module")
def pw_runs():
    print("[peaceman 100 h: oracle Toeplitz K, oracle run]", flush=True)
    K = OracleSystem(bp8_geometry(50.0, 800.0, 400.0, 400.0)).fault_stiffness_toeplitz()
    n2f = int(round(math.sqrt(K.shape[0] // 2)))
    o = OracleBP8(K, n2f, dict(dz=50.0), injection="peaceman")
    ro = o.run(0.0, HOURS * 3600.0, 3600.0)
    print(f"[peaceman 100 h: device run] oracle naccept {ro['naccept']}", flush=True)
    m = bp8.model_from_stiffness(K, n2f, par=bp8.benchmark_parameters(dz=50.0), injection="peaceman")
    sol = bp8.run_bp8(m, tspan=(0.0, HOURS * 3600.0), saveat=3600.0)
    print(f"[peaceman 100 h] device naccept {sol.naccept} in {sol.ms:.0f} ms", flush=True)
    return m, sol, ro


def test_peaceman_100h_matches_oracle(pw_runs):
    m, sol, ro = pw_runs
    nf = m.nf
    ic = (m.n2f // 2) * (m.n2f + 1)
    assert sol.retcode == "Success"
    assert np.array_equal(sol.t, ro["t"])
    assert abs(sol.naccept - ro["naccept"]) <= 0.005 * ro["naccept"], (sol.naccept, ro["naccept"])
    rel = lambda a, b: np.abs(a - b) / np.abs(b)  # noqa: E731
    # hour 0 is the initial state (p = 0): compare from the first save on
    assert np.all(rel(sol.u[1:, -1], ro["u"][1:, -1]) <= 1e-6)                       # p_well
    assert np.all(rel(sol.u[1:, 3 * nf + ic], ro["u"][1:, 3 * nf + ic]) <= 1e-6)     # well-cell p
    dp = np.abs(sol.u[:, 3 * nf:4 * nf] - ro["u"][:, 3 * nf:4 * nf])
    assert np.all(dp <= 1e-6 * np.abs(ro["u"][:, 3 * nf:4 * nf]) + 1e-3)
    vmax = bp8.saved_vmax(m, len(sol.t))
    d = rel(vmax, ro["vmax"])
    print(f"  V_max max rel diff {d.max():.2e} (median {np.median(d):.2e})", flush=True)
    assert np.all(d <= 1e-6)


def test_peaceman_100h_published_state(pw_runs):
    m, sol, ro = pw_runs
    nf, P = m.nf, m.par
    ic = (m.n2f // 2) * (m.n2f + 1)
    u = sol.u[-1]
    p = u[3 * nf:4 * nf]
    assert int(np.argmax(p)) == ic
    assert f"{p[ic] / 1e6:.2f}" == "25.85"
    assert f"{(P['sigma0'] - p.max()) / 1e6:.2f}" == "-0.85"
    assert abs(sol.naccept - 31249) <= 0.005 * 31249
    ref = eq25(P["r_well"], HOURS * 3600.0, P)
    assert abs(u[-1] - ref) <= 2e-3 * ref
    rep = bp8.effective_stress_report(m)
    assert rep["bound"] and rep["nodes"] == 1 and rep["sigma_lowest"] < -0.8e6
