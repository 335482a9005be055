"""Parity of the device BP8 time loop (through the C ABI) with the CPU oracle.

Bars: rhs!/evaluate! entries within 1e-10 relative (the Newton stops at |d ln V| < 1e-12;
the K product and A_p differ from the oracle only in summation order); Tsit5 runs within
1e-6 relative in slip and pressure, and in V_max at every saved time within 1e-6 or, where the
trajectory itself is that ill-conditioned, twice the oracle's own spread under ulp-level
scalings of K (the step sequences agree, the states differ by rounding amplified by the
adaptive controller);
the published 100 h scalars (PROGRESS.md:584-593) end to end with K built on the device.
"""
import math

import numpy as np
import pytest

import eqd
from eqd import _lib, bp8
from oracle.oracle import OracleBP8, OracleSystem, bp8_geometry

pytestmark = pytest.mark.gpu


def synthetic_K(n2f, seed=7):
    nf = n2f * n2f
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((2 * nf, 2 * nf)) * 1e6
    return -(B @ B.T) / (2 * nf) - 4e8 * np.eye(2 * nf)


def close(a, b, rtol, atol=0.0):
    return np.all(np.abs(a - b) <= rtol * np.abs(b) + atol)


@pytest.fixture(scope="module")
def toeplitz_K():
    o = OracleSystem(bp8_geometry(50.0, 800.0, 400.0, 400.0))
    return o.fault_stiffness_toeplitz()


@pytest.mark.parametrize("injection", ["gaussian", "peaceman"])
def test_rhs_and_evaluate_match_oracle(injection):
    print(f"[bp8 rhs {injection}]", flush=True)
    n2f = 17
    K = synthetic_K(n2f)
    par = bp8.benchmark_parameters(dz=50.0)
    m = bp8.model_from_stiffness(K, n2f, par=par, injection=injection)
    o = OracleBP8(K, n2f, dict(dz=50.0), injection=injection)
    u0 = bp8.initial_state(m)
    assert np.array_equal(u0, o.initial_state())
    rng = np.random.default_rng(11)
    nf = m.nf
    for trial in range(4):
        u = u0.copy()
        u[:2 * nf] += 1e-3 * rng.standard_normal(2 * nf)
        u[2 * nf:3 * nf] += 0.1 * rng.standard_normal(nf)
        u[3 * nf:4 * nf] = 1e6 * rng.random(nf)
        if injection == "peaceman":
            u[4 * nf] = 2e6
        t = 3600.0 * trial * 40  # crosses t_off = 100 h
        du = bp8.rhs(np.zeros_like(u), u, m, t)
        duo = o.rhs(u, t)
        for lo, hi, atol in [(0, 2 * nf, 1e-30), (2 * nf, 3 * nf, 1e-14), (3 * nf, 4 * nf, 1e-9)]:
            assert close(du[lo:hi], duo[lo:hi], 1e-10, atol * (1 + np.abs(duo[lo:hi]).max())), (trial, lo)
        if injection == "peaceman":
            assert close(du[4 * nf:], duo[4 * nf:], 1e-12)
        c = bp8.evaluate(m, u, t)
        co = o.evaluate(u, t)
        for k in ("V2", "V3", "Vmag"):
            assert close(c[k], co[k], 1e-10, 1e-30), k
        for k in ("tau2", "tau3"):
            assert close(c[k], co[k], 1e-12, 1e-6), k


def test_operators_match_oracle():
    n2f = 17
    K = synthetic_K(n2f)
    m = bp8.model_from_stiffness(K, n2f, par=bp8.benchmark_parameters(dz=50.0))
    o = OracleBP8(K, n2f, dict(dz=50.0))
    A, src, w = m.operators()
    so, wo = o.source()
    assert np.allclose(A, o.Ap(), rtol=0, atol=1e-15 * np.abs(o.Ap()).max())
    assert np.array_equal(src, so) and np.array_equal(w, wo)
    assert np.array_equal(m.K, K)


def test_newton_failure_is_reported_and_cleared():
    n2f = 17
    m = bp8.model_from_stiffness(synthetic_K(n2f), n2f, par=bp8.benchmark_parameters(dz=50.0))
    u = bp8.initial_state(m)
    bad = u.copy()
    bad[2 * m.nf + 5 * n2f + 5] = np.nan  # an active node: the Newton iterate never settles
    with pytest.raises(eqd.EqdError) as e:
        bp8.rhs(np.zeros_like(u), bad, m, 0.0)
    assert e.value.code == _lib.EQD_ERR_NEWTON
    bp8.rhs(np.zeros_like(u), u, m, 0.0)  # the flag does not leak into the next call


def _run_pair(K, tf, saveat, injection="gaussian"):
    n2f = int(round(math.sqrt(K.shape[0] // 2)))
    m = bp8.model_from_stiffness(K, n2f, par=bp8.benchmark_parameters(dz=50.0), injection=injection)
    o = OracleBP8(K, n2f, dict(dz=50.0), injection=injection)
    sol = bp8.run_bp8(m, tspan=(0.0, tf), saveat=saveat)
    ro = o.run(0.0, tf, saveat)
    return m, sol, ro


SCALINGS = (1 + 1e-15, 1 - 1e-15, 1 + 4e-15, 1 - 4e-15)


def vmax_envelope(K, ro, tf, saveat):
    """Per save, the bound a device V_max(t) is held to: max(1e-6, 2 x the oracle's own spread
    when K is scaled by 1 +- 1e-15 and 1 +- 4e-15 -- a few ulps, the size of any change of
    summation order). Where the trajectory is well conditioned the spread is ~1e-9 and the bound
    is 1e-6; only near the V_max minimum after the injection stops does it rise above."""
    n2f = int(round(math.sqrt(K.shape[0] // 2)))
    spread = np.zeros_like(ro["vmax"])
    for f in SCALINGS:
        r = OracleBP8(K * f, n2f, dict(dz=50.0)).run(0.0, tf, saveat)
        assert np.array_equal(r["t"], ro["t"])
        spread = np.maximum(spread, np.abs(r["vmax"] - ro["vmax"]) / ro["vmax"])
    return np.maximum(1e-6, 2 * spread), spread


def test_run_100h_matches_oracle(toeplitz_K):
    print("[bp8 run 100 h toeplitz]", flush=True)
    tf = 100 * 3600.0
    m, sol, ro = _run_pair(toeplitz_K, tf, 3600.0)
    nf = m.nf
    assert sol.retcode == "Success"
    assert np.array_equal(sol.t, ro["t"])
    assert abs(sol.naccept - ro["naccept"]) <= 2 and abs(sol.nreject - ro["nreject"]) <= 2
    assert close(sol.u[:, :2 * nf], ro["u"][:, :2 * nf], 1e-6, 1e-12)
    assert close(sol.u[:, 2 * nf:3 * nf], ro["u"][:, 2 * nf:3 * nf], 1e-8)
    assert close(sol.u[:, 3 * nf:], ro["u"][:, 3 * nf:], 1e-6, 1e-3)
    vmax = bp8.saved_vmax(m, len(sol.t))
    bound, spread = vmax_envelope(toeplitz_K, ro, tf, 3600.0)
    d = np.abs(vmax - ro["vmax"]) / ro["vmax"]
    print(f"  V_max: max rel diff {d.max():.2e}, oracle spread max {spread.max():.2e}, saves held to 1e-6: "
          f"{int((bound == 1e-6).sum())} of {len(bound)}", flush=True)
    assert np.all(d <= bound)
    assert np.all(bound == 1e-6)  # over the first 100 h the 1e-6 bar applies at every save


def test_run_30day_published_gs(toeplitz_K):
    """PROGRESS.md:671 (30 days, saveat 300, :toeplitz): peak V 1.10e-6 at 2.22 days, final
    slip 0.0401 m, peak p(0,0) 13.17 MPa -- and the oracle's trajectory: V_max(t) at every save
    within max(1e-6, 2 x the oracle's own spread under K x (1 +- 1e-15), (1 +- 4e-15)). Near day 19
    (after the injection stops, V ~ 2e-8 m/s) a last-bit change of K moves the oracle itself by
    ~1e-6 (tests/test_oracle_bp8.py::test_30day_vmax_conditioning); everywhere else the 1e-6 bar
    holds: the oracle's spread leaves 8337 of the 8641 saves at 1e-6 (all but days 18.5-25.9)."""
    print("[bp8 run 30 d toeplitz]", flush=True)
    tf = 30 * 86400.0
    m, sol, ro = _run_pair(toeplitz_K, tf, 300.0)
    nf = m.nf
    ic = (m.n2f // 2) * (m.n2f + 1)
    assert len(sol.t) == 8641 and sol.t[-1] == tf
    vmax = bp8.saved_vmax(m, len(sol.t))
    j = int(np.argmax(vmax))
    assert f"{vmax[j]:.2e}" == "1.10e-06" and f"{sol.t[j] / 86400:.2f}" == "2.22"
    assert f"{sol.u[-1, ic]:.4f}" == "0.0401"
    assert f"{sol.u[:, 3 * nf + ic].max() / 1e6:.2f}" == "13.17"
    bound, spread = vmax_envelope(toeplitz_K, ro, tf, 300.0)
    d = np.abs(vmax - ro["vmax"]) / ro["vmax"]
    n6 = int((bound == 1e-6).sum())
    k = int(np.argmax(d / bound))
    print(f"  V_max: max rel diff {d.max():.2e} at day {sol.t[int(np.argmax(d))] / 86400:.2f}; worst ratio to bound "
          f"{(d / bound)[k]:.2f} at day {sol.t[k] / 86400:.2f}; saves held to 1e-6: {n6} of {len(bound)}; "
          f"oracle spread max {spread.max():.2e}", flush=True)
    assert np.all(d <= bound)
    assert n6 >= 8300
    assert close(sol.u[:, :2 * nf], ro["u"][:, :2 * nf], 1e-6, 1e-12)


def test_build_model_exact_on_device_reproduces_published_100h():
    """build_model(dz=50, L_fault=800, L_normal=400, stiffness=:exact) with K built and kept
    on the device; 100 h with saveat 3600 (PROGRESS.md:584-593)."""
    print("[bp8 build_model exact 100 h]", flush=True)
    m = bp8.build_model(dz=50.0, L_fault=800.0, L_normal=400.0, stiffness="exact", precision="mixed")
    nf = m.nf
    ic = (m.n2f // 2) * (m.n2f + 1)
    K = m.K
    assert f"{K[ic, ic]:.5e}" == "-4.38404e+08"
    sol = bp8.run_bp8(m, tspan=(0.0, 100 * 3600.0), saveat=3600.0)
    u = sol.u[-1]
    c = bp8.evaluate(m, u, sol.t[-1])
    assert f"{u[ic]:.5e}" == "3.99595e-02"
    assert f"{c['Vmag'].max():.5e}" == "1.66456e-07"
    assert abs(u[3 * nf + ic] / 1e6 - 13.16528) <= 1.5e-5
    rep = bp8.effective_stress_report(m)
    assert not rep["bound"] and rep["sigma_lowest"] > 1e6


def test_build_model_toeplitz_matches_oracle(toeplitz_K):
    m = bp8.build_model(dz=50.0, L_fault=800.0, L_normal=400.0)
    K = m.K
    assert np.linalg.norm(K - toeplitz_K) <= 1e-8 * np.linalg.norm(toeplitz_K)


def test_peaceman_10h_matches_oracle(toeplitz_K):
    print("[bp8 run 10 h peaceman]", flush=True)
    m, sol, ro = _run_pair(toeplitz_K, 10 * 3600.0, 600.0, injection="peaceman")
    nf = m.nf
    assert np.array_equal(sol.t, ro["t"])
    assert abs(sol.naccept - ro["naccept"]) <= max(3, ro["naccept"] // 1000)
    assert close(sol.u[:, 3 * nf:], ro["u"][:, 3 * nf:], 1e-6, 1e-3)
    assert close(sol.u[:, :2 * nf], ro["u"][:, :2 * nf], 1e-5, 1e-12)
