"""K s from the separation kernel of a Toeplitz K (the time loop of a build_model(stiffness =
:toeplitz) model) against the dense product of the same K.

A Toeplitz K (src/FaultResponse.jl:378-390) is a function of node separation in each of its four
blocks, so the device keeps 4 (2 n2f - 1)^2 kernel values instead of streaming 2 rows of K per
node. The entries are K's own, only the summation order differs: rhs! within 1e-12 relative of
the dense-K model (model_from_stiffness with m.K) and of the CPU oracle; 100 h trajectories
within the bars of tests/test_gpu_bp8.py."""
import numpy as np
import pytest

from eqd import bp8
from oracle.oracle import OracleBP8

pytestmark = pytest.mark.gpu


def _states(m, rng, n=4):
    nf = m.nf
    u0 = bp8.initial_state(m)
    for trial in range(n):
        u = u0.copy()
        u[:2 * nf] += 1e-3 * rng.standard_normal(2 * nf)
        u[2 * nf:3 * nf] += 0.1 * rng.standard_normal(nf)
        u[3 * nf:4 * nf] = 1e6 * rng.random(nf)
        yield u, 3600.0 * 40 * trial


@pytest.mark.parametrize("dz,l_f", [(50.0, 400.0), (20.0, 800.0)])
def test_kernel_rhs_matches_dense(dz, l_f):
    print(f"[toeplitz K s dz={dz} l_f={l_f}]", flush=True)
    par = bp8.benchmark_parameters(l_f=l_f, dz=dz)
    m = bp8.build_model(par=par, dz=dz, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    K = m.K
    md = bp8.model_from_stiffness(K, m.n2f, par=m.par)
    nf = m.nf
    rng = np.random.default_rng(5)
    for u, t in _states(m, rng):
        du = bp8.rhs(np.zeros_like(u), u, m, t)
        dd = bp8.rhs(np.zeros_like(u), u, md, t)
        cd = bp8.evaluate(md, u, t)
        c = bp8.evaluate(m, u, t)
        # a stress component that nearly cancels (tens of Pa out of 1e7-Pa terms) carries the
        # sum's rounding (~1e-15 of the terms): stresses get that absolute floor, slip-rate
        # components the same floor relative to |V| (V_c = |V| T_c / |T|)
        tfloor = 1e-14 * np.abs(cd["tau2"]).max()
        np.testing.assert_allclose(c["tau2"], cd["tau2"], rtol=1e-12, atol=tfloor)
        np.testing.assert_allclose(c["tau3"], cd["tau3"], rtol=1e-12, atol=tfloor)
        vm = np.concatenate([cd["Vmag"], cd["Vmag"]])
        assert np.all(np.abs(du[:2 * nf] - dd[:2 * nf]) <= 1e-12 * (np.abs(dd[:2 * nf]) + vm) + 1e-30)
        for lo, hi, atol in [(2 * nf, 3 * nf, 1e-14), (3 * nf, 4 * nf, 1e-9)]:
            assert np.all(np.abs(du[lo:hi] - dd[lo:hi]) <= 1e-12 * np.abs(dd[lo:hi]) + atol)


def test_kernel_run_100h_matches_oracle():
    print("[toeplitz K s 100 h]", flush=True)
    m = bp8.build_model(dz=50.0, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    sol = bp8.run_bp8(m, tspan=(0.0, 100 * 3600.0), saveat=3600.0)
    o = OracleBP8(m.K, m.n2f, dict(dz=50.0))
    ro = o.run(0.0, 100 * 3600.0, 3600.0)
    nf = m.nf
    assert np.array_equal(sol.t, ro["t"])
    assert abs(sol.naccept - ro["naccept"]) <= 2
    vmax = bp8.saved_vmax(m, len(sol.t))
    assert np.all(np.abs(vmax - ro["vmax"]) <= 1e-5 * ro["vmax"])
    assert np.all(np.abs(sol.u[:, :2 * nf] - ro["u"][:, :2 * nf]) <= 1e-6 * np.abs(ro["u"][:, :2 * nf]) + 1e-12)
