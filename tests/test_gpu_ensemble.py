"""Trajectory ensembles (members in lockstep stages, K S as one FP64 MFMA GEMM per stage for a
dense K, as one batched FFT convolution for a :toeplitz model)
through the C ABI, member by member against the CPU oracle run with that member's Q0, a
and sigma0.

Bars: the same as the single-trajectory device loop against the oracle -- saved slip and
pressure within 1e-6 relative and V_max within 1e-5 at every saved time over 100 h, the
save grid identical, the step counts within 2; V_max at the saved states within 1e-10 of
evaluate! on a one-member model with the member's parameters (the GEMM and the GEMV sum
in different orders, the Newton stops at |d ln V| < 1e-12).
"""
import numpy as np
import pytest

import eqd
from eqd import bp8
from eqd.ensemble import Ensemble, perturbed_members
from oracle.oracle import OracleBP8, OracleSystem, bp8_geometry

pytestmark = pytest.mark.gpu


def close(a, b, rtol, atol=0.0):
    return np.all(np.abs(a - b) <= rtol * np.abs(b) + atol)


@pytest.fixture(scope="module")
def toeplitz_K():
    o = OracleSystem(bp8_geometry(50.0, 800.0, 400.0, 400.0))
    return o.fault_stiffness_toeplitz()


def member_par(base, mp, j):
    return dict(dz=50.0, Q0=float(mp["Q0"][j]), a=float(mp["a"][j]), sigma0=float(mp["sigma0"][j]))


@pytest.mark.parametrize("injection,members,hours,saveat", [("gaussian", 5, 100, 3600.0),
                                                            ("peaceman", 3, 10, 600.0)])
def test_members_match_oracle(toeplitz_K, injection, members, hours, saveat):
    print(f"[ensemble {injection} M={members} {hours} h]", flush=True)
    n2f = 17
    nf = n2f * n2f
    par = bp8.benchmark_parameters(dz=50.0)
    m = bp8.model_from_stiffness(toeplitz_K, n2f, par=par, injection=injection)
    mp = perturbed_members(par, members, seed=3)
    ens = Ensemble(m, members, **mp)
    tf = hours * 3600.0
    stats = ens.run(tspan=(0.0, tf), saveat=saveat)
    for j in range(members):
        o = OracleBP8(toeplitz_K, n2f, member_par(par, mp, j), injection=injection)
        ro = o.run(0.0, tf, saveat)
        sol = ens.solution(j)
        assert stats[j].retcode == 0
        assert np.array_equal(sol.t, ro["t"]), j
        tol_steps = 2 if injection == "gaussian" else max(3, ro["naccept"] // 1000)
        assert abs(stats[j].naccept - ro["naccept"]) <= tol_steps, (j, stats[j].naccept, ro["naccept"])
        pw = injection == "peaceman"
        d = [np.max(np.abs(sol.u[:, a:b] - ro["u"][:, a:b]) / (np.abs(ro["u"][:, a:b]) + f))
             for a, b, f in [(0, 2 * nf, 1e-12), (2 * nf, 3 * nf, 0.0), (3 * nf, sol.u.shape[1], 1e-3)]]
        print(f"  member {j}: naccept {stats[j].naccept} / {ro['naccept']}, max rel diff slip {d[0]:.2e} "
              f"ln theta {d[1]:.2e} p {d[2]:.2e}", flush=True)
        assert close(sol.u[:, :2 * nf], ro["u"][:, :2 * nf], 1e-5 if pw else 1e-6, 1e-12), j
        assert close(sol.u[:, 2 * nf:3 * nf], ro["u"][:, 2 * nf:3 * nf], 1e-7 if pw else 1e-8), j
        assert close(sol.u[:, 3 * nf:], ro["u"][:, 3 * nf:], 1e-6, 1e-3), j
    vmax, _ = ens.saved_vmax(len(ro["t"]))
    for j in range(members):
        o = OracleBP8(toeplitz_K, n2f, member_par(par, mp, j), injection=injection)
        ro = o.run(0.0, tf, saveat)
        assert close(vmax[j], ro["vmax"], 1e-5), j


def test_members_differ_and_base_member_matches_single_run(toeplitz_K):
    """Member 0 carries the base parameters: its trajectory is the single-model device run's
    to rounding; perturbed members give different trajectories."""
    print("[ensemble base member vs run_bp8]", flush=True)
    n2f = 17
    nf = n2f * n2f
    par = bp8.benchmark_parameters(dz=50.0)
    m = bp8.model_from_stiffness(toeplitz_K, n2f, par=par)
    mp = perturbed_members(par, 4, seed=9)
    ens = Ensemble(m, 4, **mp)
    ens.run(tspan=(0.0, 50 * 3600.0), saveat=3600.0)
    sol = bp8.run_bp8(m, tspan=(0.0, 50 * 3600.0), saveat=3600.0)
    s0 = ens.solution(0)
    assert np.array_equal(s0.t, sol.t)
    assert close(s0.u[:, :2 * nf], sol.u[:, :2 * nf], 1e-8, 1e-14)
    assert close(s0.u[:, 3 * nf:], sol.u[:, 3 * nf:], 1e-8, 1e-6)
    s1 = ens.solution(1)
    assert not close(s1.u[-1, :2 * nf], sol.u[-1, :2 * nf], 1e-3)


@pytest.mark.parametrize("members", [1, 70])
def test_saved_vmax_matches_evaluate(toeplitz_K, members):
    """One GEMM per save over all members (ragged: 578 rows, 1 or 70 member columns) against
    evaluate! of single models holding each member's parameters."""
    print(f"[ensemble saved_vmax M={members}]", flush=True)
    n2f = 17
    par = bp8.benchmark_parameters(dz=50.0)
    m = bp8.model_from_stiffness(toeplitz_K, n2f, par=par)
    mp = perturbed_members(par, members, seed=5)
    ens = Ensemble(m, members, **mp)
    ens.run(tspan=(0.0, 5 * 3600.0), saveat=1800.0)
    vmax, _ = ens.saved_vmax(11)
    for j in sorted({0, members // 2, members - 1}):
        mj = bp8.model_from_stiffness(toeplitz_K, n2f, par=bp8.benchmark_parameters(**member_par(par, mp, j)))
        sol = ens.solution(j)
        ref = np.array([bp8.evaluate(mj, u, t)["Vmag"].max() for t, u in zip(sol.t, sol.u)])
        assert close(vmax[j], ref, 1e-10), j


def test_per_member_reltol_and_maxiters(toeplitz_K):
    print("[ensemble reltol / maxiters]", flush=True)
    n2f = 17
    par = bp8.benchmark_parameters(dz=50.0)
    m = bp8.model_from_stiffness(toeplitz_K, n2f, par=par)
    ens = Ensemble(m, 2)
    st = ens.run(tspan=(0.0, 20 * 3600.0), saveat=3600.0, member_reltol=[1e-8, 1e-6])
    assert st[0].naccept > st[1].naccept
    o = OracleBP8(toeplitz_K, n2f, dict(dz=50.0))
    ro = o.run(0.0, 20 * 3600.0, 3600.0, reltol=1e-6)
    assert abs(st[1].naccept - ro["naccept"]) <= 2
    st = ens.run(tspan=(0.0, 20 * 3600.0), saveat=3600.0, maxiters=5)
    assert all(s.retcode == 1 for s in st)


def test_bad_arguments(toeplitz_K):
    n2f = 17
    m = bp8.model_from_stiffness(toeplitz_K, n2f, par=bp8.benchmark_parameters(dz=50.0))
    with pytest.raises(eqd.EqdError):
        Ensemble(m, 0)
    with pytest.raises(eqd.EqdError):
        Ensemble(m, 3, a=[0.01, -0.01, 0.01])
    with pytest.raises(eqd.EqdError):
        Ensemble(m, 3, Q0=[1.0, 2.0])
    ens = Ensemble(m, 2)
    with pytest.raises(eqd.EqdError):
        ens.run(tspan=(0.0, 3600.0), saveat=600.0, member_reltol=[1e-8, 0.0])


def test_production_scale_ensemble_matches_single_models():
    """The dz = 20 m ensemble configuration (nf = 6561, K 13122^2 from the device Toeplitz build,
    K S as one 256 x 256 FFT convolution batched over the members): member 0 (base parameters)
    against run_bp8 of the same model (its own FFT K s), and saved V_max of two perturbed members
    against evaluate! of one-member dense-K models with their parameters."""
    print("[ensemble dz=20 nf=6561 M=16 3 h]", flush=True)
    par = bp8.benchmark_parameters(l_f=800.0, dz=20.0)
    m = bp8.build_model(par=par, dz=20.0, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    mp = perturbed_members(m.par, 16, seed=11)
    ens = Ensemble(m, 16, **mp)
    ens.run(tspan=(0.0, 3 * 3600.0), saveat=1800.0)
    s0 = ens.solution(0)
    single = bp8.run_bp8(m, tspan=(0.0, 3 * 3600.0), saveat=1800.0)
    nf = m.nf
    assert np.array_equal(s0.t, single.t)
    assert close(s0.u[:, :2 * nf], single.u[:, :2 * nf], 1e-6, 1e-12)
    assert close(s0.u[:, 3 * nf:4 * nf], single.u[:, 3 * nf:4 * nf], 1e-6, 1e-3)
    n = len(s0.t)
    vmax, _ = ens.saved_vmax(n)
    K = m.K
    for j in (5, 15):
        mj = bp8.model_from_stiffness(K, m.n2f, par=bp8.benchmark_parameters(
            l_f=800.0, dz=20.0, Q0=float(mp["Q0"][j]), a=float(mp["a"][j]), sigma0=float(mp["sigma0"][j])))
        sol = ens.solution(j)
        ref = np.array([bp8.evaluate(mj, u, t)["Vmag"].max() for t, u in zip(sol.t, sol.u)])
        assert close(vmax[j], ref, 1e-10), j
        del mj
