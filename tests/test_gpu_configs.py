"""Every benchmark configuration run by a device test, each against the CPU oracle.

* dz = 100 m, the CPU-runnable configuration in its legal form (L_fault = 800, L_normal = 800:
  n1 = 5 at L_normal = 400 is rejected by build_model, src/BP8.jl:199-202), l_f = L_fault = 800:
  the exact K (289 nodes, 578 columns) in FP64 and mixed precision within 1e-8 of the oracle's,
  then 30 days of BP8-QD-GS with V_max(t) within 1e-6 of the oracle's integration of the same K
  at every save (or twice the oracle's own spread under ulp-level scalings of K, where larger).
* dz = 50 m, L = (800, 400), l_f = 800 (the headline grid: 1089 nodes, 2178 columns): 64 seeded
  columns of the oracle's K (tests/golden/c2_columns.npz) against the device's FP64 and mixed
  builds of the same columns, 1e-8 (Frobenius over the sample and per column).
* The serving mix over the three serving grids, l_f = L_fault: (100, 800, 800), (50, 800, 400)
  and (25, 1600, 800): every request's result bitwise equal alone and in the mix, and the
  dz = 100 and 50 m results within 1e-8 of the oracle.
* dz = 20 m, L = (800, 400), l_f = 800 with the Peaceman well (nf = 6561, K 13122^2 from the
  device Toeplitz build): ensemble members against single-trajectory device runs with the
  member's parameters, member 0 against the oracle's integration of the same K.
* The Toeplitz K at l_f = L_fault (the patch reaches the Dirichlet ring): the device's within
  1e-8 of the oracle's on the ring grid and at dz = 50 m, L = (800, 400).
"""
import math
import os

import numpy as np
import pytest

import eqd
from eqd import bp8
from eqd.ensemble import Ensemble, perturbed_members
from eqd.serve import Server
from oracle.oracle import Geometry, OracleBP8, OracleSystem, bp8_geometry

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MU, NU = 32.04e9, 0.25
LAM = 2 * MU * NU / (1 - 2 * NU)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# --------------------------------------------------------------------------------------------
# dz = 100 m, L = (800, 800), l_f = 800: exact K and 30 days

@pytest.fixture(scope="module")
def c1_oracle_K():
    print("[configs: dz=100 l_f=800 oracle exact K]", flush=True)
    return OracleSystem(bp8_geometry(100.0, 800.0, 800.0, 800.0)).fault_stiffness(nthreads=8)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c1_exact_K(c1_oracle_K, prec):
    print(f"[configs: dz=100 l_f=800 device exact K {prec}]", flush=True)
    par = bp8.benchmark_parameters(l_f=800.0, dz=100.0)
    m = bp8.build_model(par=par, dz=100.0, L_fault=800.0, L_normal=800.0, stiffness="exact", precision=prec)
    assert m.nf == 289
    K = m.K
    assert K.shape == (578, 578)
    assert rel(K, c1_oracle_K) <= 1e-8
    rep = eqd.solver_report(m.fe.rs)
    assert rep["solves"] == 578 and rep["unconverged"] == 0


def test_c1_30day_gs(c1_oracle_K):
    print("[configs: dz=100 l_f=800 30 days GS]", flush=True)
    K = c1_oracle_K
    n2f = 17
    tf, saveat = 30 * 86400.0, 3600.0
    par = bp8.benchmark_parameters(l_f=800.0, dz=100.0)
    m = bp8.model_from_stiffness(K, n2f, par=par)
    sol = bp8.run_bp8(m, tspan=(0.0, tf), saveat=saveat)
    ro = OracleBP8(K, n2f, dict(l_f=800.0, dz=100.0)).run(0.0, tf, saveat)
    assert sol.retcode == "Success" and np.array_equal(sol.t, ro["t"])
    assert abs(sol.naccept - ro["naccept"]) <= 2
    spread = np.zeros_like(ro["vmax"])
    for f in (1 + 1e-15, 1 - 1e-15, 1 + 4e-15, 1 - 4e-15):
        r = OracleBP8(K * f, n2f, dict(l_f=800.0, dz=100.0)).run(0.0, tf, saveat)
        spread = np.maximum(spread, np.abs(r["vmax"] - ro["vmax"]) / ro["vmax"])
    bound = np.maximum(1e-6, 2 * spread)
    vmax = bp8.saved_vmax(m, len(sol.t))
    d = np.abs(vmax - ro["vmax"]) / ro["vmax"]
    print(f"  naccept {sol.naccept} / {ro['naccept']}; V_max max rel diff {d.max():.2e}; saves held to 1e-6: "
          f"{int((bound == 1e-6).sum())} of {len(bound)}", flush=True)
    assert np.all(d <= bound)
    nf = m.nf
    assert np.all(np.abs(sol.u[:, :2 * nf] - ro["u"][:, :2 * nf]) <= 1e-6 * np.abs(ro["u"][:, :2 * nf]) + 1e-12)


# --------------------------------------------------------------------------------------------
# the headline grid: 64 seeded columns of the oracle's 2178-column K

@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c2_columns_against_oracle_fixture(prec):
    print(f"[configs: dz=50 l_f=800 64 fixture columns {prec}]", flush=True)
    z = np.load(os.path.join(HERE, "golden", "c2_columns.npz"))
    cols, Kref = z["cols"], z["K"]
    fe = eqd.FaultElasticity(eqd.bp8_grid_pair(50.0, 800.0, 400.0), LAM, MU, l_f=800.0, precision=prec)
    assert fe.op.nf == 1089 and fe.op.ntot == 58806
    K = eqd.fault_stiffness(fe, cols=cols)
    assert K.shape == Kref.shape == (2178, 64)
    assert rel(K, Kref) <= 1e-8
    per_col = np.linalg.norm(K - Kref, axis=0) / np.linalg.norm(Kref, axis=0)
    print(f"  Frobenius {rel(K, Kref):.2e}, worst column {per_col.max():.2e}", flush=True)
    assert per_col.max() <= 1e-8
    assert eqd.solver_report(fe.rs)["unconverged"] == 0


# --------------------------------------------------------------------------------------------
# serving over the three serving grids

SERVE_SPECS = [(100.0, 800.0, 800.0), (50.0, 800.0, 400.0), (25.0, 1600.0, 800.0)]


def serve_stream(seed=20261203, n=14):
    """A seeded request stream over the three grids (weights 0.5 / 0.35 / 0.15): k ~ U{1..64}
    unit-slip columns or a forward solve with a Gaussian slip bump (amplitude 1e-3 m, width
    U[100, 400] m), as the serving benchmark draws them; arrivals are given as the loop step at
    which each request is submitted (Poisson gaps of mean 1 step)."""
    rng = np.random.default_rng(seed)
    nfs = [int(2 * Lf / dz) + 1 for dz, Lf, _ in SERVE_SPECS]
    reqs, step = [], 0
    for _ in range(n):
        step += int(rng.poisson(1.0))
        gi = int(rng.choice(3, p=[0.5, 0.35, 0.15]))
        n2f = nfs[gi]
        nf = n2f * n2f
        if rng.random() < 0.5:
            k = int(rng.integers(1, 65)) if gi < 2 else int(rng.integers(1, 9))
            reqs.append((step, gi, "columns", rng.choice(2 * nf, size=k, replace=False), None))
        else:
            Lf = SERVE_SPECS[gi][1]
            x = np.linspace(-Lf, Lf, n2f)
            X, Y = np.meshgrid(x, x, indexing="ij")
            w = rng.uniform(100.0, 400.0)
            c2, c3 = rng.uniform(-Lf, Lf, 2)
            bump = (1e-3 * np.exp(-((X - c2) ** 2 + (Y - c3) ** 2) / (2 * w * w))).ravel(order="F")
            z = np.zeros(nf)
            reqs.append((step, gi, "forward", None, (bump, z) if rng.random() < 0.5 else (z, bump)))
    return reqs


def make_server(only=None):
    """A server holding the three grids (or only grid `only`, at the same pool size); returns it and
    the grid ids indexed like SERVE_SPECS."""
    sv = Server(precision="mixed")
    gids = {}
    for gi, (dz, Lf, Ln) in enumerate(SERVE_SPECS):
        if only is None or gi == only:
            gids[gi] = sv.add_grid(eqd.bp8_grid_pair(dz, Lf, Ln), LAM, MU, l_f=Lf,
                                   max_columns=64 if dz == 25.0 else 128)
    return sv, gids


def submit(sv, gids, req):
    _, gi, kind, cols, slip = req
    return sv.submit_columns(gids[gi], cols) if kind == "columns" else sv.submit_forward(gids[gi], *slip)


def fetch(sv, gids, rid, req):
    _, gi, kind, cols, _ = req
    out, _, unc = sv.result(rid, gids[gi], 1 if kind == "forward" else len(cols))
    assert unc == 0
    return out


def test_c4_serving_mix_alone_and_oracle():
    reqs = serve_stream()
    print(f"[configs: serving mix of {len(reqs)} requests over dz = 100 / 50 / 25 m]", flush=True)
    sv, gids = make_server()
    rids, i, step = [], 0, 0
    while i < len(reqs):  # requests arrive while earlier ones are in flight
        while i < len(reqs) and reqs[i][0] <= step:
            rids.append(submit(sv, gids, reqs[i]))
            i += 1
        sv.step()
        step += 1
    sv.drain(timeout_s=240.0)
    mixed = [fetch(sv, gids, rid, r) for rid, r in zip(rids, reqs)]
    del sv
    print("  mix drained; running each request alone", flush=True)
    for j, r in enumerate(reqs):
        alone, ga = make_server(only=r[1])
        rid = submit(alone, ga, r)
        alone.drain(timeout_s=120.0)
        assert np.array_equal(fetch(alone, ga, rid, r), mixed[j]), (j, r[1], r[2])
        del alone
    print("  alone == mix bitwise; oracle check on the dz = 100 and 50 m requests", flush=True)
    oracles = {}
    for j, r in enumerate(reqs):
        _, gi, kind, cols, slip = r
        if gi == 2:
            continue
        dz, Lf, Ln = SERVE_SPECS[gi]
        if gi not in oracles:
            oracles[gi] = OracleSystem(bp8_geometry(dz, Lf, Ln, Lf))
        o = oracles[gi]
        if kind == "columns":
            sel = cols[:8]
            ref = o.fault_stiffness_cols(list(map(int, sel)), nthreads=8)
            got = mixed[j][:, :len(sel)]
        else:
            d2, d3 = o.shear_traction(*slip)
            ref = np.concatenate([d2, d3])[:, None]
            got = mixed[j].reshape(ref.shape)
        assert rel(got, ref) <= 1e-8, (j, gi, kind, rel(got, ref))


# --------------------------------------------------------------------------------------------
# dz = 20 m Peaceman-well ensemble (nf = 6561)

def test_c5_peaceman_members():
    print("[configs: dz=20 l_f=800 Peaceman ensemble, 4 members, 1 h]", flush=True)
    par = bp8.benchmark_parameters(l_f=800.0, dz=20.0)
    m = bp8.build_model(par=par, dz=20.0, L_fault=800.0, L_normal=400.0, injection="peaceman",
                        stiffness="toeplitz", precision="mixed")
    assert m.nf == 6561 and m.state_len == 4 * 6561 + 1
    M, tf, saveat = 4, 3600.0, 900.0
    mp = perturbed_members(m.par, M, seed=5)
    ens = Ensemble(m, M, **mp)
    stats = ens.run(tspan=(0.0, tf), saveat=saveat)
    assert all(s.retcode == 0 for s in stats)
    K = m.K
    nf = m.nf
    for j in range(M):
        pj = bp8.benchmark_parameters(l_f=800.0, dz=20.0, Q0=float(mp["Q0"][j]), a=float(mp["a"][j]),
                                      sigma0=float(mp["sigma0"][j]))
        mj = bp8.model_from_stiffness(K, m.n2f, par=pj, injection="peaceman")
        single = bp8.run_bp8(mj, tspan=(0.0, tf), saveat=saveat)
        sol = ens.solution(j)
        assert np.array_equal(sol.t, single.t)
        assert abs(stats[j].naccept - single.naccept) <= max(2, single.naccept // 200), (j, stats[j].naccept)
        for a, b, floor in ((0, 2 * nf, 1e-12), (3 * nf, 4 * nf + 1, 1e-3)):
            assert np.all(np.abs(sol.u[:, a:b] - single.u[:, a:b]) <= 1e-6 * np.abs(single.u[:, a:b]) + floor), (j, a)
        print(f"  member {j}: {stats[j].naccept} steps (single run {single.naccept})", flush=True)
        del mj
    print("  member 0 against the oracle (dense K GEMV on the host)", flush=True)
    o = OracleBP8(K, m.n2f, dict(l_f=800.0, dz=20.0, Q0=float(mp["Q0"][0]), a=float(mp["a"][0]),
                                 sigma0=float(mp["sigma0"][0])), injection="peaceman")
    ro = o.run(0.0, tf, saveat)
    s0 = ens.solution(0)
    assert np.array_equal(s0.t, ro["t"])
    assert abs(stats[0].naccept - ro["naccept"]) <= max(2, ro["naccept"] // 200)
    assert np.all(np.abs(s0.u[:, 3 * nf:] - ro["u"][:, 3 * nf:]) <= 1e-6 * np.abs(ro["u"][:, 3 * nf:]) + 1e-3)
    vmax, _ = ens.saved_vmax(len(s0.t))
    assert np.all(np.abs(vmax[0] - ro["vmax"]) <= 1e-6 * ro["vmax"])


# --------------------------------------------------------------------------------------------
# Toeplitz K with l_f = L_fault

@pytest.mark.parametrize("case", [("ring", (9, 17, 400.0, 800.0, 800.0)), ("dz50", (9, 33, 400.0, 800.0, 800.0))])
def test_toeplitz_lf_equals_L_fault(case):
    name, (n1, n23, Ln, Lf, lf) = case
    print(f"[configs: Toeplitz K at l_f = L_fault, {name}]", flush=True)
    ref = OracleSystem(Geometry(n1, n23, Ln, Lf, LAM, MU, lf)).fault_stiffness_toeplitz()
    for prec in ("fp64", "mixed"):
        fe = eqd.FaultElasticity(eqd.GridPair(n1, n23, Ln, Lf), LAM, MU, l_f=lf, precision=prec)
        K = eqd.fault_stiffness_toeplitz(fe)
        assert eqd.solver_report(fe.rs)["solves"] == 10
        assert rel(K, ref) <= 1e-8, (prec, rel(K, ref))
