"""K s from the separation kernel of a Toeplitz K (the time loop of a build_model(stiffness =
:toeplitz) model) against the dense product of the same K.

A Toeplitz K (src/FaultResponse.jl:378-390) is a function of node separation in each of its four
blocks, so the device keeps 4 (2 n2f - 1)^2 kernel values instead of streaming 2 rows of K per
node, and forms K s either as the correlation summed directly or as a zero-padded 2D FFT
convolution. The entries are K's own; the FFT rounds differently from a sum: K s within 1e-12 of
the dense product K @ s (norm-wise and in the largest entry), for random slip, unit slip at the
patch corners (the separations that reach the padding) and a smooth bump; rhs! within 1e-12
relative of the dense-K model (model_from_stiffness with m.K) by either method; 100 h
trajectories and FFT-convolution ensembles within the bars of tests/test_gpu_bp8.py and
tests/test_gpu_ensemble.py."""
import numpy as np
import pytest

import eqd
from eqd import bp8
from eqd.ensemble import Ensemble, perturbed_members
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


def _slips(nf, n2f, rng):
    yield "random", 1e-3 * rng.standard_normal(2 * nf)
    s = np.zeros(2 * nf)
    s[0] = 1.0                     # s2 at the (x2, x3) = (-l_f, -l_f) corner
    yield "corner s2", s
    s = np.zeros(2 * nf)
    s[2 * nf - 1] = 1.0            # s3 at the (+l_f, +l_f) corner
    yield "corner s3", s
    s = np.zeros(2 * nf)
    s[n2f - 1] = -0.5              # s2 at (+l_f, -l_f), s3 at (-l_f, +l_f)
    s[nf + nf - n2f] = 2.0
    yield "corners s2 s3", s
    x = np.linspace(-1.0, 1.0, n2f)
    bump = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / 0.1).ravel()
    yield "bump", np.concatenate([0.8 * bump, -0.6 * bump])


@pytest.mark.parametrize("dz,l_f", [(50.0, 400.0), (20.0, 800.0)])
def test_ks_product_matches_dense(dz, l_f):
    """K s by each method against numpy's dense K @ s of the same K, within 1e-12."""
    print(f"[toeplitz K s product dz={dz} l_f={l_f}]", flush=True)
    par = bp8.benchmark_parameters(l_f=l_f, dz=dz)
    m = bp8.build_model(par=par, dz=dz, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    K = m.K
    nf, n2f = m.nf, m.n2f
    rng = np.random.default_rng(17)
    methods = ["fft", "direct"] if n2f > 48 else ["fft"]
    got = {}
    slips = list(_slips(nf, n2f, rng))
    for method in methods:
        m.set_ks_method(method)
        for name, s in slips:
            out = m.ks(s)
            ref = K @ s
            e2 = np.linalg.norm(out - ref) / np.linalg.norm(ref)
            emax = np.abs(out - ref).max() / np.abs(ref).max()
            print(f"  {method:6s} {name:14s} norm {e2:.2e}  max {emax:.2e}", flush=True)
            assert e2 <= 1e-12 and emax <= 1e-12, (method, name)
            got[method, name] = out
        print(f"  {method}: {m.ks_timing(50) * 1e3:.1f} us per K s", flush=True)
    if len(methods) == 2:
        for name, _ in slips:
            d = got["fft", name] - got["direct", name]
            assert np.abs(d).max() <= 1e-12 * np.abs(got["direct", name]).max()
    # a dense-K model has no separation kernel; a small patch's direct K s lives in its stage kernel
    md = bp8.model_from_stiffness(K, n2f, par=m.par)
    with pytest.raises(eqd.EqdError):
        md.set_ks_method("fft")
    if n2f <= 48:
        m.set_ks_method("direct")
        with pytest.raises(eqd.EqdError):
            m.ks(np.zeros(2 * nf))


@pytest.mark.parametrize("dz,l_f,method", [(50.0, 400.0, "direct"), (50.0, 400.0, "fft"), (20.0, 800.0, "fft"),
                                           (20.0, 800.0, "direct")])
def test_kernel_rhs_matches_dense(dz, l_f, method):
    print(f"[toeplitz K s rhs dz={dz} l_f={l_f} {method}]", flush=True)
    par = bp8.benchmark_parameters(l_f=l_f, dz=dz)
    m = bp8.build_model(par=par, dz=dz, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    m.set_ks_method(method)
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


@pytest.mark.parametrize("method", ["direct", "fft"])
def test_kernel_run_100h_matches_oracle(method):
    print(f"[toeplitz K s 100 h {method}]", flush=True)
    m = bp8.build_model(dz=50.0, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    m.set_ks_method(method)
    sol = bp8.run_bp8(m, tspan=(0.0, 100 * 3600.0), saveat=3600.0)
    o = OracleBP8(m.K, m.n2f, dict(dz=50.0))
    ro = o.run(0.0, 100 * 3600.0, 3600.0)
    nf = m.nf
    assert np.array_equal(sol.t, ro["t"])
    assert abs(sol.naccept - ro["naccept"]) <= 2
    vmax = bp8.saved_vmax(m, len(sol.t))
    assert np.all(np.abs(vmax - ro["vmax"]) <= 1e-5 * ro["vmax"])
    assert np.all(np.abs(sol.u[:, :2 * nf] - ro["u"][:, :2 * nf]) <= 1e-6 * np.abs(ro["u"][:, :2 * nf]) + 1e-12)


def test_fft_ensemble_members_match_oracle_and_gemm():
    """An ensemble of a :toeplitz model forms each stage's K S as one FFT convolution batched over
    the members; on the dz = 50 m model (17 x 17 nodes, transform 64 x 64) its members match the
    oracle run with their parameters over 100 h (bars of tests/test_gpu_ensemble.py) and the
    members of the GEMM ensemble over the same model's dense K."""
    print("[toeplitz K S: FFT ensemble, 4 members, 100 h]", flush=True)
    m = bp8.build_model(dz=50.0, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    K = m.K
    nf = m.nf
    mp = perturbed_members(m.par, 4, seed=23)
    tf = 100 * 3600.0
    ens = Ensemble(m, 4, **mp)
    st = ens.run(tspan=(0.0, tf), saveat=3600.0)
    ms, flops = ens.gemm_timing(20)
    print(f"  FFT K S over 4 members: {ms * 1e3:.1f} us (flops reported {flops})", flush=True)
    assert flops == 0.0
    md = bp8.model_from_stiffness(K, m.n2f, par=m.par)
    eg = Ensemble(md, 4, **mp)
    sg = eg.run(tspan=(0.0, tf), saveat=3600.0)
    for j in range(4):
        sol, solg = ens.solution(j), eg.solution(j)
        assert st[j].retcode == 0 and np.array_equal(sol.t, solg.t)
        assert abs(st[j].naccept - sg[j].naccept) <= 2
        assert np.all(np.abs(sol.u[:, :2 * nf] - solg.u[:, :2 * nf]) <= 1e-6 * np.abs(solg.u[:, :2 * nf]) + 1e-12)
    for j in (0, 3):
        pj = dict(dz=50.0, Q0=float(mp["Q0"][j]), a=float(mp["a"][j]), sigma0=float(mp["sigma0"][j]))
        ro = OracleBP8(K, m.n2f, pj).run(0.0, tf, 3600.0)
        sol = ens.solution(j)
        assert np.array_equal(sol.t, ro["t"])
        assert abs(st[j].naccept - ro["naccept"]) <= 2
        d = np.max(np.abs(sol.u[:, :2 * nf] - ro["u"][:, :2 * nf]) / (np.abs(ro["u"][:, :2 * nf]) + 1e-12))
        print(f"  member {j}: {st[j].naccept} steps (oracle {ro['naccept']}), slip max rel diff {d:.2e}", flush=True)
        assert np.all(np.abs(sol.u[:, :2 * nf] - ro["u"][:, :2 * nf]) <= 1e-6 * np.abs(ro["u"][:, :2 * nf]) + 1e-12)
        assert np.all(np.abs(sol.u[:, 3 * nf:] - ro["u"][:, 3 * nf:]) <= 1e-6 * np.abs(ro["u"][:, 3 * nf:]) + 1e-3)
