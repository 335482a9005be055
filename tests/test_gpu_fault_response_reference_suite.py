"""The reference's FaultResponse test set (test/fault_response_test.jl) on the device, on its own
fixtures and with its own bounds.

Fixtures (test/fault_response_test.jl:10-26): lambda = 2, mu = 1; FE = the n = 13 cube pair over
L = 1.2 with l_f = 0.6 (Omega_f 7 x 7); the n = 11 unit cube pair with l_f = 0.4.
"""
import numpy as np
import pytest

import eqd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fe13():
    print("[fault response suite: n=13, L=1.2, l_f=0.6]", flush=True)
    fe = eqd.FaultElasticity(eqd.GridPair(13, 13, 1.2, 1.2), 2.0, 1.0, l_f=0.6)
    return fe, eqd.fault_stiffness(fe)


def cube11(**kw):
    return eqd.FaultElasticity(eqd.GridPair(11, 11, 1.0, 1.0), 2.0, 1.0, l_f=0.4, **kw)


def test_omega_f_lands_on_l_f_and_is_x2_fastest(fe13):
    fe, _ = fe13
    x2, x3 = eqd.fault_grid_axes(fe)
    assert np.isclose(x2[0], -0.6) and np.isclose(x2[-1], 0.6)
    assert np.isclose(x3[0], -0.6) and np.isclose(x3[-1], 0.6)
    assert eqd.frictional_node_count(fe) == len(x2) * len(x3)
    assert np.all(np.diff(x2) > 0) and np.all(np.diff(x3) > 0)
    with pytest.raises(eqd.EqdError):  # a spacing that misses +-l_f is rejected, not snapped
        eqd.FaultElasticity(eqd.GridPair(13, 13, 1.2, 1.2), 2.0, 1.0, l_f=0.55)


def test_stiffness_reproduces_a_direct_solve(fe13):
    fe, K = fe13
    nf = eqd.frictional_node_count(fe)
    k = np.arange(1, nf + 1)  # the reference's 1-based k
    s2, s3 = 0.01 * np.sin(3 * k), 0.02 * np.cos(2 * k)
    d2, d3 = eqd.shear_traction(fe, s2, s3)
    st = K @ np.concatenate([s2, s3])
    assert np.linalg.norm(st[:nf] - d2) <= 1e-8 * np.linalg.norm(d2)
    assert np.linalg.norm(st[nf:] - d3) <= 1e-8 * np.linalg.norm(d3)


def test_slip_relieves_the_stress_driving_it(fe13):
    fe, K = fe13
    nf = eqd.frictional_node_count(fe)
    assert np.all(np.diag(K) < 0)
    d2, _ = eqd.shear_traction(fe, np.ones(nf), np.zeros(nf))
    assert d2.max() < 0


def test_no_slip_gives_no_traction_change(fe13):
    fe, _ = fe13
    nf = eqd.frictional_node_count(fe)
    d2, d3 = eqd.shear_traction(fe, np.zeros(nf), np.zeros(nf))
    assert np.all(d2 == 0) and np.all(d3 == 0)


def test_sign_convention_cube11():
    fe = cube11()
    K = eqd.fault_stiffness(fe)
    assert eqd.elastic_solver_report(fe)["unconverged"] == 0
    assert np.diag(K).max() < 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_build_matches_one_column_at_a_time(precision):
    """The reference's threaded == serial (bitwise, same iteration count); here: all columns in
    one batch against one column per batch."""
    fe_all = cube11(precision=precision, tile_k=16)
    fe_one = cube11(precision=precision, tile_k=16, max_cols_in_flight=1)
    K_all = eqd.fault_stiffness(fe_all)
    K_one = eqd.fault_stiffness(fe_one)
    assert np.array_equal(K_all, K_one)
    assert eqd.solver_report(fe_all.rs)["iterations"] == eqd.solver_report(fe_one.rs)["iterations"]


def test_toeplitz_build_approximates_the_exact_one():
    fe_e, fe_t = cube11(), cube11()
    K_exact = eqd.fault_stiffness(fe_e)
    K_toep = eqd.fault_stiffness_toeplitz(fe_t)
    assert K_toep.shape == K_exact.shape
    assert np.linalg.norm(K_toep - K_exact) / np.linalg.norm(K_exact) < 0.05
    x2, x3 = eqd.fault_grid_axes(fe_e)
    n2, n3 = len(x2), len(x3)
    nf = eqd.frictional_node_count(fe_e)
    c = ((n3 + 1) // 2 - 1) * n2 + (n2 + 1) // 2 - 1  # the centre node, 0-based
    assert np.allclose(K_toep[:nf, c], K_exact[:nf, c], rtol=1e-6, atol=0)
    assert np.allclose(K_toep[nf:, nf + c], K_exact[nf:, nf + c], rtol=1e-6, atol=0)
    assert np.all(np.diag(K_toep)[:nf] < 0)
    assert abs(K_toep[c, c] - K_exact[c, c]) <= 1e-6 * abs(K_exact[c, c])
    assert abs(K_toep[0, 0] - K_exact[0, 0]) > 1e-6 * abs(K_exact[0, 0])
    assert abs(K_toep[0, 0] - K_exact[0, 0]) <= 0.05 * abs(K_exact[0, 0])
    assert eqd.solver_report(fe_t.rs)["solves"] == 10
    assert eqd.solver_report(fe_e.rs)["solves"] == 2 * nf


def test_threaded_build_with_duplicated_solvers_matches_serial():
    """The reference's threaded build (src/FaultResponse.jl:265-281): one duplicated solver per task,
    the operator shared read-only, the columns split in a strided partition; here two host threads
    drive two duplicated solvers (each with its own stream) at once. Bitwise equal to the serial
    build, with the same total iteration count."""
    import threading
    fe = cube11(precision="mixed", tile_k=16)
    K_ser = eqd.fault_stiffness(fe)
    it_ser = eqd.solver_report(fe.rs)["iterations"]
    fe2 = cube11(precision="mixed", tile_k=16)
    nf2 = K_ser.shape[0]
    parts = [list(range(0, nf2, 2)), list(range(1, nf2, 2))]
    solvers = [fe2.rs, eqd.duplicate(fe2.rs)]
    out = [None, None]
    errs = []

    def work(i):
        try:
            out[i] = eqd.fault_stiffness(fe2, cols=parts[i], solver=solvers[i])
        except Exception as e:  # surfaced below
            errs.append(e)

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    K_thr = np.zeros_like(K_ser)
    for i in range(2):
        K_thr[:, parts[i]] = out[i]
    assert np.array_equal(K_thr, K_ser)
    merged = eqd.merge_stats(solvers[0], solvers[1])
    assert eqd.solver_report(merged)["iterations"] == it_ser
