"""The oracle against every pin the reference publishes for this path.

Sources: nnz counts PERFORMANCE.md:69-73 (structural: pin the stencil footprint);
zero-diagonal count src/ElasticitySplitNode.jl:343-344; rank(P) PERFORMANCE.md:324-326;
singular-CG invariants test/elasticity_split_node_test.jl:149-177; K sign and linearity
test/fault_response_test.jl:54-88; threaded == serial test/fault_response_test.jl:100-112;
Toeplitz guards test/fault_response_test.jl:128-171; K_self = -4.38404e8
(PROGRESS.md:586, Delta z = 50 m, L = (800, 400), l_f = 400, :exact); the Galerkin spectrum and
CG iteration counts on scripts/split_node_spd.jl's geometry and right-hand side
(src/ElasticitySplitNode.jl:290-292); mean CG iterations, K asymmetry, max diag(K) and the Toeplitz
error at production size (PROGRESS.md:365-367, SYMMETRIC_SAT.md:288-291, PERFORMANCE.md:230-233)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.oracle import OP_A, OP_HPDSAT, OP_P, Geometry, OracleSystem, bp8_geometry


def csr(o, op):
    rp, col, val = o.export(op)
    return sp.csr_matrix((val, col, rp), shape=(o.ntot, o.ntot))


def cube(n, L=1.0, l_f=0.4):
    return OracleSystem(Geometry(n, n, L, L, 2.0, 1.0, l_f))


def test_nnz_pins():
    o11 = cube(11)
    assert (o11.nnz_A, o11.nnz_HPDSAT) == (172898, 214202)
    o15 = cube(15, l_f=3 / 7)
    assert (o15.nnz_A, o15.nnz_HPDSAT, o15.nnz_P) == (563882, 648386, 15210)


def test_zero_diagonal_set_is_far_field():
    o = OracleSystem(Geometry(9, 13, 1.0, 1.0, 2.0, 1.0, 0.5))
    A = csr(o, OP_A)
    assert o.ntot == 9126
    d = A.diagonal()
    assert (d == 0).sum() == 3318
    assert (d >= 0).all()
    P = csr(o, OP_P)
    assert set(np.flatnonzero(d == 0)) == set(np.flatnonzero(np.asarray(P.sum(axis=1)).ravel() == 0))


def test_rank_P():
    # N = 4374 per the reference: 2 blocks of 9^3 nodes x 3 components
    o = OracleSystem(Geometry(9, 9, 1.0, 1.0, 2.0, 1.0, 0.5))
    assert o.ntot == 4374
    P = csr(o, OP_P).toarray()
    assert np.isclose(np.trace(P), 2205)
    assert np.linalg.matrix_rank(P) == 2205


def test_singular_cg_invariants():
    o = cube(11)
    A = csr(o, OP_A)
    assert abs(A - A.T).max() / abs(A).max() < 1e-12
    nf = o.nf
    s2 = np.array([0.01 * np.sin(3 * (k + 1)) for k in range(nf)])
    s3 = np.array([0.02 * np.cos(2 * (k + 1)) for k in range(nf)])
    b = o.apply(OP_HPDSAT, o.build_chi(s2, s3))
    Pb = o.apply(OP_P, b)
    assert np.linalg.norm(b - Pb) / np.linalg.norm(b) < 1e-14
    u, it, ok = o.cg(b)
    assert ok and it > 0
    assert np.linalg.norm(b - A @ u) / np.linalg.norm(b) < 1e-8
    assert np.linalg.norm(u - o.apply(OP_P, u)) / np.linalg.norm(u) < 1e-12


def test_cg_matches_direct_solve_on_range_P():
    """A library direct solve of the Galerkin system S^T A S x = S^T b (S: basis of range(P))."""
    o = cube(11)
    A = csr(o, OP_A)
    P = csr(o, OP_P)
    keep = np.flatnonzero(np.asarray(P.sum(axis=1)).ravel() != 0)
    # basis: one column per independent DOF of range(P) (pairs share a column)
    rows, cols, seen = [], [], {}
    for r in keep:
        cs = P.indices[P.indptr[r]:P.indptr[r + 1]]
        key = int(cs.min())
        if key not in seen:
            seen[key] = len(seen)
        rows.append(r); cols.append(seen[key])
    S = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(o.ntot, len(seen)))
    nf = o.nf
    s2 = np.linspace(0.1, 1.0, nf); s3 = np.cos(np.arange(nf))
    b = o.apply(OP_HPDSAT, o.build_chi(s2, s3))
    import scipy.sparse.linalg as spl
    xs = spl.splu((S.T @ A @ S).tocsc()).solve(S.T @ b)
    u, _, _ = o.cg(b)
    assert np.linalg.norm(S @ xs - u) / np.linalg.norm(u) < 1e-7


@pytest.fixture(scope="module")
def fe11():
    o = cube(11)
    return o, o.fault_stiffness(nthreads=1)


def test_stiffness_linearity_and_sign(fe11):
    o, K = fe11
    nf = o.nf
    assert np.all(np.diag(K) < 0)
    s2 = np.array([0.01 * np.sin(3 * (k + 1)) for k in range(nf)])
    s3 = np.array([0.02 * np.cos(2 * (k + 1)) for k in range(nf)])
    d2, d3 = o.shear_traction(s2, s3)
    st = K @ np.concatenate([s2, s3])
    assert np.allclose(st[:nf], d2, rtol=1e-8, atol=1e-8 * np.abs(d2).max())
    assert np.allclose(st[nf:], d3, rtol=1e-8, atol=1e-8 * np.abs(d3).max())
    z2, z3 = o.shear_traction(np.zeros(nf), np.zeros(nf))
    assert not z2.any() and not z3.any()
    d2, _ = o.shear_traction(np.ones(nf), np.zeros(nf))
    assert d2.max() < 0


def test_threaded_matches_serial_bitwise():
    ser, thr = cube(11), cube(11)
    Ks = ser.fault_stiffness(nthreads=1)
    Kt = thr.fault_stiffness(nthreads=4)
    assert np.array_equal(Ks, Kt)
    assert thr.report()["iterations"] == ser.report()["iterations"]
    assert ser.report()["solves"] == 2 * ser.nf


def test_toeplitz_guards(fe11):
    o, K = fe11
    ot = cube(11)
    Kt = ot.fault_stiffness_toeplitz()
    assert ot.report()["solves"] == 10
    nf = o.nf
    assert np.linalg.norm(Kt - K) / np.linalg.norm(K) < 0.05
    n2 = o.n2f
    c = ((o.n3f + 1) // 2 - 1) * n2 + (n2 + 1) // 2 - 1
    assert np.allclose(Kt[:nf, c], K[:nf, c], rtol=1e-6)
    assert np.allclose(Kt[nf:, nf + c], K[nf:, nf + c], rtol=1e-6)
    assert np.all(np.diag(Kt)[:nf] < 0)
    assert not np.isclose(Kt[0, 0], K[0, 0], rtol=1e-6)
    assert np.isclose(Kt[0, 0], K[0, 0], rtol=0.05)


def test_bad_l_f_rejected():
    with pytest.raises(ValueError):
        OracleSystem(Geometry(13, 13, 1.2, 1.2, 2.0, 1.0, 0.55))


def test_K_self_published_value():
    """K_self = K[ic, ic] at Delta z = 50 m, L_fault = 800, L_normal = 400, l_f = 400: -4.38404e8."""
    o = OracleSystem(bp8_geometry(50.0, 800.0, 400.0, 400.0))
    assert o.ntot == 58806 and o.nf == 289
    ic = (o.n2f // 2) + (o.n3f // 2) * o.n2f
    col = o.fault_stiffness_cols([ic])[:, 0]
    assert float(f"{col[ic]:.5e}") == -4.38404e8


def _galerkin(o):
    """S^T A S with S the 0/1 basis of range(P) (fault pairs share a column), as the reference's
    scripts/split_node_spd.jl:128-132 forms it."""
    A, P = csr(o, OP_A), csr(o, OP_P)
    keep = np.flatnonzero(np.asarray(abs(P).sum(axis=1)).ravel() != 0)
    rows, cols, seen = [], [], {}
    for r in keep:
        key = int(P.indices[P.indptr[r]:P.indptr[r + 1]].min())
        if key not in seen:
            seen[key] = len(seen)
        rows.append(r)
        cols.append(seen[key])
    S = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(o.ntot, len(seen)))
    return (S.T @ A @ S).tocsr()


def _script_rhs(o, n):
    """The right-hand side of scripts/split_node_spd.jl:68-69,168-169: HP(D+SAT) chi for the
    Gaussian slip s2 = exp(-(x2^2 + x3^2) / (2 * 0.25^2)), s3 = 0 on the whole fault face
    ([-1, 1]^2, l_f = 1 so Omega_f is every face node, x2 fastest)."""
    x = -1.0 + (2.0 / (n - 1)) * np.arange(n)
    X2, X3 = np.meshgrid(x, x, indexing="ij")
    s2 = np.exp(-(X2 ** 2 + X3 ** 2) / (2 * 0.25 ** 2)).ravel(order="F")
    return o.apply(OP_HPDSAT, o.build_chi(s2, np.zeros_like(s2)))


def test_published_spectrum_and_cg_iterations():
    """src/ElasticitySplitNode.jl:290-292 and PROGRESS.md:365-367 on the geometry of
    scripts/split_node_spd.jl:123-128 (blocks [-1,0] x [-1,1]^2 and [0,1] x [-1,1]^2, n^3 nodes
    each, lambda = 2, mu = 1): on range(A) the operator is positive definite with
    lambda in [0.101, 11.9], kappa = 118 at n = 9 (2205 reduced DOF), and CG on the singular A
    from x0 = 0 takes 71 iterations at n = 9 and 108 at n = 13 for the script's Gaussian slip."""
    o9 = OracleSystem(Geometry(9, 9, 1.0, 1.0, 2.0, 1.0, 1.0))
    G = _galerkin(o9)
    assert G.shape == (2205, 2205)
    ev = np.linalg.eigvalsh(0.5 * (G + G.T).toarray())
    assert f"{ev.min():.3g}" == "0.101" and f"{ev.max():.3g}" == "11.9"
    assert round(np.abs(ev).max() / np.abs(ev).min()) == 118
    for n, want in ((9, 71), (13, 108)):
        o = o9 if n == 9 else OracleSystem(Geometry(n, n, 1.0, 1.0, 2.0, 1.0, 1.0))
        assert o.nf == n * n
        _, it, ok = o.cg(_script_rhs(o, n))
        assert ok and it == want, (n, it)


def test_published_production_K_pins(oracle_exact_dz50_lf400):
    """Delta z = 50 m, L = (800, 400), l_f = 400 (578 columns): mean CG iterations 147
    (PROGRESS.md:365-367), ||K - K^T|| / ||K|| = 0.2003% and max diag(K) = -4.38e8
    (SYMMETRIC_SAT.md:288-291), and the Toeplitz K within 0.0155% (Frobenius) of the exact one
    (PERFORMANCE.md:230-233)."""
    K, rep = oracle_exact_dz50_lf400
    assert rep["solves"] == 578 and rep["unconverged"] == 0
    assert round(rep["mean_iterations"]) == 147
    assert f"{100 * np.linalg.norm(K - K.T) / np.linalg.norm(K):.4f}" == "0.2003"
    assert f"{np.diag(K).max():.2e}" == "-4.38e+08"
    Kt = OracleSystem(bp8_geometry(50.0, 800.0, 400.0, 400.0)).fault_stiffness_toeplitz()
    assert f"{100 * np.linalg.norm(Kt - K) / np.linalg.norm(K):.4f}" == "0.0155"


def test_published_dz100_K_pins():
    """Delta z = 100 m, L = (800, 800), l_f = 400 (SYMMETRIC_SAT.md:288-291):
    ||K - K^T|| / ||K|| = 0.1827%, max diag(K) = -2.19e8."""
    o = OracleSystem(bp8_geometry(100.0, 800.0, 800.0, 400.0))
    K = o.fault_stiffness(nthreads=4)
    assert K.shape == (162, 162)
    assert f"{100 * np.linalg.norm(K - K.T) / np.linalg.norm(K):.4f}" == "0.1827"
    assert f"{np.diag(K).max():.2e}" == "-2.19e+08"
