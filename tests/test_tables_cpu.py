"""CPU check of libeqd's host-built operator tables (no device needed).

The kernels apply the reduced operator -H P (D+SAT) P, the right-hand side
H P (D+SAT) chi(s) and the traction map 1/2 (T U- + T U+) as sums of tensor products
of 1D tables. tests/_tablemodel.py combines the exported tables the way the kernels
do; here that model is compared with the oracle's assembled CSR operators
(src/ElasticitySplitNode.jl:106-214, src/FaultResponse.jl:130-132, 201-209)."""
import numpy as np
import pytest

from _tablemodel import TableModel, from_full, to_full
from oracle.oracle import OP_A, OP_HPDSAT, OP_P, OP_T2, OP_T3, Geometry, OracleSystem

CASES = [
    (11, 11, 1.0, 1.0, 2.0, 1.0, 0.4),          # reference test cube (test/fault_response_test.jl:83)
    (9, 13, 1.0, 1.2, 2.0, 1.0, 0.6),           # unequal n1 != n23, h1 != h2
    (9, 17, 400.0, 800.0, 32.04e9, 32.04e9, 800.0),  # l_f = L_fault: Omega_f includes the ring
]


@pytest.mark.parametrize("case", CASES)
def test_tables_reproduce_assembled_operators(case):
    n1, n23, Ln, Lf, lam, mu, lf = case
    o = OracleSystem(Geometry(n1, n23, Ln, Lf, lam, mu, lf))
    tm = TableModel(n1, n23, Ln, Lf, lam, mu, lf)
    N = o.N
    rng = np.random.default_rng(7)
    v = [rng.standard_normal((tm.NY, tm.NY, tm.NX)) for _ in range(3)]
    x = to_full(tm, v, N)
    y_or = o.apply(OP_A, x)
    assert np.abs(y_or - to_full(tm, tm.apply(v), N)).max() <= 1e-13 * np.abs(y_or).max()
    # A maps any vector through P first: the merged model sees only P x
    xr = rng.standard_normal(o.ntot)
    assert np.abs(o.apply(OP_A, xr) - to_full(tm, tm.apply(from_full(tm, o.apply(OP_P, xr), N)), N)).max() \
        <= 1e-13 * np.abs(o.apply(OP_A, xr)).max()

    s2 = rng.standard_normal(o.nf); s3 = rng.standard_normal(o.nf)
    chi = o.build_chi(s2, s3)
    b_or = o.apply(OP_HPDSAT, chi)
    f2 = np.zeros((n23, n23)); f3 = np.zeros((n23, n23))
    for k in range(o.nf):
        a, b = k % o.n2f, k // o.n2f
        f2[o.i3lo + b, o.i2lo + a] = s2[k]
        f3[o.i3lo + b, o.i2lo + a] = s3[k]
    assert np.abs(b_or - to_full(tm, tm.rhs(f2, f3), N)).max() <= 1e-13 * np.abs(b_or).max()

    U = o.apply(OP_P, x) + chi
    om = [(o.i2lo + k % o.n2f) + n23 * (o.i3lo + k // o.n2f) for k in range(o.nf)]
    t2, t3 = o.apply(OP_T2, U)[om], o.apply(OP_T3, U)[om]
    e2, e3 = tm.extract(v, f2, f3, o.i2lo, o.n2f)
    assert np.abs(t2 - e2).max() <= 1e-13 * np.abs(t2).max()
    assert np.abs(t3 - e3).max() <= 1e-13 * np.abs(t3).max()


def test_table_widths_fit_kernel_limit():
    tm = TableModel(33, 129, 800.0, 1600.0, 32.04e9, 32.04e9, 1600.0)
    from eqd import _lib
    for name in ("AXs", "AXo", "MX", "Ts", "To", "D1", "D1s", "D1full", "D1ext"):
        start, width, _ = _lib.tables_export(tm.g, name)
        assert width.max() <= 16
