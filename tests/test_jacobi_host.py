"""CPU check of jacobi_preconditioner (src/ElasticitySplitNode.jl:343-357) as libeqd computes it
on the host from its operator tables (no device needed).

Pins:
* against the oracle's assembled A: M^{-1} = 1 / diag(A), A's zero diagonal entries floored to 1;
* the reference docstring's own count at n1 = 9, n23 = 13 (:345-349): 3,318 of the 9,126
  diagonal entries are exactly zero, and that set is the far-field DOF set;
* rows rm and rp of a fault pair carry the same entry (:312-315: a diagonal M commutes with P).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from eqd import _lib
from oracle.oracle import OP_A, OP_P, Geometry, OracleSystem

CASES = [
    (11, 11, 1.0, 1.0, 2.0, 1.0, 0.4),               # reference test cube (test/fault_response_test.jl:83)
    (9, 13, 1.0, 1.2, 2.0, 1.0, 0.6),                # the docstring's n1 = 9, n23 = 13 grid
    (9, 17, 800.0, 800.0, 32.04e9, 32.04e9, 800.0),  # BP8 at dz = 100 m, L_normal = 800
]


def device_minv(n1, n23, Ln, Lf, lam, mu, lf):
    g = _lib.GridDesc(n1, n23, Ln, Lf, lam, mu, lf)
    n = 6 * n1 * n23 * n23
    minv = np.zeros(n)
    _lib.check(_lib.lib().eqd_jacobi_preconditioner(C.byref(g), minv, n))
    return minv


@pytest.mark.parametrize("case", CASES)
def test_minv_is_inverse_diagonal_of_assembled_A(case):
    o = OracleSystem(Geometry(*case))
    rp, col, val = o.export(OP_A)
    A = sp.csr_matrix((val, col, rp), shape=(o.ntot, o.ntot))
    d = A.diagonal()
    assert d.min() >= 0
    ref = np.where(d == 0, 1.0, 1.0 / np.where(d == 0, 1.0, d))
    got = device_minv(*case)
    assert np.array_equal(got == 1.0, (d == 0) | (ref == 1.0))
    assert np.max(np.abs(got - ref) / ref) <= 1e-13


def test_far_field_count_matches_reference_docstring():
    n1, n23 = 9, 13
    case = (n1, n23, 1.0, 1.2, 2.0, 1.0, 0.6)
    o = OracleSystem(Geometry(*case))
    rp, col, val = o.export(OP_A)
    d = sp.csr_matrix((val, col, rp), shape=(o.ntot, o.ntot)).diagonal()
    assert o.ntot == 9126 and int(np.sum(d == 0)) == 3318
    got = device_minv(*case)
    # the floored entries are exactly the far-field DOFs: the rows P zeroes
    rp, col, val = o.export(OP_P)
    P = sp.csr_matrix((val, col, rp), shape=(o.ntot, o.ntot))
    far = np.asarray(abs(P).sum(axis=1)).ravel() == 0
    assert int(far.sum()) == 3318
    assert np.all(got[far] == 1.0) and np.all(d[~far] > 0)


@pytest.mark.parametrize("case", CASES[:2])
def test_fault_pair_rows_share_their_entry(case):
    n1, n23 = case[0], case[1]
    N = n1 * n23 * n23
    got = device_minv(*case)
    for c in range(3):
        for i3 in range(1, n23 - 1):
            for i2 in range(1, n23 - 1):
                rm = c * N + (n1 - 1) + n1 * (i2 + n23 * i3)     # face of g-
                rp = 3 * N + c * N + 0 + n1 * (i2 + n23 * i3)    # face of g+
                assert got[rm] == got[rp]


def test_rejects_wrong_length_and_small_grids():
    g = _lib.GridDesc(11, 11, 1.0, 1.0, 2.0, 1.0, 0.4)
    buf = np.zeros(10)
    assert _lib.lib().eqd_jacobi_preconditioner(C.byref(g), buf, 10) == _lib.EQD_ERR_ARG
    g = _lib.GridDesc(7, 11, 1.0, 1.0, 2.0, 1.0, 0.4)
    buf = np.zeros(6 * 7 * 121)
    assert _lib.lib().eqd_jacobi_preconditioner(C.byref(g), buf, buf.size) == _lib.EQD_ERR_GRID
