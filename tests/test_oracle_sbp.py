"""Pins of the oracle's SBP order-4 stencils (oracle/sbp4.cpp) by the identities that
define them. The reference takes the stencils from Diffinitive's
standard_diagonal.toml (src/BP8.jl:195), which is not available; these identities
plus the published stiffness scalars (test_oracle_pins.py) pin the restatement."""
import numpy as np
import pytest

from oracle.oracle import sbp1d


@pytest.mark.parametrize("n,h", [(9, 1.0), (13, 0.2), (33, 50.0)])
def test_first_derivative_sbp_identity(n, h):
    H, D1, _, _, _ = sbp1d(n, h)
    Q = np.diag(H) @ D1 + D1.T @ np.diag(H)
    B = np.zeros((n, n)); B[0, 0] = -1; B[-1, -1] = 1
    assert np.abs(Q - B).max() < 1e-12


@pytest.mark.parametrize("n,h", [(9, 1.0), (17, 0.5), (33, 50.0)])
def test_second_derivative_sbp_identity(n, h):
    H, D1, D2, dl, dr = sbp1d(n, h)
    e0 = np.zeros(n); e0[0] = 1
    eN = np.zeros(n); eN[-1] = 1
    M = -(np.diag(H) @ D2 + np.outer(e0, dl) - np.outer(eN, dr))
    assert np.abs(M - M.T).max() < 1e-12 * np.abs(M).max()
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > -1e-12 * np.abs(M).max()


@pytest.mark.parametrize("n", [9, 12, 21])
def test_polynomial_exactness(n):
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    H, D1, D2, dl, dr = sbp1d(n, h)
    for k in range(3):  # D1 exact to degree 2 in the interior, closures degree 2 (order 4 / 2)
        assert np.allclose(D1 @ x ** k, k * x ** max(k - 1, 0), atol=1e-9)
    assert np.allclose(D2 @ x ** 2, 2 * np.ones(n), atol=1e-8)
    assert np.allclose(D2 @ x, 0, atol=1e-8)
    # boundary derivative is third-order: exact on cubics
    for k in range(4):
        assert abs(dl @ x ** k - (k * x[0] ** (k - 1) if k else 0)) < 1e-8
        assert abs(dr @ x ** k - k * x[-1] ** (k - 1) if k else dr @ x ** k) < 1e-8
    assert H.sum() == pytest.approx(x[-1] - x[0], rel=1e-14)


def test_interior_accuracy_order_four():
    errs = []
    for n in (41, 81):
        h = 1.0 / (n - 1)
        x = np.arange(n) * h
        _, D1, D2, _, _ = sbp1d(n, h)
        sl = slice(8, n - 8)
        errs.append(np.abs((D1 @ np.sin(x))[sl] - np.cos(x)[sl]).max())
    assert errs[0] / errs[1] > 14  # ~2^4
