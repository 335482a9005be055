"""The production-size grid (dz = 25 m, L_fault = 1600, L_normal = 800, l_f = 1600: 3,294,918 DOF,
16,641 fault nodes, 1024 workgroups per tile, four-segment pencils walked in x2 blocks) through
properties that do not need the CPU oracle at this size:

* mixed-precision columns within 1e-8 (relative, Frobenius) of the same columns in FP64;
* the Toeplitz build's measured columns (the centre source, both directions) equal the exact build's
  columns to 1e-6 (test/fault_response_test.jl:128-171 pins the same on its small fixture);
* negative self-stiffness at every sampled node and zero slip -> exactly zero traction;
* batching invariance at this grid: a column built alone equals the same column in a batch, bitwise;
* four columns of the seeded exact subset within 1e-8 of the CPU oracle (a committed fixture: the
  oracle needs ~80 s per column at this size).
"""
import numpy as np
import pytest

import eqd

pytestmark = pytest.mark.gpu

MU = 32.04e9
LAM = MU  # nu = 0.25


@pytest.fixture(scope="module")
def grid():
    return eqd.bp8_grid_pair(25.0, 1600.0, 800.0)


def test_mixed_matches_fp64_and_batches_are_bitwise(grid):
    print("[production dz=25: 8 columns mixed vs fp64]", flush=True)
    fe = eqd.FaultElasticity(grid, LAM, MU, l_f=1600.0, precision="mixed", tile_k=64)
    nf = eqd.frictional_node_count(fe)
    assert fe.Ntot == 3294918 and nf == 16641
    rng = np.random.default_rng(20261204)
    cols = np.sort(rng.choice(2 * nf, size=8, replace=False))
    K = eqd.fault_stiffness(fe, cols=cols)
    assert eqd.solver_report(fe.rs)["unconverged"] == 0
    fe64 = eqd.FaultElasticity(grid, LAM, MU, l_f=1600.0, precision="fp64", tile_k=8)
    K64 = eqd.fault_stiffness(fe64, cols=cols)
    assert np.linalg.norm(K - K64) <= 1e-8 * np.linalg.norm(K64)
    for j, c in enumerate(cols):
        assert K[c, j] < 0
    fe1 = eqd.FaultElasticity(grid, LAM, MU, l_f=1600.0, precision="mixed", tile_k=64)
    K1 = eqd.fault_stiffness(fe1, cols=cols[3:4])
    assert np.array_equal(K1[:, 0], K[:, 3])


def test_toeplitz_measured_columns_are_exact(grid):
    print("[production dz=25: toeplitz centre columns]", flush=True)
    fe = eqd.FaultElasticity(grid, LAM, MU, l_f=1600.0, precision="mixed", tile_k=64)
    nf = eqd.frictional_node_count(fe)
    n2 = int(round(np.sqrt(nf)))
    c = ((n2 + 1) // 2 - 1) * n2 + (n2 + 1) // 2 - 1
    KT = eqd.fault_stiffness_toeplitz(fe)
    assert eqd.solver_report(fe.rs)["solves"] == 10
    fe2 = eqd.FaultElasticity(grid, LAM, MU, l_f=1600.0, precision="mixed", tile_k=64)
    Ke = eqd.fault_stiffness(fe2, cols=[c, nf + c])
    assert np.allclose(KT[:nf, c], Ke[:nf, 0], rtol=1e-6, atol=1e-6 * np.abs(Ke[:nf, 0]).max())
    assert np.allclose(KT[nf:, nf + c], Ke[nf:, 1], rtol=1e-6, atol=1e-6 * np.abs(Ke[nf:, 1]).max())
    assert np.all(np.diag(KT) < 0)
    z2, z3 = eqd.shear_traction(fe2, np.zeros(nf), np.zeros(nf))
    assert not z2.any() and not z3.any()


@pytest.mark.parametrize("prec,tile_k", [("mixed", 64), ("fp64", 8)])
def test_exact_subset_against_oracle_fixture(grid, prec, tile_k):
    """Four columns of the production benchmark's seeded 512-column exact subset against the CPU
    oracle's FP64 columns (tests/golden/c3_columns.npz, written by make_c3_columns.py): within
    1e-8 relative, per column."""
    import os
    print(f"[production dz=25: 4 columns {prec} vs oracle fixture]", flush=True)
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_columns.npz"))
    cols, Kref = z["cols"], z["K"]
    fe = eqd.FaultElasticity(grid, LAM, MU, l_f=1600.0, precision=prec, tile_k=tile_k)
    K = eqd.fault_stiffness(fe, cols=cols)
    assert K.shape == Kref.shape == (33282, 4)
    err = np.linalg.norm(K - Kref, axis=0) / np.linalg.norm(Kref, axis=0)
    print(f"  per-column rel err {err}", flush=True)
    assert err.max() <= 1e-8
