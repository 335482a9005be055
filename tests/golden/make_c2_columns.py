"""Writes tests/golden/c2_columns.npz: 64 seeded columns of the oracle's exact FP64 K on the
headline grid (dz = 50 m, L_fault = 800, L_normal = 400, l_f = L_fault = 800: 1089 Omega_f
nodes, 2178 unit-slip columns, 58,806 DOF), the configuration the K-columns/s metric is quoted
on. The oracle (oracle/libeqd_oracle.so, a CPU restatement of src/FaultResponse.jl:229-287) runs
each column by CG to rtol 1e-10 from x0 = 0.

    python tests/golden/make_c2_columns.py

Contents: cols (int32, 0-based, sorted; drawn without replacement from 0..2177 with
numpy.random.default_rng(20261203)), K (2178 x 64 float64, K[:, j] is column cols[j]),
iterations (the oracle's total CG iterations over the 64 solves)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import OracleSystem, bp8_geometry  # noqa: E402


def main():
    o = OracleSystem(bp8_geometry(50.0, 800.0, 400.0, 800.0))
    assert o.nf == 1089 and o.ntot == 58806
    cols = np.sort(np.random.default_rng(20261203).choice(2 * o.nf, size=64, replace=False)).astype(np.int32)
    K = o.fault_stiffness_cols(cols.tolist(), nthreads=min(8, os.cpu_count() or 1))
    rep = o.report()
    assert rep["unconverged"] == 0
    np.savez_compressed(os.path.join(HERE, "c2_columns.npz"), cols=cols, K=K, iterations=np.int64(rep["iterations"]))
    print(f"wrote c2_columns.npz: {len(cols)} columns, {rep['iterations']} CG iterations")


if __name__ == "__main__":
    main()
