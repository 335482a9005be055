"""Writes tests/golden/c3_columns.npz: 4 columns of the oracle's exact FP64 K on the production
grid (dz = 25 m, L_fault = 1600, L_normal = 800, l_f = L_fault = 1600: 16,641 Omega_f nodes,
33,282 unit-slip columns, 3,294,918 DOF). The columns are the first 4 (in draw order) of the
512-column seeded subset the production benchmark builds: numpy.random.default_rng(20261203)
.choice(33282, 512, replace=False). The oracle (oracle/libeqd_oracle.so, a CPU restatement of
src/FaultResponse.jl:229-287) takes ~50 s to assemble the system and ~80 s per column per thread.

    python tests/golden/make_c3_columns.py

Contents: cols (int32, 0-based), K (33282 x 4 float64, K[:, j] is column cols[j]), iterations
(the oracle's total CG iterations over the 4 solves)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import OracleSystem, bp8_geometry  # noqa: E402


def main():
    o = OracleSystem(bp8_geometry(25.0, 1600.0, 800.0, 1600.0))
    assert o.nf == 16641 and o.ntot == 3294918
    cols = np.random.default_rng(20261203).choice(2 * o.nf, size=512, replace=False)[:4].astype(np.int32)
    K = o.fault_stiffness_cols(cols.tolist(), nthreads=4)
    rep = o.report()
    assert rep["unconverged"] == 0
    np.savez_compressed(os.path.join(HERE, "c3_columns.npz"), cols=cols, K=K, iterations=np.int64(rep["iterations"]))
    print(f"wrote c3_columns.npz: columns {cols.tolist()}, {rep['iterations']} CG iterations")


if __name__ == "__main__":
    main()
