"""The section 4 output files from device-evaluated saves (write_outputs, src/BP8.jl:605-774),
checked against the CPU oracle's evaluate! of the same saves.

Bars: every station row's slip, pressure and state columns equal the saved state to the printed
digits (7 significant); slip rates, shear stresses, max slip rate and moment rate within 1e-6
relative of the oracle (the Newton stops at |d ln V| < 1e-12; printing keeps 7 digits); Darcy
velocities within 1e-6 relative of -(k / viscosity) D1 p with the oracle's SBP-4 D1 (plus a
rounding floor of 1e-12 of the sum's terms)."""
import math
import os

import numpy as np
import pytest

from eqd import bp8
from oracle.oracle import OracleBP8, sbp1d

pytestmark = pytest.mark.gpu


def _data(path):
    rows = []
    with open(path) as f:
        lines = f.read().splitlines()
    start = lines.index("# Here is the time-series data.") + 1
    for line in lines[start:]:
        rows.append([float(v) for v in line.split()])
    return lines, np.array(rows)


def test_write_outputs_matches_oracle(tmp_path):
    print("[write_outputs 20 h dz=50 toeplitz]", flush=True)
    m = bp8.build_model(dz=50.0, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    saveat = 600.0
    sol = bp8.run_bp8(m, tspan=(0.0, 20 * 3600.0), saveat=saveat)
    out = str(tmp_path / "bp8")
    bp8.write_outputs(m, sol, out, modeler="test", profile_dt=3600.0)
    files = sorted(os.listdir(out))
    assert len(files) == 9 + 1 + 10
    nf, n2f = m.nf, m.n2f
    o = OracleBP8(m.K, n2f, dict(dz=50.0))
    ev = [o.evaluate(u, t) for u, t in zip(sol.u, sol.t)]
    _, _, w = m.operators()
    H, D1, _, _, _ = sbp1d(n2f, 50.0)
    kfac = -m.par["k"] / m.par["viscosity"]
    for name, sx2, sx3 in bp8.station_locations():
        lines, d = _data(os.path.join(out, name + ".dat"))
        assert lines[1] == "# problem=SEAS Benchmark BP8-QD-GS"
        assert any(l.startswith("# num_time_steps=") for l in lines)
        i = int(np.argmin(np.abs(m.x2 - sx2)))
        j = int(np.argmin(np.abs(m.x3 - sx3)))
        k = i + n2f * j
        assert d.shape == (len(sol.t), 11)
        np.testing.assert_allclose(d[:, 0], sol.t, rtol=1e-12)
        np.testing.assert_allclose(d[:, 1], sol.u[:, k], rtol=1e-6, atol=1e-30)
        np.testing.assert_allclose(d[:, 7], sol.u[:, 3 * nf + k] / 1e6, rtol=1e-6, atol=1e-12)
        np.testing.assert_allclose(d[:, 10], sol.u[:, 2 * nf + k] / math.log(10), rtol=1e-6)
        V2 = np.array([e["V2"][k] for e in ev])
        tau2 = np.array([e["tau2"][k] for e in ev])
        lv = np.log10(np.maximum(np.abs(V2), 1e-300))
        np.testing.assert_allclose(d[:, 3], lv, rtol=1e-6)
        np.testing.assert_allclose(d[:, 5], tau2 / 1e6, rtol=1e-6)
        P = sol.u[:, 3 * nf:4 * nf].reshape(len(sol.t), n2f, n2f, order="F")  # [save, a, b]
        q2 = kfac * np.einsum("j,sj->s", D1[i], P[:, :, j])
        q3 = kfac * np.einsum("j,sj->s", D1[j], P[:, i, :])
        # absolute floor: rounding of the D1 sum, whose terms are ~ |kfac| max|p| / dz (the centre
        # station's q vanishes by symmetry and the summation order sets the residue there)
        qat = 1e-12 * abs(kfac) * np.abs(P).max() / 50.0
        np.testing.assert_allclose(d[:, 8], q2, rtol=1e-6, atol=qat)
        np.testing.assert_allclose(d[:, 9], q3, rtol=1e-6, atol=qat)
    lines, g = _data(os.path.join(out, "global.dat"))
    vmax = np.array([e["Vmag"].max() for e in ev])
    mom = np.array([m.par["mu"] * np.sum(w * e["Vmag"]) for e in ev])
    np.testing.assert_allclose(g[:, 1], np.log10(vmax), rtol=1e-6)
    np.testing.assert_allclose(g[:, 2], mom, rtol=1e-6)
    # the same as the dedicated saved_vmax pass, to the printed digits of its log10
    np.testing.assert_allclose(g[:, 1], np.log10(bp8.saved_vmax(m, len(sol.t))), rtol=1e-6)
    with open(os.path.join(out, "shear_stress_2_strike.dat")) as f:
        plines = f.read().splitlines()
    data = [l for l in plines[plines.index("# Here are the data") + 1:]]
    first = np.array([float(v) for v in data[0].split()])
    np.testing.assert_allclose(first[2:], m.x2, rtol=1e-7)
    assert len(data) == 1 + 21  # 0 h .. 20 h every hour
    row = np.array([float(v) for v in data[-1].split()])
    j0 = int(np.argmin(np.abs(m.x3)))
    tau_line = ev[-1]["tau2"][np.arange(n2f) + n2f * j0] / 1e6
    np.testing.assert_allclose(row[2:], tau_line, rtol=1e-6)


def test_saved_outputs_rejects_bad_indices():
    m = bp8.build_model(dz=50.0, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    bp8.run_bp8(m, tspan=(0.0, 3600.0), saveat=600.0)
    with pytest.raises(Exception):
        bp8.saved_outputs(m, 2, [m.nf])
    with pytest.raises(Exception):
        bp8.saved_outputs(m, 100, [0])
