"""The run_bp8 command-line driver end to end on the device (scripts/run_bp8.jl:1-79): build_model,
run_bp8 with saveat = 300 s, the summary and write_outputs, over one day at dz = 50 m.

The summary's figures are checked against the oracle's Tsit5 run of the same K over the same
day (peak slip rate, its time, final slip at the origin within the bars of tests/test_gpu_bp8.py),
and the output directory holds the 20 section 4 files write_outputs makes."""
import io
import os

import numpy as np
import pytest

from eqd import run_bp8 as cli
from eqd import bp8
from oracle.oracle import OracleBP8

pytestmark = pytest.mark.gpu


def test_cli_gs_one_day_matches_oracle(tmp_path):
    print("[run_bp8 CLI: gs 50 800 400, 1 day]", flush=True)
    out = io.StringIO()
    res = cli.main(["gs", "50", "800", "400", "--outdir", str(tmp_path / "gs"), "--tf_days", "1"], out=out)
    text = out.getvalue()
    print(text, flush=True)
    assert res["retcode"] == "Success" and res["nsaved"] == 289
    assert len(os.listdir(tmp_path / "gs")) == 9 + 1 + 10
    for line in ("peak slip rate", "final slip (0,0)", "final max slip", "peak pressure (0,0)", "resolution:",
                 "outputs written"):
        assert line in text
    # the same model again (the K build is deterministic) for its K and n2f; the oracle over one day
    m = bp8.build_model(dz=50.0, L_fault=800.0, L_normal=400.0)
    ro = OracleBP8(m.K, m.n2f, dict(dz=50.0)).run(0.0, 86400.0, 300.0)
    assert np.array_equal(ro["t"], np.arange(289) * 300.0)
    vpk = ro["vmax"].max()
    assert abs(res["vpeak"] - vpk) <= 1e-5 * vpk
    # the save the summary names holds the oracle's peak (to the same bar)
    assert ro["vmax"][int(round(res["tpeak"] / 300.0))] >= vpk * (1 - 1e-5)
    ic = (m.n2f // 2) * (m.n2f + 1)
    assert abs(res["final_slip_centre"] - ro["u"][-1, ic]) <= 1e-6 * abs(ro["u"][-1, ic])
    assert abs(res["ppeak"] - ro["u"][:, 3 * m.nf + ic].max()) <= 1e-6 * ro["u"][:, 3 * m.nf + ic].max()
    assert f"{res['vpeak']:.4E}" in text


def test_cli_pw_imex_writes_outputs(tmp_path):
    print("[run_bp8 CLI: pw 50 800 400 --alg imex, 1 day]", flush=True)
    out = io.StringIO()
    res = cli.main(["PW", "50", "800", "400", "--outdir", str(tmp_path / "pw"), "--tf_days", "1", "--alg", "imex"],
                   out=out)
    print(out.getvalue(), flush=True)
    assert res["retcode"] == "Success" and res["nsaved"] == 289
    assert len(os.listdir(tmp_path / "pw")) == 9 + 1 + 10
    # a day of injection at the Peaceman well raises the origin's pressure by megapascals
    assert res["ppeak"] > 1e6
    with open(tmp_path / "pw" / "global.dat") as f:
        head = f.read(2000)
    assert "# modeler=" in head
