"""The reference's published domain-sweep scalars, end to end on the device.

scripts/bp8_domain_convergence.jl:61-80 builds each model with stiffness=:exact at
dz = 50 m, l_f = 400 m, runs run_bp8 for 100 h with saveat = 3600 s and prints K_self =
K[ic, ic], the centre slip and V_max = max |V| of evaluate! at the last save. PROGRESS.md
tabulates the results for L_normal in 400..1600 at L_fault = 800 (:584-593) and for L_fault
in 800..1600 at L_normal = 1200 (:617-621); tests/golden/reference_scalars.json holds them.
Here the K build (mixed-precision batched CG) and the 100 h Tsit5 loop run on the device and
must reproduce every printed digit of K_self and the centre slip; V_max, which the adaptive
step sequence amplifies (exp(tau / (a sigma))), within one unit of its sixth digit. The
grids cover every pencil shape of the stencil: one x1 segment (L_normal = 400) and two to
four segments with x1 halos (L_normal >= 600).
"""
import json
import os

import pytest

from eqd import bp8

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_scalars.json")))


def _rows():
    out = []
    for r in GOLD["coupled_100h_gs_exact_dz50_lf400"]["rows"]:
        out.append(pytest.param(800.0, float(r["L_normal"]), r, id=f"Lf800_Ln{r['L_normal']}"))
    for r in GOLD["coupled_100h_gs_exact_dz50_lf400_Lnormal1200"]["rows"]:
        if r["L_fault"] != 800:  # (800, 1200) is already a row of the first sweep
            out.append(pytest.param(float(r["L_fault"]), 1200.0, r, id=f"Lf{r['L_fault']}_Ln1200"))
    return out


def _printed_equal(v, ref, ulp_last=0.0):
    """v printed as %.5e equals ref, or lies within ulp_last units of its last digit."""
    if f"{v:.5e}" == f"{ref:.5e}":
        return True
    unit = 10.0 ** (int(f"{ref:.5e}".split("e")[1]) - 5)
    return abs(v - ref) <= (ulp_last + 0.5) * unit if ulp_last else False


@pytest.mark.parametrize("L_fault,L_normal,row", _rows())
def test_domain_sweep_row(L_fault, L_normal, row):
    print(f"[published 100 h row L_fault={L_fault} L_normal={L_normal}]", flush=True)
    m = bp8.build_model(dz=50.0, L_fault=L_fault, L_normal=L_normal, stiffness="exact", precision="mixed")
    assert m.grid_info["elastic_dofs"] == row["dof"]
    nf = m.nf
    ic = (m.n2f // 2) * (m.n2f + 1)
    K = m.K
    assert _printed_equal(K[ic, ic], row["K_self"]), K[ic, ic]
    sol = bp8.run_bp8(m, tspan=(0.0, 100 * 3600.0), saveat=3600.0)
    u = sol.u[-1]
    c = bp8.evaluate(m, u, sol.t[-1])
    assert _printed_equal(u[ic], row["slip_00"]), u[ic]
    assert _printed_equal(c["Vmag"].max(), row["V_max"], ulp_last=1.0), c["Vmag"].max()
    # p(0,0) = 13.16528 MPa on every row (bit-identical in the reference): the pressure does
    # not see the elastic domain; two units of the last digit as for the production row.
    assert abs(u[3 * nf + ic] / 1e6 - GOLD["coupled_100h_gs_exact_dz50_lf400"]["p_00_MPa"]) <= 1.5e-5
