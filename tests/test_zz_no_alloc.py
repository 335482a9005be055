"""eqd_shear_traction_noalloc (the forward solve whose Krylov workspace is a static device array)
through the C ABI, against eqd_shear_traction on solvers with the same options: equal results
(the same kernels in the same order: bitwise) on two calls in a row, the second on a smaller
grid, in FP64 and mixed precision; the scratch array keeps the last call's contents, so the
second call shows that nothing stale is read. A grid whose workspace exceeds the array is
refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

import eqd
from eqd import _lib

pytestmark = pytest.mark.gpu

MU = 32.04e9


def slip(nf, seed):
    rng = np.random.default_rng(seed)
    return 1e-3 * rng.standard_normal(nf), 1e-3 * rng.standard_normal(nf)


def noalloc(fe, s2, s3):
    nf = fe.op.nf
    d2, d3 = np.zeros(nf), np.zeros(nf)
    rc = _lib.lib().eqd_shear_traction_noalloc(fe.rs.handle, np.ascontiguousarray(s2), np.ascontiguousarray(s3),
                                               d2, d3)
    _lib.check(rc)
    return d2, d3


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_matches_shear_traction_two_calls(prec):
    print(f"[no-alloc forward solve {prec}]", flush=True)
    cases = [(eqd.bp8_grid_pair(50.0, 800.0, 400.0), MU, 400.0, 1),   # 43,245 stored unknowns
             (eqd.GridPair(11, 11, 1.0, 1.0), 2.0, 0.4, 2)]            # a smaller grid second
    for g, mu, lf, seed in cases:
        fe_ref = eqd.FaultElasticity(g, mu, mu, l_f=lf, precision=prec)
        fe_na = eqd.FaultElasticity(g, mu, mu, l_f=lf, precision=prec)
        s2, s3 = slip(fe_ref.op.nf, seed)
        r2, r3 = eqd.shear_traction(fe_ref, s2, s3)
        a2, a3 = noalloc(fe_na, s2, s3)
        print(f"  nf={fe_ref.op.nf}: max |diff| {max(np.abs(a2 - r2).max(), np.abs(a3 - r3).max()):.3e}", flush=True)
        assert np.array_equal(a2, r2) and np.array_equal(a3, r3)
        ra, rb = eqd.solver_report(fe_ref.rs), eqd.solver_report(fe_na.rs)
        assert ra["iterations"] == rb["iterations"] and rb["solves"] == 1 and rb["unconverged"] == 0
        # the solver's own workspace is untouched: a regular solve afterwards still agrees
        b2, b3 = eqd.shear_traction(fe_na, s2, s3)
        assert np.array_equal(b2, r2) and np.array_equal(b3, r3)


def test_noalloc_refuses_oversized_grid():
    print("[no-alloc forward solve: dz = 25 m grid refused]", flush=True)
    fe = eqd.FaultElasticity(eqd.bp8_grid_pair(25.0, 1600.0, 800.0), MU, MU, l_f=1600.0, precision="fp64")
    nf = fe.op.nf
    with pytest.raises(eqd.EqdError) as e:
        noalloc(fe, np.zeros(nf), np.zeros(nf))
    assert e.value.code == _lib.EQD_ERR_RANGE and "static scratch holds" in str(e.value)
