"""Host-side pieces of the BP8 boundary that need no device: parameters and the
build_model legality errors (src/BP8.jl:189-204), raised before any device call."""
import numpy as np
import pytest

import eqd
from eqd import _lib, bp8
from oracle.oracle import bp8_default_params


def test_default_parameters_match_table1_and_oracle():
    p = bp8.benchmark_parameters()
    assert p == bp8_default_params()
    # Table 1 spot checks (src/BP8.jl:50-87)
    assert p["mu"] == 32.04e9 and p["nu"] == 0.25 and p["a"] == 0.016 and p["b"] == 0.010
    assert p["sigma0"] == 25e6 and p["tau_init"] == 14.6e6 and p["t_off"] == 100 * 3600.0
    assert bp8.lame_lambda(p) == pytest.approx(p["mu"])


def test_unknown_parameter_rejected():
    with pytest.raises(eqd.EqdError):
        bp8.benchmark_parameters(not_a_field=1.0)


@pytest.mark.parametrize("kw,code", [
    (dict(dz=50.0, L_fault=800.0, L_normal=400.0, injection="point"), _lib.EQD_ERR_ARG),
    (dict(dz=30.0, L_fault=800.0, L_normal=400.0), _lib.EQD_ERR_GRID),     # 30 does not divide 400
    (dict(dz=50.0, L_fault=300.0, L_normal=400.0), _lib.EQD_ERR_GRID),     # L_fault < l_f
    (dict(dz=100.0, L_fault=800.0, L_normal=400.0), _lib.EQD_ERR_GRID),    # n1 = 5 < 9
    (dict(dz=50.0, L_fault=800.0, L_normal=400.0, stiffness="dense"), _lib.EQD_ERR_ARG),
])
def test_build_model_legality(kw, code):
    with pytest.raises(eqd.EqdError) as e:
        bp8.build_model(**kw)
    assert e.value.code == code


def test_model_from_stiffness_checks_before_device():
    K = np.zeros((2 * 64, 2 * 64))
    with pytest.raises(eqd.EqdError) as e:   # Omega_f 8 x 8 < 9
        bp8.model_from_stiffness(K, 8)
    assert e.value.code == _lib.EQD_ERR_GRID
    with pytest.raises(eqd.EqdError) as e:
        bp8.model_from_stiffness(np.zeros((10, 10)), 17)
    assert e.value.code == _lib.EQD_ERR_ARG


def test_derived_parameters_match_table1():
    """test/bp8_test.jl:17-25: nu = 0.25 makes lambda = mu; c_s = sqrt(mu / rho) reproduces the
    tabulated 3.464 km/s; alpha = k / (phi beta eta) reproduces 0.05 m^2/s."""
    p = bp8.benchmark_parameters()
    assert bp8.lame_lambda(p) == pytest.approx(p["mu"], rel=1e-12)
    assert np.sqrt(p["mu"] / p["rho"]) == pytest.approx(p["c_s"], rel=1e-3)
    assert p["k"] / (p["phi"] * p["beta"] * p["viscosity"]) == pytest.approx(p["alpha"], rel=1e-12)
