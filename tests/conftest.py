"""Shared test setup: import paths for the oracle and the eqd host package, the gpu marker."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "earthquakediffinitive.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device (MI355X)")


import pytest  # noqa: E402


@pytest.fixture(scope="session")
def oracle_exact_dz50_lf400():
    """The oracle's exact K at dz = 50 m, L = (800, 400), l_f = 400 (578 columns, ~25 s on 8
    threads) and its CG statistics; shared by every pin on that build."""
    from oracle.oracle import OracleSystem, bp8_geometry
    o = OracleSystem(bp8_geometry(50.0, 800.0, 400.0, 400.0))
    K = o.fault_stiffness(nthreads=min(8, os.cpu_count() or 1))
    return K, o.report()
