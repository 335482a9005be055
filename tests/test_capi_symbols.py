"""libeqd.so loads without a device and exports every symbol include/eqd.h declares."""
import ctypes
import os
import re

import pytest

from eqd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "eqd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(eqd_[A-Za-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing
    assert sorted(_lib.EXPORTED) == syms


def test_version_and_host_only_entry_points():
    L = _lib.lib()
    assert b"gfx950" in L.eqd_version()
    g = _lib.GridDesc(9, 9, 1.0, 1.0, 2.0, 1.0, 0.5)
    start, width, coef = _lib.tables_export(g, "AXs")
    assert len(start) == 2 * 9 - 3
    # grid legality is checked before any device call (src/BP8.jl:199-204)
    bad = _lib.GridDesc(5, 17, 400.0, 800.0, 1.0, 1.0, 800.0)
    assert L.eqd_tables_export(ctypes.byref(bad), 0, None, None, None) == -1
    buf = ctypes.create_string_buffer(256)
    L.eqd_last_error(buf, 256)
    assert b"at least 9 points" in buf.value


def test_operator_create_rejects_bad_patch_before_touching_device():
    import eqd
    with pytest.raises(eqd.EqdError) as e:
        eqd.FaultElasticity(eqd.GridPair(13, 13, 1.2, 1.2), 2.0, 1.0, l_f=0.55)
    assert e.value.code == _lib.EQD_ERR_GRID
    with pytest.raises(eqd.EqdError) as e:
        eqd.bp8_grid_pair(100.0, 800.0, 400.0)   # dz = 100 m, L = (800, 400): n1 = 5 < 9
    assert e.value.code == _lib.EQD_ERR_GRID


def test_operator_create_rejects_grid_beyond_tile_offsets():
    """A 64-column FP64 tile vector is addressed with 32-bit byte offsets: a grid whose tile would
    span 2^31 B or more is refused with EQD_ERR_GRID before any device call (dz = 20 m on the
    L = (1600, 800) domain: 1,997,199 stored nodes, 3.07 GB per tile); dz = 25 m there (1,016,127
    stored nodes, 1.56 GB) is below the limit and gets as far as the device check."""
    import eqd
    with pytest.raises(eqd.EqdError) as e:
        eqd.FaultElasticity(eqd.bp8_grid_pair(20.0, 1600.0, 800.0), 32.04e9, 32.04e9, l_f=1600.0)
    assert e.value.code == _lib.EQD_ERR_GRID and "grid too large" in str(e.value)
    try:
        eqd.FaultElasticity(eqd.bp8_grid_pair(25.0, 1600.0, 800.0), 32.04e9, 32.04e9, l_f=1600.0)
    except eqd.EqdError as e2:
        assert e2.code == _lib.EQD_ERR_DEVICE  # no GPU in this process
