"""Argument handling of the run_bp8 command-line driver (scripts/run_bp8.jl:14-20): positional
injection, dz, L_fault, L_normal with the reference's defaults and output directory name."""
import os

import pytest

from eqd import run_bp8 as cli


def test_defaults_are_the_reference_scripts():
    a = cli.parse_args([])
    assert (a.injection, a.dz, a.L_fault, a.L_normal) == ("gaussian", 50.0, 800.0, 400.0)
    assert a.alg == "tsit5" and a.tf_days is None and a.outdir is None
    assert cli.default_outdir(a.injection, a.dz, a.L_fault, a.L_normal) == os.path.join(
        "output", "BP8-QD-GS_dz50_Lf800_Ln400")


@pytest.mark.parametrize("word,inj", [("pw", "peaceman"), ("PW", "peaceman"), ("Pw", "peaceman"),
                                      ("gs", "gaussian"), ("GS", "gaussian"), ("anything", "gaussian")])
def test_injection_word(word, inj):
    # scripts/run_bp8.jl:14: lowercase(ARGS[1]) == "pw" selects the well, anything else the source
    assert cli.parse_args([word]).injection == inj


def test_positional_grid_and_outdir_name():
    a = cli.parse_args(["pw", "25", "1600", "800"])
    assert (a.injection, a.dz, a.L_fault, a.L_normal) == ("peaceman", 25.0, 1600.0, 800.0)
    assert cli.default_outdir(a.injection, a.dz, a.L_fault, a.L_normal).endswith("BP8-QD-PW_dz25_Lf1600_Ln800")
    a = cli.parse_args(["gs", "100", "800", "800", "--alg", "imex", "--tf_days", "2", "--outdir", "x"])
    assert (a.alg, a.tf_days, a.outdir) == ("imex", 2.0, "x")
    with pytest.raises(SystemExit):
        cli.parse_args(["gs", "fifty"])
