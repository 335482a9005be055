"""Host-side pieces of the section 4 output path (no GPU): station list, nearest-node rule,
profile times, and the header formatting helpers (src/BP8.jl:553-774)."""
import numpy as np

from eqd.outputs import _fmt_real, _profile_times, _safelog10, nearest_node, station_locations


def test_station_names_and_coordinates():
    # src/BP8.jl:553-561: fltst_strk<+-XXX>dp<+-XXX> for the nine points of BP8-QD section 4.1
    st = station_locations()
    assert [s[0] for s in st] == [
        "fltst_strk+000dp+000", "fltst_strk-200dp+000", "fltst_strk+000dp+200", "fltst_strk+200dp+000",
        "fltst_strk+000dp-200", "fltst_strk-200dp-200", "fltst_strk-200dp+200", "fltst_strk+200dp-200",
        "fltst_strk+200dp+200"]
    assert st[6][1:] == (-200.0, 200.0)


def test_nearest_node_takes_the_first_minimum():
    ax = np.array([-100.0, -50.0, 0.0, 50.0, 100.0])
    assert nearest_node(ax, 0.0) == 2
    assert nearest_node(ax, 25.0) == 2  # tie between 0 and 50: Julia's argmin keeps the first
    assert nearest_node(ax, 200.0) == 4


def test_profile_times_follow_the_range_rule():
    # collect(0.0:dt:t_end), then t_end appended when the range stops short (src/BP8.jl:705-707)
    assert _profile_times(7200.0, 3600.0) == [0.0, 3600.0, 7200.0]
    assert _profile_times(9000.0, 3600.0) == [0.0, 3600.0, 7200.0, 9000.0]
    assert _profile_times(0.0, 3600.0) == [0.0]


def test_formatting_helpers():
    assert _fmt_real(50.0) == "50.0" and _fmt_real(-400.0) == "-400.0"
    assert _safelog10(0.0) == -300.0
    assert abs(_safelog10(-1e-7) + 7.0) < 1e-12
