"""Simulation-directory files in the Java tool chain's formats (flowpro_v2_amd.results_io), CPU only.

java.lang.Double.toString known answers from its specification (JDK 19+ javadoc): plain notation
for 1e-3 <= |x| < 1e7 with at least one digit after the point, computerised scientific notation
otherwise; Double.MIN_VALUE is "4.9E-324" (the two-digit rule for one-digit shortest decimals),
Double.MAX_VALUE "1.7976931348623157E308".  Every written double parses back to itself.
"""
import math
import os

import numpy as np
import pytest

from flowpro_v2_amd import results_io as rio
from flowpro_v2_amd.solver import State


@pytest.mark.parametrize("x,ref", [
    (1.0, "1.0"), (0.1, "0.1"), (100.0, "100.0"), (1e7, "1.0E7"), (9999999.0, "9999999.0"), (1e-3, "0.001"),
    (9.999999999999998e-4, "9.999999999999998E-4"), (1e-4, "1.0E-4"), (123456789.0, "1.23456789E8"),
    (-0.0, "-0.0"), (0.0, "0.0"), (math.inf, "Infinity"), (-math.inf, "-Infinity"), (math.nan, "NaN"),
    (1.7976931348623157e308, "1.7976931348623157E308"), (5e-324, "4.9E-324"), (2.2250738585072014e-308,
    "2.2250738585072014E-308"), (-2.5, "-2.5"), (1234567.125, "1234567.125"), (1e21, "1.0E21"), (1e-5, "1.0E-5"),
    (3.0e10, "3.0E10"), (0.30000000000000004, "0.30000000000000004"), (1e23, "1.0E23"),
])
def test_double_to_string_known_answers(x, ref):
    assert rio.java_double(x) == ref


def test_double_to_string_round_trips():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.standard_normal(3000) * 10.0 ** rng.integers(-300, 300, 3000), rng.random(2000),
                         np.nextafter(0.0, 1.0) * rng.integers(1, 1000, 200)])
    for x in xs:
        s = rio.java_double(x)
        assert float(s) == x, s
        mant = s.split("E")[0]
        assert "." in mant and mant[-1] != "." and not mant.startswith(".")


def test_matrix_and_properties_round_trip(tmp_path):
    A = [np.array([1.0, -2.5e-7, 3.25]), np.array([4e12, 0.0])]
    p = str(tmp_path / "W.txt")
    rio.save_matrix(p, A)
    text = open(p).read().splitlines()
    assert text[0] == "1.0 -2.5E-7 3.25 " and text[1] == "4.0E12 0.0 "
    B = rio.load_matrix(p)
    assert all(np.array_equal(a, b) for a, b in zip(A, B))
    q = str(tmp_path / "x.txt")
    rio.save_properties(q, [("a", "1.0E-5"), ("key with:colon", "v=1 #2"), ("residuumNormalisation", "1.0,2.0")])
    lines = open(q).read().splitlines()
    assert lines[0].startswith("#")
    assert "key\\ with\\:colon=v\\=1 \\#2" in lines
    assert rio.load_properties(q) == {"a": "1.0E-5", "key with:colon": "v=1 #2", "residuumNormalisation": "1.0,2.0"}


def test_state_round_trip(tmp_path):
    st = State(cfl=5.0)
    st.t, st.steps, st.residuum, st.current_cfl, st.dt = 0.125, 17, 3.5e-4, 4.75, 0.0078125
    st.normalisation, st.n_norm_steps = [1.5, 2.0, 0.25], 5
    p = str(tmp_path / "state.txt")
    rio.save_state(p, st, order=3, cpu_ms=1234)
    props = rio.load_properties(p)
    assert props["t"] == "0.125" and props["steps"] == "17" and props["CPU"] == "1234"
    assert props["residuumNormalisation"] == "1.5,2.0,0.25" and props["order"] == "3"
    st2 = State(cfl=5.0)
    assert rio.load_state(p, st2) == 3
    assert (st2.t, st2.steps, st2.current_cfl, st2.normalisation, st2.n_norm_steps) == (0.125, 17, 4.75, [1.5, 2.0, 0.25], 5)


def test_append_residuum(tmp_path):
    p = str(tmp_path / "residuum.txt")
    rio.append_residuum(p, 1.0, 0.05, 1234.0, [1.0, 0.5, 2.5e-3])
    rio.append_residuum(p, 0.9, 0.1, 2468.0)
    assert open(p).read().splitlines() == ["1.0 0.05 1234.0 1.0 0.5 0.0025", "0.9 0.1 2468.0"]
