"""Pins of the field schedule in the oracle (fo_wave: wave_module.F:8-67 and the
hysteresis-loop sweep of hysteresis_loop.F:47-49).

The sweep is pinned by the reference's own output: every row of
22example-BaTiO3-new-param-hl/360.hl prints the step and the E_field_add that
step used (tests/golden/bto32_hysteresis_360K.json).  The waves have no
shipped output, so they are pinned by their closed forms: the triangular
cosine starts at +A, crosses zero at a quarter period and reaches -A at half
a period; the sine is the cosine a quarter period later; both are periodic
and bounded by |A|; the ramps are linear to and from the amplitude.  A
planted off-by-one step must fail the .hl pin.
"""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def hl():
    return json.load(open(os.path.join(HERE, "golden", "bto32_hysteresis_360K.json")))


def test_hysteresis_sweep_matches_reference_hl_output():
    d = hl()
    p = d["param"]
    A, n, na = p["external_E_field"], p["n_thermalize"], p["n_average"]
    for row in d["rows"]:
        E = oracle.wave("hysteresis", A, 0, n, na, row[0])
        assert np.allclose(np.round(E, 12), row[2:5], rtol=0, atol=5e-13), row[:5]
    # before the sweep the field is the amplitude (initialize_E_field.F:58-66)
    assert np.array_equal(oracle.wave("hysteresis", A, 0, n, na, 5), A)


def test_planted_off_by_one_fails():
    d = hl()
    p = d["param"]
    A, n, na = p["external_E_field"], p["n_thermalize"], p["n_average"]
    bad = 0
    for row in d["rows"][1:]:
        E = oracle.wave("hysteresis", A, 0, n, na, row[0] - 1000)
        bad += not np.allclose(np.round(E, 12), row[2:5], rtol=0, atol=5e-13)
    assert bad == len(d["rows"]) - 1


@pytest.mark.parametrize("period", [4, 40, 1000])
def test_triangular_closed_forms(period):
    A = np.array([0.001, -0.002, 0.003])
    n = 7
    c = lambda i: oracle.wave("triangular_cos", A, period, n, 0, i)
    s = lambda i: oracle.wave("triangular_sin", A, period, n, 0, i)
    assert np.allclose(c(n), A) and np.allclose(c(n + period // 2), -A)
    assert np.allclose(c(n + period // 4), 0) and np.allclose(c(n + 3 * period // 4), 0)
    assert np.allclose(s(n), 0) and np.allclose(s(n + period // 4), A) and np.allclose(s(n + 3 * period // 4), -A)
    for i in range(n, n + 2 * period):
        assert np.allclose(s(i), c(i + 3 * period // 4))
        assert np.allclose(c(i), c(i + period))
        assert np.all(np.abs(c(i)) <= np.abs(A) + 1e-18)
    # before n_thermalize the wave holds its value at i = 0
    assert np.array_equal(c(0), c(n))


def test_ramps_and_zero_amplitude():
    A = [0.0, 0.0, 0.002]
    for i in (0, 10, 25, 50, 80):
        on = oracle.wave("ramping_on", A, 50, 100, 0, i)
        off = oracle.wave("ramping_off", A, 50, 100, 0, i)
        assert on[2] == pytest.approx(0.002 * min(i, 50) / 50, abs=1e-18)
        assert off[2] == pytest.approx(0.002 * max(50 - i, 0) / 50, abs=1e-18)
    for kind in oracle.WAVES:
        assert np.array_equal(oracle.wave(kind, [0, 0, 0], 8, 3, 10, 5), [0, 0, 0])
