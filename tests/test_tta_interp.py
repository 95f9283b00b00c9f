"""The equal-error interpolation of the time-to-accuracy comparison
(corrector.interp_time_at_error): piecewise linear in log time against log
error, exact on power laws, the end point just beyond either end, None
further outside the measured range."""
import importlib

import numpy as np
import pytest

import problems_helper  # noqa: F401  (puts the repository on sys.path)

C = importlib.import_module("hpc-reading-course_amd.corrector")


def test_power_law_is_exact():
    # error = 1e-3 (t/0.1)^-2: a straight line in log-log
    pts = [(t, 1e-3 * (t / 0.1) ** -2) for t in (0.1, 0.3, 0.9)]
    for t in (0.15, 0.5, 0.8):
        assert C.interp_time_at_error(pts, 1e-3 * (t / 0.1) ** -2) == pytest.approx(t, rel=1e-12)


def test_brackets_and_outside():
    pts = [(0.172, 4.9e-4), (0.563, 3.2e-5), (0.30, 1.5e-4)]
    t = C.interp_time_at_error(pts, 7.65e-5)
    assert 0.30 < t < 0.563
    assert C.interp_time_at_error(pts, 1e-6) is None
    assert C.interp_time_at_error(pts, 1e-2) is None
    assert C.interp_time_at_error(pts, 4.9e-4) == pytest.approx(0.172)


def test_end_points_within_tolerance():
    """A target a hair beyond the curve's end (the same grid solved by the
    FP32 ensemble: 3.1965e-5 against the plain path's 3.1984e-5) takes the end
    point's time; 2% beyond it is outside the curve."""
    pts = [(0.179, 4.90e-4), (0.265, 1.03e-4), (0.569, 3.1984e-5)]
    assert C.interp_time_at_error(pts, 3.1965e-5) == pytest.approx(0.569)
    assert C.interp_time_at_error(pts, 4.93e-4) == pytest.approx(0.179)
    assert C.interp_time_at_error(pts, 3.1984e-5 * 0.98) is None
    assert C.interp_time_at_error(pts, 3.1965e-5, rtol=0.0) is None
