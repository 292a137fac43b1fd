"""The surrogate gate's spread calibration (host logic; no GPU)."""
from __future__ import annotations

import numpy as np

from kwgpu import pipeline


def _held_out(n, seed):
    """Errors that grow with the spread, plus a few outliers at small spread."""
    rng = np.random.default_rng(seed)
    spread = rng.uniform(0.0, 1e-4, n)
    err = spread * rng.normal(0.0, 0.5, n)
    return err, spread


def test_bound_keeps_every_calibration_error_under_half_the_bar():
    err, spread = _held_out(200_000, 1)
    thr = pipeline.calibrate_spread(err, spread, bar=1e-5, cap=1e-4, margin=0.5)
    under = spread <= thr
    assert under.sum() >= pipeline.MIN_CALIBRATED
    assert np.max(np.abs(err[under])) <= 0.5e-5
    # the largest such bound: the next option by spread breaks the margin or the cap
    nxt = np.min(spread[spread > thr])
    assert nxt > 1e-4 or np.max(np.abs(err[spread <= nxt])) > 0.5e-5


def test_too_few_calibration_options_serve_nothing():
    err, spread = _held_out(200_000, 2)
    # one tiny-spread option with a tiny error, then a large error right after it
    i = np.argsort(spread)
    err[i[1]] = 1.0
    assert pipeline.calibrate_spread(err, spread, bar=1e-4, cap=1e-4) == -1.0
    # the same data with the count requirement waived gives the one-option bound
    assert pipeline.calibrate_spread(err, spread, bar=1e-4, cap=1e-4, min_count=1) == spread[i[0]]


def test_cap_limits_the_bound():
    err = np.zeros(5000)
    spread = np.linspace(0.0, 1.0, 5000)
    thr = pipeline.calibrate_spread(err, spread, bar=1e-4, cap=0.5)
    assert thr <= 0.5 and thr == np.max(spread[spread <= 0.5])
