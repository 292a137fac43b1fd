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


def test_scale_cap_excludes_shared_biases():
    """Where the baseline scale is large the members share a bias that the spread does
    not show: the gate's scale cap must leave those options out, and the cap chosen is
    the one serving the most held-out options within half the bar."""
    rng = np.random.default_rng(3)
    n = 100_000
    scale = rng.uniform(0.0, 0.2, n)
    spread = rng.uniform(0.0, 5e-5, n)
    err = spread * rng.normal(0.0, 0.4, n)
    biased = scale > 0.05
    err[biased] += 4e-3 * scale[biased] * rng.choice([-1.0, 1.0], biased.sum())
    thr, cap = pipeline.calibrate_gate(err, spread, scale, bar=1e-4, cap=5e-5)
    assert 0 < thr <= 5e-5 and cap <= 0.05
    served = (spread <= thr) & (scale <= cap)
    assert served.sum() >= pipeline.MIN_CALIBRATED and not np.any(served & biased)
    assert np.max(np.abs(err[served])) <= 0.5e-4
    # without the cap a small spread would admit biased options beyond the margin
    thr_all = pipeline.calibrate_spread(err, spread, bar=1e-4, cap=5e-5)
    assert thr_all < thr


def test_no_cap_with_enough_options_serves_nothing():
    err = np.full(500, 1.0)
    spread = np.zeros(500)
    scale = np.zeros(500)
    assert pipeline.calibrate_gate(err, spread, scale) == (-1.0, -1.0)


def test_inference_cost_picks_the_cap_that_saves_most():
    """Each candidate (option under the scale cap) costs `cost` FD solves of inference and
    each option served saves one. Here every cap serves its options with spread under
    the bound, but the wider caps add candidates that are seldom served: a cost makes the
    narrow cap win, and a cost above the served share turns serving off."""
    rng = np.random.default_rng(4)
    n = 200_000
    scale = rng.uniform(0.0, 0.5, n)  # beyond the largest finite cap
    # options above 0.02 have large spreads: candidates there are hardly ever served
    spread = np.where(scale <= 0.02, rng.uniform(0.0, 1e-5, n), rng.uniform(0.0, 1e-3, n))
    err = np.zeros(n)
    thr0, cap0 = pipeline.calibrate_gate(err, spread, scale, bar=1e-4, cap=5e-5)
    assert cap0 == float("inf")  # serving the most: no cap
    thr1, cap1 = pipeline.calibrate_gate(err, spread, scale, bar=1e-4, cap=5e-5, cost=0.1)
    assert cap1 == 0.02 and thr1 > 0
    served = np.sum((scale <= cap1) & (spread <= thr1))
    # the saving of the chosen cap is positive and beats the uncapped one's
    assert served - 0.1 * np.sum(scale <= cap1) > (
        np.sum((spread <= thr0)) - 0.1 * n)
    assert pipeline.calibrate_gate(err, spread, scale, bar=1e-4, cap=5e-5, cost=2.0) == (-1.0, -1.0)


def test_small_saving_serves_nothing():
    """A gate that would save less than min_gain of the FD solves serves nothing: a small
    served share is not worth the error tail a batch can show beyond the calibration."""
    rng = np.random.default_rng(5)
    n = 200_000
    scale = rng.uniform(0.0, 0.2, n)
    spread = rng.uniform(0.0, 1e-3, n)  # about 1 % of options under a 1e-5 spread bound
    err = np.zeros(n)
    thr, cap = pipeline.calibrate_gate(err, spread, scale, bar=1e-4, cap=1e-5)
    assert thr > 0
    served = np.mean((spread <= thr) & (scale <= cap))
    assert 0.005 < served < 0.02
    assert pipeline.calibrate_gate(err, spread, scale, bar=1e-4, cap=1e-5,
                                   min_gain=2 * served) == (-1.0, -1.0)
    assert pipeline.calibrate_gate(err, spread, scale, bar=1e-4, cap=1e-5,
                                   min_gain=served / 2)[0] == thr


def test_classes_calibrate_apart():
    """Two classes, one with errors ten times the other's: the one gate is held to the
    noisy class's errors, while calibrated apart the quiet class serves far more. Every
    class keeps every calibration error it serves within half the bar, a class with too
    few held-out options stays closed, and a total saving under min_gain closes all."""
    rng = np.random.default_rng(6)
    n = 400_000
    cls = rng.integers(0, 2, n)  # class 0 quiet, 1 noisy; classes 2.. are empty
    spread = rng.uniform(0.0, 5e-5, n)
    err = spread * rng.normal(0.0, 0.2, n) * np.where(cls == 0, 1.0, 10.0)
    scale = np.zeros(n)
    g = pipeline.calibrate_classes(err, spread, scale, cls, t_edges=(0.5,), m_edges=(),
                                   bar=1e-4, cap=5e-5, min_count=1000)
    assert len(g["spread_max"]) == 4 and g["spread_max"][2:] == [-1.0, -1.0]
    b0, b1 = g["spread_max"][:2]
    assert b0 > b1 > 0
    for c, b in ((0, b0), (1, b1)):
        m = (cls == c) & (spread <= b)
        assert m.sum() >= 1000 and np.abs(err[m]).max() <= 0.5e-4
    one, _ = pipeline.calibrate_gate(err, spread, scale, bar=1e-4, cap=5e-5)
    assert one <= b1 < b0
    served = np.mean(spread <= np.where(cls == 0, b0, b1))
    assert served > 1.5 * np.mean(spread <= one)
    closed = pipeline.calibrate_classes(err, spread, scale, cls, t_edges=(0.5,), m_edges=(),
                                        bar=1e-4, cap=5e-5, min_count=1000, min_gain=2 * served)
    assert max(closed["spread_max"]) == -1.0


def test_gate_mask_follows_the_class_bounds():
    import kwgpu
    o = np.zeros(6, dtype=kwgpu.OPTION_DTYPE)
    o["k"], o["s"] = 100.0, [90.0, 110.0, 90.0, 110.0, 100.0, 100.0]
    o["t"], o["w"] = [0.2, 0.2, 3.0, 3.0, 1.0, 1.0], [-1, -1, 1, 1, -1, 1]
    g = {"t_edges": [0.5, 2.0], "m_edges": [0.0], "spread_max": [1e-5] * 12,
         "scale_max": [0.05] * 12}
    g["spread_max"][(1 * 3 + 2) * 2 + 1] = -1.0  # calls, T >= 2, S >= K: closed
    spread = np.full(6, 5e-6)
    scale = np.full(6, 0.01)
    assert pipeline.gate_mask(g, o, spread, scale).tolist() == [True, True, True, False, True, True]
    scale[0] = 0.1
    assert not pipeline.gate_mask(g, o, spread, scale)[0]
