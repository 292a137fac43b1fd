"""numpy restatement of the Barone-Adesi & Whaley (1987) early-exercise premium with a
continuous dividend yield, used to check the device function the surrogate's target is
taken relative to. The critical spot by bisection (64 halvings in log S over [K/1000, K]
for puts, [K, 10^4 K] for calls); the device solves the same equation over the same
bracket by the Illinois method, so the two agree only if both find its root."""
from __future__ import annotations

import numpy as np
from scipy.special import ndtr


def european(o, s=None):
    """(price, Phi(w d1)) of the European option at spot s (default o["s"])."""
    s = o["s"] if s is None else s
    k, t, z, r, q = o["k"], o["t"], o["z"], o["r"], o["q"]
    w = np.where(o["w"] < 0, -1.0, 1.0)
    sq = z * np.sqrt(t)
    d1 = (np.log(s / k) + (r - q + 0.5 * z * z) * t) / sq
    nd1 = ndtr(w * d1)
    return w * (s * np.exp(-q * t) * nd1 - k * np.exp(-r * t) * ndtr(w * (d1 - sq))), nd1


def critical_spot(o):
    k, t, z, r, q = o["k"], o["t"], o["z"], o["r"], o["q"]
    w = np.where(o["w"] < 0, -1.0, 1.0)
    z2 = z * z
    with np.errstate(divide="ignore", invalid="ignore"):
        mk = np.where(r > 0, 2 * r / (z2 * -np.expm1(-r * t)), 2 / (z2 * t))
    nn = 2 * (r - q) / z2
    qw = 0.5 * (-(nn - 1) + w * np.sqrt((nn - 1) ** 2 + 4 * mk))
    eq = np.exp(-q * t)
    lo = np.where(w < 0, np.log(k * 1e-3), np.log(k))
    hi = np.where(w < 0, np.log(k), np.log(k * 1e4))
    for _ in range(64):
        mid = 0.5 * (lo + hi)
        s = np.exp(mid)
        e, nd1 = european(o, s)
        g = w * (s - k) - e - w * (1 - eq * nd1) * s / qw
        below = np.where(w < 0, g > 0, g < 0)
        lo, hi = np.where(below, mid, lo), np.where(below, hi, mid)
    return np.exp(0.5 * (lo + hi)), qw


def baw_premium(o):
    w = np.where(o["w"] < 0, -1.0, 1.0)
    none = ((w > 0) & ~(o["q"] > 0)) | ((w < 0) & ~(o["r"] > 0))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        ss, qw = critical_spot(o)
        _, nd1s = european(o, ss)
        e, _ = european(o)
        a = w * (ss / qw) * (1 - np.exp(-o["q"] * o["t"]) * nd1s)
        cont = np.where(w < 0, o["s"] > ss, o["s"] < ss)
        prem = np.where(cont, a * (o["s"] / ss) ** qw, w * (o["s"] - o["k"]) - e)
    return np.where(none, 0.0, prem)
