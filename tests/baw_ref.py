"""numpy restatement of the surrogate's analytic baseline premium, used to check the
device functions the surrogate's target is taken relative to.

baw_premium: the Barone-Adesi & Whaley (1987) early-exercise premium with a continuous
dividend yield. The critical spot by bisection (64 halvings in log S over [K/1000, K]
for puts, [K, 10^4 K] for calls); the device solves the same equation over the same
bracket by the Illinois method, so the two agree only if both find its root.

baseline_premium: Ju & Zhong's (1999) correction of it, the baseline the device uses:
the continuation-side premium A (S/S*)^lambda becomes A (S/S*)^lambda / (1 - chi) with
chi = b log(S/S*)^2 + c log(S/S*) (1 + chi where chi > 0), S* unchanged."""
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


def _m(o):
    """alpha / h = 2r / (sigma^2 (1 - e^{-rT})), with its limit 2 / (sigma^2 T) at r = 0."""
    r, t, z2 = o["r"], o["t"], o["z"] ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(r > 0, 2 * r / (z2 * -np.expm1(-r * t)), 2 / (z2 * t))


def critical_spot(o):
    k, t, z, r, q = o["k"], o["t"], o["z"], o["r"], o["q"]
    w = np.where(o["w"] < 0, -1.0, 1.0)
    nn = 2 * (r - q) / (z * z)
    qw = 0.5 * (-(nn - 1) + w * np.sqrt((nn - 1) ** 2 + 4 * _m(o)))
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


def _premium(o, chi_of=None):
    w = np.where(o["w"] < 0, -1.0, 1.0)
    none = ((w > 0) & ~(o["q"] > 0)) | ((w < 0) & ~(o["r"] > 0))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        ss, qw = critical_spot(o)
        _, nd1s = european(o, ss)
        e, _ = european(o)
        a = w * (ss / qw) * (1 - np.exp(-o["q"] * o["t"]) * nd1s)
        x = np.log(o["s"] / ss)
        f = 1.0
        if chi_of is not None:
            chi = chi_of(o, ss, a, w, x)
            f = np.where(chi > 0, 1 + chi, 1 / (1 - chi))
        cont = np.where(w < 0, o["s"] > ss, o["s"] < ss)
        prem = np.where(cont, a * (o["s"] / ss) ** qw * f, w * (o["s"] - o["k"]) - e)
    return np.where(none, 0.0, prem)


def baw_premium(o):
    return _premium(o)


def _jz_chi(o, ss, a, w, x):
    """Ju & Zhong's chi at x = log(S/S*): b x^2 + c x (0 where b or c is not finite or
    A <= 0)."""
    k, t, z, r, q = o["k"], o["t"], o["z"], o["r"], o["q"]
    z2 = z * z
    h = np.where(r > 0, -np.expm1(-r * t), 0.0)
    m = _m(o)
    beta = 2 * (r - q) / z2
    dd = (beta - 1) ** 2 + 4 * m
    sd = np.sqrt(dd)
    sq = z * np.sqrt(t)
    d1 = (np.log(ss / k) + (r - q + 0.5 * z2) * t) / sq
    dens = np.exp(-0.5 * d1 * d1) / np.sqrt(2 * np.pi)
    det = (ss * np.exp(-q * t) * dens * z / (2 * np.sqrt(t)) - w * q * ss * np.exp(-q * t) * ndtr(w * d1)
           + w * r * k * np.exp(-r * t) * ndtr(w * (d1 - sq)))
    h_r = np.where(r > 0, h / np.where(r > 0, r, 1.0), t)
    b = -(1 - h) * m * m / (2 * dd)
    c = -(m * h_r * det / (w * sd * a) + (1 - h) * m / (w * sd) - (1 - h) * w * m * m / (dd * sd))
    ok = (a > 0) & np.isfinite(b) & np.isfinite(c)
    return np.where(ok, (b * x + c) * x, 0.0)


def baseline_premium(o):
    return _premium(o, _jz_chi)
