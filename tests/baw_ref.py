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


# ---- the control variate: the premium from the exercise boundary ------------------------
# The American put's exercise boundary b(tau) by Andersen, Lake & Offengenden's fixed point
# (form "A"), restated step for step as the device computes it: 6 Chebyshev-Lobatto nodes
# in sqrt(tau), 8 Jacobi steps from X (b_T / X)^{sqrt(tau / T)}, a 17-point tanh-sinh rule;
# calls by put-call symmetry (swap S and K, r and q).
TS_H = 0.375
_TS_K = np.arange(-8, 9) * TS_H
TS_A = 1.0 / (1.0 + np.exp(-np.pi * np.sinh(_TS_K)))            # (1 + tanh(pi/2 sinh kh)) / 2
TS_W = 0.25 * np.pi * TS_H * np.cosh(_TS_K) / np.cosh(0.5 * np.pi * np.sinh(_TS_K)) ** 2
BN, BIT = 6, 8


def _ncdf(x):
    return ndtr(x)


def _cheb_fit(H):
    i = np.arange(BN + 1)
    c = np.cos(np.outer(i, i) * np.pi / BN)
    w = np.ones(BN + 1)
    w[0] = w[-1] = 0.5
    a = (2.0 / BN) * (H * w) @ c.T
    a[..., 0] *= 0.5
    a[..., -1] *= 0.5
    return a


def _b_of(a, x0, st, xi):
    """b at xi = sqrt(u): X exp(-sqrt(H)), H by Clenshaw over z = 1 - 2 xi / sqrt T."""
    z = 1.0 - 2.0 * xi / st
    b1 = np.zeros_like(z)
    b2 = np.zeros_like(z)
    for k in range(BN, 0, -1):
        b1, b2 = 2 * z * b1 - b2 + a[..., k], b1
    h = z * b1 - b2 + a[..., 0]
    return x0 * np.exp(-np.sqrt(np.maximum(h, 0.0)))


def boundary_premium(o):
    """Early-exercise premium from the exercise boundary, per option (price units)."""
    o = np.atleast_1d(o)
    call = ~(o["w"] < 0)
    r = np.where(call, o["q"], o["r"]).astype(float)
    q = np.where(call, o["r"], o["q"]).astype(float)
    s, T = o["z"].astype(float), o["t"].astype(float)
    none = ~((o["k"] > 0) & (T > 0) & (s > 0)) | ~(r > 0)
    out = np.zeros(len(o))
    for n in np.nonzero(~none)[0]:
        out[n] = _premium_one(float(r[n]), float(q[n]), float(s[n]), float(T[n]), bool(call[n]),
                              float(o["k"][n]), float(o["s"][n]))
    return out


def _premium_one(r, q, s, T, call, K, S):
    st = np.sqrt(T)
    x0 = min(1.0, r / q) if q > 0 else 1.0
    u = np.zeros(1, dtype=[("t", "<f8"), ("k", "<f8"), ("z", "<f8"), ("r", "<f8"), ("q", "<f8"),
                           ("s", "<f8"), ("e", "u1"), ("w", "i1")])
    u["t"], u["k"], u["z"], u["r"], u["q"], u["s"], u["e"], u["w"] = T, 1.0, s, r, q, 1.0, 1, -1
    bT = min(float(critical_spot(u)[0][0]), x0)
    f = 0.5 * (1.0 - np.cos(np.arange(BN + 1) * np.pi / BN))
    tau = T * f * f
    b = x0 * (bT / x0) ** f
    b[0] = x0
    for _ in range(BIT):
        a = _cheb_fit(np.log(b / x0) ** 2)
        ti = tau[1:, None]
        u1, dt = ti * TS_A[None, :], ti * TS_A[::-1][None, :]
        sd = s * np.sqrt(dt)
        dm = (np.log(b[1:, None] / _b_of(a, x0, st, np.sqrt(u1))) + (r - q - 0.5 * s * s) * dt) / sd
        ni = np.sum(TS_W * np.exp(r * u1) * _ncdf(dm), 1)
        di = np.sum(TS_W * np.exp(q * u1) * _ncdf(dm + sd), 1)
        sd0 = s * np.sqrt(tau[1:])
        d0 = (np.log(b[1:]) + (r - q - 0.5 * s * s) * tau[1:]) / sd0
        v = np.exp(-(r - q) * tau[1:]) * (_ncdf(d0) + r * tau[1:] * ni) / (_ncdf(d0 + sd0) + q * tau[1:] * di)
        b[1:] = np.minimum(v, x0)
    a = _cheb_fit(np.log(b / x0) ** 2)
    kk = S if call else K
    x = (K if call else S) / kk
    if x <= _b_of(a, x0, st, np.array(st)):
        sd = s * st
        dm = (np.log(x) + (r - q - 0.5 * s * s) * T) / sd
        eu = np.exp(-r * T) * _ncdf(-dm) - x * np.exp(-q * T) * _ncdf(-dm - sd)
        return kk * ((1.0 - x) - eu)
    u1, dt = T * TS_A, T * TS_A[::-1]
    sd = s * np.sqrt(dt)
    dm = (np.log(x / _b_of(a, x0, st, np.sqrt(u1))) + (r - q - 0.5 * s * s) * dt) / sd
    p = np.sum(TS_W * (r * np.exp(-r * dt) * _ncdf(-dm) - q * x * np.exp(-q * dt) * _ncdf(-dm - sd)))
    return kk * T * p
