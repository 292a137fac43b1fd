"""Seeded synthetic portfolios for the batched American FD pricer.

A portable SplitMix64 stream (state += 0x9E3779B97F4A7C15, then the standard
mix) with U = (x >> 11) * 2**-53, so that Python, C++ and HIP code can all
regenerate the same options. Per option the draws come in the order
K, T, sigma, r, q, w (w = put when the top bit of the raw draw is 0), then any
extra draws a workload needs. All options are American with spot 100.

Workloads (sizes are the ones the benchmarks and parity tests use):
  plain_512     128 options, 512 x-points x 512 t-points      (seed 0)
  plain_1024    32768 options, 1024 x-points x 1024 t-points  (seed 1)
  uneven        65536 options, x in {256,512,1024,2048},
                t-points = max(32, ceil(365 T)), sigma U[0.05,0.8],
                2% low-vol/long-maturity stress slice          (seed 2)
  surrogate     training box K U[40,200], T U[1/52,5]         (seed 3)
  serving       1M options, ~10% outside the training box      (seed 4)
"""
from __future__ import annotations

import math

import numpy as np

OPTION_DTYPE = np.dtype(
    [("t", "<f8"), ("k", "<f8"), ("z", "<f8"), ("r", "<f8"), ("q", "<f8"), ("s", "<f8"),
     ("e", "u1"), ("w", "i1")],
    align=True,
)
assert OPTION_DTYPE.itemsize == 56

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """First `count` outputs of SplitMix64 started at state `seed`."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def unit(x: np.ndarray) -> np.ndarray:
    return (x >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def make_options(n: int, seed: int, *, k=(50.0, 150.0), t=(1.0 / 12.0, 3.0), z=(0.1, 0.6),
                 r=(0.0, 0.1), q=(0.0, 0.08), spot=100.0, american=True, extra=0):
    """n options from the per-option draw order K, T, sigma, r, q, w (+ `extra` draws).

    Returns (options, extra_draws) where extra_draws has shape (n, extra) of raw uint64.
    """
    per = 6 + extra
    raw = splitmix64(seed, n * per).reshape(n, per)
    u = unit(raw)
    opts = np.zeros(n, dtype=OPTION_DTYPE)
    opts["k"] = k[0] + (k[1] - k[0]) * u[:, 0]
    opts["t"] = t[0] + (t[1] - t[0]) * u[:, 1]
    opts["z"] = z[0] + (z[1] - z[0]) * u[:, 2]
    opts["r"] = r[0] + (r[1] - r[0]) * u[:, 3]
    opts["q"] = q[0] + (q[1] - q[0]) * u[:, 4]
    opts["w"] = np.where((raw[:, 5] >> np.uint64(63)) == 0, -1, 1).astype(np.int8)
    opts["s"] = spot
    opts["e"] = 1 if american else 0
    return opts, raw[:, 6:]


def plain(n: int = 128, seed: int = 0):
    """The plain-method distributions: K U[50,150], T U[1/12,3], sigma U[0.1,0.6],
    r U[0,0.1], q U[0,0.08], 50/50 put/call, spot 100, American."""
    opts, _ = make_options(n, seed)
    return opts


X_BUCKETS = np.array([256, 512, 1024, 2048], dtype=np.int32)


def uneven(n: int = 65536, seed: int = 2, stress_every: int = 50):
    """Uneven-grid batch: x-points from X_BUCKETS via the top two bits of the 7th draw,
    t-points = max(32, ceil(365 T)), sigma U[0.05, 0.8]; every `stress_every`-th option
    (2%) is moved into the low-vol/long-maturity corner sigma U[0.05,0.08], T U[2.5,3]
    using the 8th and 9th draws."""
    opts, extra = make_options(n, seed, z=(0.05, 0.8), extra=3)
    x_dims = X_BUCKETS[(extra[:, 0] >> np.uint64(62)).astype(np.int64)]
    stress = np.arange(n) % stress_every == 0
    us = unit(extra[:, 1:3])
    opts["z"] = np.where(stress, 0.05 + 0.03 * us[:, 0], opts["z"])
    opts["t"] = np.where(stress, 2.5 + 0.5 * us[:, 1], opts["t"])
    t_dims = np.maximum(32, np.ceil(365.0 * opts["t"])).astype(np.int32)
    return opts, x_dims.astype(np.int32), t_dims


# Surrogate training box (inputs outside it are routed to the FD solver).
SURROGATE_BOX = {"k": (40.0, 200.0), "t": (1.0 / 52.0, 5.0), "z": (0.1, 0.6), "r": (0.0, 0.1),
                 "q": (0.0, 0.08)}


def surrogate_train(n: int, seed: int = 3):
    b = SURROGATE_BOX
    opts, _ = make_options(n, seed, k=b["k"], t=b["t"], z=b["z"], r=b["r"], q=b["q"])
    return opts


def serving(n: int = 1 << 20, seed: int = 4, outside_every: int = 10):
    """Serving batch: the training box, with every `outside_every`-th option (~10%) pushed
    outside it (T in (5,6] or sigma in (0.6,0.8], chosen by the 7th draw's top bit)."""
    b = SURROGATE_BOX
    opts, extra = make_options(n, seed, k=b["k"], t=b["t"], z=b["z"], r=b["r"], q=b["q"],
                               extra=2)
    out = np.arange(n) % outside_every == 0
    which = (extra[:, 0] >> np.uint64(63)) == 0
    u = unit(extra[:, 1])
    opts["t"] = np.where(out & which, 5.0 + u, opts["t"])
    opts["z"] = np.where(out & ~which, 0.6 + 0.2 * u, opts["z"])
    return opts


def serving_grids(opts, seed: int = 5):
    """Per-option grid sizes for a serving batch, by the uneven batch's rule: x-points
    from X_BUCKETS by the top two bits of a SplitMix64 stream of its own (so the options
    themselves are unchanged), t-points = max(32, ceil(365 T))."""
    raw = splitmix64(seed, len(opts))
    x_dims = X_BUCKETS[(raw >> np.uint64(62)).astype(np.int64)].astype(np.int32)
    return x_dims, t_points_for(opts["t"])


def t_points_for(t: np.ndarray) -> np.ndarray:
    return np.maximum(32, np.ceil(365.0 * np.asarray(t))).astype(np.int32)


def cost_point_steps(x_dims, t_dims) -> int:
    """Sum over options of x-points * (t-points - 1)."""
    return int(np.sum(np.asarray(x_dims, np.int64) * (np.asarray(t_dims, np.int64) - 1)))


def fd_flops_per_option(x_dim: int, t_dim: int) -> float:
    """Algorithmic work of one plain FD solve: 10 flops per point-step (3-point RHS 5,
    forward substitution 3, back-substitution 2) with the time-invariant operator
    factored once."""
    return 10.0 * x_dim * (t_dim - 1)


def is_finite_price(p) -> bool:
    return math.isfinite(float(p))
