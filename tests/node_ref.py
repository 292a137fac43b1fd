"""numpy restatement of the surrogate's mesh geometry: the sinh x-grid of an option's FD
grid (src/Pricer/kwFd1d.cpp:105-125, density 0.25 and scale 50 by default), the node
bracket of log(S/K) (src/Math/kwFd1d.cpp:139-158: the first node x_j >= log(S/K)), and the
damping n / mu of Crank-Nicolson's highest-frequency modes on the strike's cell after the
march's n = t_dim - 1 steps, mu = (sigma^2 / 2) dt / dx^2, below which the surrogate
neither labels nor serves an option."""
from __future__ import annotations

import numpy as np

DAMP_MIN = 10.0


def nodes(o, x_dim, density=0.25, scale=50.0):
    """Mesh nodes, one row per option: x_j = density sinh(yMin (1 - j/(N-1)) + yMax j/(N-1))."""
    xmax = scale * o["z"] * np.sqrt(o["t"])
    ymin, ymax = np.arcsinh(-xmax / density), np.arcsinh(xmax / density)
    j = np.arange(x_dim)[None, :] / (x_dim - 1.0)
    return density * np.sinh(ymin[:, None] * (1 - j) + ymax[:, None] * j)


def bracket(grid, x):
    """Index of the first node >= x in each row (x_dim when none is)."""
    ge = grid >= np.asarray(x)[:, None]
    return np.where(ge.any(axis=1), np.argmax(ge, axis=1), grid.shape[1])


def kink_damping(o, x_dim, t_dim, density=0.25, scale=50.0):
    g = nodes(o, x_dim, density, scale)
    jk = bracket(g, np.zeros(len(o)))
    r = np.arange(len(o))
    dx = g[r, jk] - g[r, jk - 1]
    dt = o["t"] / (t_dim - 1.0)
    mu = 0.5 * o["z"] ** 2 * dt / dx ** 2
    return (t_dim - 1.0) / mu
