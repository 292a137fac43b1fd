"""CPU oracle for the batched American FD pricer -- TEST INFRASTRUCTURE ONLY.

ctypes bindings over liboracle.so (kw_oracle.cpp, an FP64 restatement of the
reference pricer) and, where it was built, _ref/libkwref.so (the reference's own
library sources compiled in place behind ref_shim.cpp).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker; the product path never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libkwref.so")

_lib = None
_ref = None

OPTION_DTYPE = np.dtype(
    [("t", "<f8"), ("k", "<f8"), ("z", "<f8"), ("r", "<f8"), ("q", "<f8"), ("s", "<f8"),
     ("e", "u1"), ("w", "i1")],
    align=True,
)

_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_D = ctypes.c_double
_I64 = ctypes.c_int64
_I = ctypes.c_int


def build(reference_tree: str | None = None) -> None:
    """Compile liboracle.so, and _ref/libkwref.so when the reference tree exists."""
    subprocess.check_call(["make", "-s", "-B", "-C", HERE, "all"])
    ref = reference_tree or "/root/reference"
    if os.path.isdir(os.path.join(ref, "src")):
        subprocess.check_call(["make", "-s", "-B", "-C", HERE, "ref", f"REF={ref}"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", HERE, "all"])
        L = ctypes.CDLL(LIB_PATH)
        L.kwo_fd1d_price.argtypes = [_P, _SZ, _D, _D, _I64, _I64, _I, _P, ctypes.c_char_p, _SZ]
        L.kwo_fd1d_price_status.argtypes = [_P, _SZ, _D, _D, _I64, _I64, _I, _P, _P]
        L.kwo_fd1d_price_uneven.argtypes = [_P, _SZ, _D, _D, _P, _P, _I, _P, _P]
        L.kwo_bs_price.argtypes = [_P, _SZ, _P]
        L.kwo_fd1d_bs_price.argtypes = [_P, _SZ, _D, _D, _I64, _I64, _I, _P, ctypes.c_char_p, _SZ]
        L.kwo_fd1d_solve_one.argtypes = [_P, _D, _D, _I64, _I64, _P, _P]
        L.kwo_solve_tridiagonal.argtypes = [_I, _P, _P, _P, _P, _P]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_LIB_PATH)
        R.kwref_price.argtypes = [ctypes.c_char_p, _P, _SZ, _D, _D, ctypes.c_longlong,
                                  ctypes.c_longlong, _P, ctypes.c_char_p, _SZ]
        R.kwref_threads.restype = ctypes.c_int
        _ref = R
    return _ref


def _opts(opts) -> np.ndarray:
    a = np.ascontiguousarray(opts)
    assert a.dtype.itemsize == 56
    return a


def fd1d_price(opts, x_dim=512, t_dim=512, density=0.25, scale=50.0, nthreads=0):
    """Returns (prices, error_string). Mirrors Fd1d_Pricer::price error behaviour."""
    a = _opts(opts)
    out = np.zeros(len(a), np.float64)
    err = ctypes.create_string_buffer(512)
    rc = lib().kwo_fd1d_price(a.ctypes.data, len(a), density, scale, t_dim, x_dim, nthreads,
                              out.ctypes.data, err, 512)
    return out, (err.value.decode() if rc else "")


def fd1d_price_status(opts, x_dim=512, t_dim=512, density=0.25, scale=50.0, nthreads=0):
    a = _opts(opts)
    out = np.zeros(len(a), np.float64)
    st = np.zeros(len(a), np.int32)
    lib().kwo_fd1d_price_status(a.ctypes.data, len(a), density, scale, t_dim, x_dim, nthreads,
                                out.ctypes.data, st.ctypes.data)
    return out, st


def fd1d_price_uneven(opts, x_dims, t_dims, density=0.25, scale=50.0, nthreads=0):
    a = _opts(opts)
    xd = np.ascontiguousarray(x_dims, np.int32)
    td = np.ascontiguousarray(t_dims, np.int32)
    out = np.zeros(len(a), np.float64)
    st = np.zeros(len(a), np.int32)
    lib().kwo_fd1d_price_uneven(a.ctypes.data, len(a), density, scale, td.ctypes.data,
                                xd.ctypes.data, nthreads, out.ctypes.data, st.ctypes.data)
    return out, st


def bs_price(opts):
    a = _opts(opts)
    out = np.zeros(len(a), np.float64)
    lib().kwo_bs_price(a.ctypes.data, len(a), out.ctypes.data)
    return out


def fd1d_bs_price(opts, x_dim=512, t_dim=512, density=0.25, scale=50.0, nthreads=0):
    a = _opts(opts)
    out = np.zeros(len(a), np.float64)
    err = ctypes.create_string_buffer(512)
    rc = lib().kwo_fd1d_bs_price(a.ctypes.data, len(a), density, scale, t_dim, x_dim, nthreads,
                                 out.ctypes.data, err, 512)
    return out, (err.value.decode() if rc else "")


def solve_one(opt, x_dim=512, t_dim=512, density=0.25, scale=50.0):
    a = _opts(np.asarray([opt], dtype=OPTION_DTYPE) if not isinstance(opt, np.ndarray) else opt)
    x = np.zeros(x_dim)
    v = np.zeros(x_dim)
    lib().kwo_fd1d_solve_one(a.ctypes.data, density, scale, t_dim, x_dim, x.ctypes.data,
                             v.ctypes.data)
    return x, v


def solve_tridiagonal(al, a, au, y):
    n = len(a)
    arrs = [np.ascontiguousarray(v, np.float64) for v in (al, a, au, y)]
    x = np.zeros(n)
    rc = lib().kwo_solve_tridiagonal(n, *[v.ctypes.data for v in arrs], x.ctypes.data)
    return x, rc


def ref_threads() -> int:
    """Worker threads the compiled reference's thread pool runs on this host."""
    return int(ref_lib().kwref_threads())


def ref_price(pricer, opts, x_dim=512, t_dim=512, density=0.25, scale=50.0):
    """The compiled reference (only where _ref/libkwref.so was built)."""
    a = _opts(opts)
    out = np.zeros(len(a), np.float64)
    err = ctypes.create_string_buffer(512)
    rc = ref_lib().kwref_price(pricer.encode(), a.ctypes.data, len(a), density, scale, t_dim,
                               x_dim, out.ctypes.data, err, 512)
    return out, (err.value.decode() if rc else "")


# ---- Greeks by central bumps (test/portfolio.py:35-56) ----------------------------
# delta = (V(s(1+b)) - V(s(1-b))) / (2 b s); gamma = (V(s(1+b)) - 2 V + V(s(1-b))) / (b s)^2;
# vega, theta, rho are the first differences in z, t, r divided by 100, 360 and 100.

GREEK_FIELDS = ("price", "delta", "gamma", "vega", "theta", "rho")
BUMPED = ("s", "z", "t", "r")


def bump_sets(opts, bump=0.01):
    """The nine option arrays the bump formulas price: the options as given, then each
    of s, z, t, r times (1 + bump) and times (1 - bump)."""
    a = _opts(opts)
    sets = [a.copy()]
    for f in BUMPED:
        for sgn in (1.0, -1.0):
            b = a.copy()
            b[f] = (1.0 + sgn * bump) * a[f]
            sets.append(b)
    return sets


def greeks_from_prices(opts, v, bump=0.01):
    """Greeks from the (9, n) prices of bump_sets, evaluated in the order the formulas
    are written (a zero rate, volatility or expiry gives NaN where the script would
    divide by zero)."""
    a = _opts(opts)
    v = np.asarray(v, np.float64)
    out = np.zeros(len(a), dtype=[(f, "<f8") for f in GREEK_FIELDS])
    with np.errstate(divide="ignore", invalid="ignore"):
        def first(k, f):
            return (v[k] - v[k + 1]) / (2.0 * bump * a[f]) + 0.0
        out["price"] = v[0]
        out["delta"] = first(1, "s")
        out["gamma"] = (v[1] - 2.0 * v[0] + v[2]) / (bump * a["s"]) ** 2 + 0.0
        out["vega"] = first(3, "z") / 100.0
        out["theta"] = first(5, "t") / 360.0
        out["rho"] = first(7, "r") / 100.0
    return out


def fd1d_greeks(opts, x_dim=512, t_dim=512, bump=0.01, density=0.25, scale=50.0, nthreads=0):
    """(greeks, status, prices[9, n]) of the FD pricer by the bump formulas; status is the
    first non-zero status of an option's nine prices."""
    sets = bump_sets(opts, bump)
    v, st = zip(*[fd1d_price_status(s, x_dim, t_dim, density, scale, nthreads) for s in sets])
    v, st = np.stack(v), np.stack(st)
    status = np.zeros(st.shape[1], np.int32)
    for k in range(st.shape[0] - 1, -1, -1):  # the lowest-numbered failing price wins
        status = np.where(st[k] != 0, st[k], status)
    return greeks_from_prices(opts, v, bump), status, v
