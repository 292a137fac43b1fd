"""ctypes binding of the kwgpu C-ABI (include/kwgpu.h) -- the gfx950 FD pricer.

Loads lib/libkwgpu.so from this package's directory and raises if it is missing:
there is no CPU fallback. Pricing names mirror the reference's pricer interface
(`Pricer::price(assets, prices)` in src/Pricer/kwPricer.h:12-22); options are
numpy arrays of OPTION_DTYPE, byte-identical to kw::Option (56 bytes).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .workloads import OPTION_DTYPE

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libkwgpu.so")

OK, ERR_INVALID, ERR_HIP, ERR_RANGE, ERR_NOMEM, ERR_UNSUPPORTED = range(6)
STATUS_OK, STATUS_OUT_OF_GRID, STATUS_BAD_INPUT, STATUS_NONFINITE = range(4)
FP64, FP32_MIXED = 0, 1
LAYOUT_AUTO, LAYOUT_LANE_SOA, LAYOUT_WAVE_REG = 0, 1, 2

EXPORTED_SYMBOLS = (
    "kwgpu_fd1d_params_default", "kwgpu_create", "kwgpu_destroy", "kwgpu_last_error",
    "kwgpu_fd1d_price", "kwgpu_fd1d_price_device", "kwgpu_bs_price", "kwgpu_bs_price_device",
    "kwgpu_fd1d_bs_price", "kwgpu_synchronize", "kwgpu_stream",
)


class Fd1dParams(ctypes.Structure):
    _fields_ = [
        ("density", ctypes.c_double), ("scale", ctypes.c_double), ("theta", ctypes.c_double),
        ("x_dim", ctypes.c_int32), ("t_dim", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("layout", ctypes.c_int32), ("x_dims", ctypes.c_void_p), ("t_dims", ctypes.c_void_p),
    ]


_lib = None


def lib():
    """The loaded C-ABI library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"kwgpu: {LIB_PATH} not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P, SZ = ctypes.c_void_p, ctypes.c_size_t
        L.kwgpu_fd1d_params_default.argtypes = [ctypes.POINTER(Fd1dParams)]
        L.kwgpu_fd1d_params_default.restype = None
        L.kwgpu_create.argtypes = [ctypes.POINTER(P), ctypes.c_int]
        L.kwgpu_destroy.argtypes = [P]
        L.kwgpu_destroy.restype = None
        L.kwgpu_last_error.argtypes = [P]
        L.kwgpu_last_error.restype = ctypes.c_char_p
        L.kwgpu_fd1d_price.argtypes = [P, P, SZ, ctypes.POINTER(Fd1dParams), P, P]
        L.kwgpu_fd1d_price_device.argtypes = [P, P, SZ, ctypes.POINTER(Fd1dParams), P, P, P]
        L.kwgpu_bs_price.argtypes = [P, P, SZ, P]
        L.kwgpu_bs_price_device.argtypes = [P, P, SZ, P, P]
        L.kwgpu_fd1d_bs_price.argtypes = [P, P, SZ, ctypes.POINTER(Fd1dParams), P, P]
        L.kwgpu_synchronize.argtypes = [P]
        L.kwgpu_stream.argtypes = [P]
        L.kwgpu_stream.restype = P
        _lib = L
    return _lib


def params(x_dim=512, t_dim=512, density=0.25, scale=50.0, theta=0.5, precision=FP64,
           layout=LAYOUT_AUTO, x_dims=None, t_dims=None) -> Fd1dParams:
    """Fd1d_Pricer config (src/Pricer/kwFd1d.cpp:9-19). x_dims / t_dims: integer
    addresses (device or host, matching the entry point) or None."""
    p = Fd1dParams()
    lib().kwgpu_fd1d_params_default(ctypes.byref(p))
    p.density, p.scale, p.theta = density, scale, theta
    p.x_dim, p.t_dim = x_dim, t_dim
    p.precision, p.layout = precision, layout
    p.x_dims = x_dims
    p.t_dims = t_dims
    return p


class KwgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


def as_options(opts) -> np.ndarray:
    a = np.ascontiguousarray(opts)
    if a.dtype != OPTION_DTYPE:
        a = a.astype(OPTION_DTYPE)
    return a


class Context:
    """A kwgpu context: one HIP stream and its device buffers on `device`."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().kwgpu_create(ctypes.byref(h), device)
        if rc != OK:
            raise KwgpuError(rc, f"kwgpu_create(device={device}) failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().kwgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def last_error(self) -> str:
        return lib().kwgpu_last_error(self.h).decode()

    @property
    def stream(self) -> int:
        return lib().kwgpu_stream(self.h)

    def synchronize(self):
        rc = lib().kwgpu_synchronize(self.h)
        if rc != OK:
            raise KwgpuError(rc, self.last_error)

    # -- host-pointer entry points (synchronous, reference semantics) ---------------
    def fd1d_price(self, opts, x_dim=512, t_dim=512, *, x_dims=None, t_dims=None, **kw):
        """Returns (prices, status, rc, error). rc/error follow Fd1d_Pricer::price."""
        a = as_options(opts)
        n = len(a)
        prices = np.full(n, np.nan)
        status = np.zeros(n, np.int32)
        xd = None if x_dims is None else np.ascontiguousarray(x_dims, np.int32)
        td = None if t_dims is None else np.ascontiguousarray(t_dims, np.int32)
        p = params(x_dim, t_dim, x_dims=None if xd is None else xd.ctypes.data,
                   t_dims=None if td is None else td.ctypes.data, **kw)
        rc = lib().kwgpu_fd1d_price(self.h, a.ctypes.data, n, ctypes.byref(p),
                                    prices.ctypes.data, status.ctypes.data)
        return prices, status, rc, (self.last_error if rc else "")

    def fd1d_bs_price(self, opts, x_dim=512, t_dim=512, **kw):
        a = as_options(opts)
        n = len(a)
        prices = np.full(n, np.nan)
        status = np.zeros(n, np.int32)
        p = params(x_dim, t_dim, **kw)
        rc = lib().kwgpu_fd1d_bs_price(self.h, a.ctypes.data, n, ctypes.byref(p),
                                       prices.ctypes.data, status.ctypes.data)
        return prices, status, rc, (self.last_error if rc else "")

    def bs_price(self, opts):
        a = as_options(opts)
        prices = np.full(len(a), np.nan)
        rc = lib().kwgpu_bs_price(self.h, a.ctypes.data, len(a), prices.ctypes.data)
        if rc != OK:
            raise KwgpuError(rc, self.last_error)
        return prices

    # -- device entry points (asynchronous on `stream`) ----------------------------
    def fd1d_price_device(self, d_opts: int, n: int, d_prices: int, d_status: int | None,
                          p: Fd1dParams, stream: int | None = None):
        rc = lib().kwgpu_fd1d_price_device(self.h, d_opts, n, ctypes.byref(p), d_prices,
                                           d_status, stream)
        if rc != OK:
            raise KwgpuError(rc, self.last_error)

    def bs_price_device(self, d_opts: int, n: int, d_prices: int, stream: int | None = None):
        rc = lib().kwgpu_bs_price_device(self.h, d_opts, n, d_prices, stream)
        if rc != OK:
            raise KwgpuError(rc, self.last_error)
