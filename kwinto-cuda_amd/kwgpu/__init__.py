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
STATUS_OK, STATUS_OUT_OF_GRID, STATUS_BAD_INPUT, STATUS_NONFINITE, STATUS_PRECISION = range(5)
FP64, FP32_MIXED = 0, 1
LAYOUT_AUTO, LAYOUT_LANE_SOA, LAYOUT_WAVE_REG = 0, 1, 2

EXPORTED_SYMBOLS = (
    "kwgpu_fd1d_params_default", "kwgpu_create", "kwgpu_destroy", "kwgpu_last_error",
    "kwgpu_fd1d_price", "kwgpu_fd1d_price_device", "kwgpu_fd1d_price_device_noalloc",
    "kwgpu_bs_price", "kwgpu_bs_price_device",
    "kwgpu_fd1d_bs_price", "kwgpu_fd1d_greeks", "kwgpu_synchronize", "kwgpu_stream",
    "kwgpu_surrogate_train_params_default", "kwgpu_surrogate_create", "kwgpu_surrogate_destroy",
    "kwgpu_surrogate_train", "kwgpu_surrogate_labels", "kwgpu_surrogate_fit",
    "kwgpu_surrogate_predict",
    "kwgpu_surrogate_price", "kwgpu_surrogate_price_gated",
    "kwgpu_surrogate_get_params", "kwgpu_surrogate_set_params", "kwgpu_surrogate_loss_grad",
    "kwgpu_surrogate_features", "kwgpu_surrogate_baseline", "kwgpu_surrogate_save", "kwgpu_surrogate_load",
    "kwgpu_fd1d_autotune", "kwgpu_tune_save", "kwgpu_tune_load",
)
SURROGATE_PARAMS = 199937
GREEKS_DTYPE = np.dtype([(f, "<f8") for f in ("price", "delta", "gamma", "vega", "theta", "rho")])


class SurrogateTrainParams(ctypes.Structure):
    _fields_ = [
        ("epochs", ctypes.c_int32), ("batch", ctypes.c_int32), ("lr", ctypes.c_float),
        ("lr_final", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
        ("eps", ctypes.c_float), ("seed", ctypes.c_uint64),
    ]


GATE_MAX_T, GATE_MAX_M = 8, 4
GATE_MAX_CLASSES = 2 * GATE_MAX_T * GATE_MAX_M


class Gate(ctypes.Structure):
    """kwgpu_gate: per-class bounds over put/call x maturity bands x log-moneyness bands."""
    _fields_ = [
        ("n_t", ctypes.c_int32), ("n_m", ctypes.c_int32),
        ("t_edges", ctypes.c_double * (GATE_MAX_T - 1)),
        ("m_edges", ctypes.c_double * (GATE_MAX_M - 1)),
        ("spread_max", ctypes.c_double * GATE_MAX_CLASSES),
        ("scale_max", ctypes.c_double * GATE_MAX_CLASSES),
    ]


def gate(t_edges, m_edges, spread_max, scale_max) -> Gate:
    """A Gate from its inner band edges and per-class bounds, indexed
    (c * n_t + a) * n_m + b for c = 0 (puts) / 1 (calls), T band a, log(S/K) band b."""
    g = Gate()
    g.n_t, g.n_m = len(t_edges) + 1, len(m_edges) + 1
    k = 2 * g.n_t * g.n_m
    if g.n_t > GATE_MAX_T or g.n_m > GATE_MAX_M or len(spread_max) != k or len(scale_max) != k:
        raise ValueError("gate: band counts or bound arrays do not fit")
    for i, v in enumerate(t_edges):
        g.t_edges[i] = v
    for i, v in enumerate(m_edges):
        g.m_edges[i] = v
    for i in range(k):
        g.spread_max[i], g.scale_max[i] = spread_max[i], scale_max[i]
    return g


def gate_class(opts, t_edges, m_edges) -> np.ndarray:
    """Class index of each option under a gate with these band edges (as the device
    assigns it: a value equal to an edge falls in the band above)."""
    c = np.where(np.asarray(opts["w"]) < 0, 0, 1)
    a = np.searchsorted(np.asarray(t_edges, float), opts["t"], side="right")
    b = np.searchsorted(np.asarray(m_edges, float), np.log(opts["s"] / opts["k"]), side="right")
    return (c * (len(t_edges) + 1) + a) * (len(m_edges) + 1) + b


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
        # One HIP runtime serves the whole process and the first one loaded wins: load
        # torch's (it works only with its own) before this library when torch is present.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
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
        L.kwgpu_fd1d_price_device_noalloc.argtypes = [P, P, SZ, ctypes.POINTER(Fd1dParams), P, P,
                                                      P]
        L.kwgpu_bs_price.argtypes = [P, P, SZ, P]
        L.kwgpu_bs_price_device.argtypes = [P, P, SZ, P, P]
        L.kwgpu_fd1d_bs_price.argtypes = [P, P, SZ, ctypes.POINTER(Fd1dParams), P, P]
        L.kwgpu_fd1d_greeks.argtypes = [P, P, SZ, ctypes.POINTER(Fd1dParams), ctypes.c_double,
                                        P, P]
        L.kwgpu_synchronize.argtypes = [P]
        L.kwgpu_stream.argtypes = [P]
        L.kwgpu_stream.restype = P
        L.kwgpu_surrogate_train_params_default.argtypes = [ctypes.POINTER(SurrogateTrainParams)]
        L.kwgpu_surrogate_train_params_default.restype = None
        L.kwgpu_surrogate_create.argtypes = [P, ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(P)]
        L.kwgpu_surrogate_destroy.argtypes = [P]
        L.kwgpu_surrogate_destroy.restype = None
        L.kwgpu_surrogate_train.argtypes = [P, P, P, P, SZ, ctypes.POINTER(SurrogateTrainParams), P]
        L.kwgpu_surrogate_labels.argtypes = [P, P, SZ, ctypes.POINTER(Fd1dParams), P, P, P]
        L.kwgpu_surrogate_fit.argtypes = [P, P, P, SZ, ctypes.POINTER(Fd1dParams),
                                          ctypes.POINTER(SurrogateTrainParams), P,
                                          ctypes.POINTER(ctypes.c_size_t)]
        L.kwgpu_surrogate_predict.argtypes = [P, P, P, SZ, ctypes.POINTER(Fd1dParams), P, P, P, P]
        L.kwgpu_surrogate_price.argtypes = [P, P, P, SZ, ctypes.c_double, ctypes.c_double,
                                            ctypes.POINTER(Fd1dParams), P, P,
                                            ctypes.POINTER(ctypes.c_size_t)]
        L.kwgpu_surrogate_price_gated.argtypes = [P, P, P, SZ, ctypes.POINTER(Gate),
                                                  ctypes.POINTER(Fd1dParams), P, P,
                                                  ctypes.POINTER(ctypes.c_size_t)]
        L.kwgpu_surrogate_get_params.argtypes = [P, P, ctypes.c_int32, P]
        L.kwgpu_surrogate_set_params.argtypes = [P, P, ctypes.c_int32, P]
        L.kwgpu_surrogate_loss_grad.argtypes = [P, P, ctypes.c_int32, P, P, SZ, P, P]
        L.kwgpu_surrogate_features.argtypes = [P, P, SZ, ctypes.c_int32, ctypes.c_int32, P]
        L.kwgpu_surrogate_baseline.argtypes = [P, P, SZ, P]
        L.kwgpu_surrogate_save.argtypes = [P, P, ctypes.c_char_p]
        L.kwgpu_surrogate_load.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(P)]
        L.kwgpu_fd1d_autotune.argtypes = [P, P, SZ, ctypes.POINTER(Fd1dParams), P, P,
                                          ctypes.c_char_p, SZ]
        L.kwgpu_tune_save.argtypes = [P, ctypes.c_char_p]
        L.kwgpu_tune_load.argtypes = [P, ctypes.c_char_p]
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

    def fd1d_bs_price(self, opts, x_dim=512, t_dim=512, *, x_dims=None, t_dims=None, **kw):
        """Fd1d_BlackScholes_Pricer::price: (prices, status, rc, error)."""
        a = as_options(opts)
        n = len(a)
        prices = np.full(n, np.nan)
        status = np.zeros(n, np.int32)
        xd = None if x_dims is None else np.ascontiguousarray(x_dims, np.int32)
        td = None if t_dims is None else np.ascontiguousarray(t_dims, np.int32)
        p = params(x_dim, t_dim, x_dims=None if xd is None else xd.ctypes.data,
                   t_dims=None if td is None else td.ctypes.data, **kw)
        rc = lib().kwgpu_fd1d_bs_price(self.h, a.ctypes.data, n, ctypes.byref(p),
                                       prices.ctypes.data, status.ctypes.data)
        return prices, status, rc, (self.last_error if rc else "")

    def fd1d_greeks(self, opts, x_dim=512, t_dim=512, bump=0.01, *, x_dims=None, t_dims=None,
                    **kw):
        """(greeks, status, rc, error): greeks is a structured array with fields
        price, delta, gamma, vega, theta, rho (central bumps of relative size `bump`)."""
        a = as_options(opts)
        n = len(a)
        out = np.full(n, np.nan, dtype=GREEKS_DTYPE)
        status = np.zeros(n, np.int32)
        xd = None if x_dims is None else np.ascontiguousarray(x_dims, np.int32)
        td = None if t_dims is None else np.ascontiguousarray(t_dims, np.int32)
        p = params(x_dim, t_dim, x_dims=None if xd is None else xd.ctypes.data,
                   t_dims=None if td is None else td.ctypes.data, **kw)
        rc = lib().kwgpu_fd1d_greeks(self.h, a.ctypes.data, n, ctypes.byref(p), bump,
                                     out.ctypes.data, status.ctypes.data)
        return out, status, rc, (self.last_error if rc else "")

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

    def fd1d_price_device_noalloc(self, d_opts: int, n: int, d_prices: int, d_status: int | None,
                                  p: Fd1dParams, stream: int | None = None):
        """kwgpu_fd1d_price_device_noalloc: p.x_dims / p.t_dims are host addresses here."""
        rc = lib().kwgpu_fd1d_price_device_noalloc(self.h, d_opts, n, ctypes.byref(p), d_prices,
                                                   d_status, stream)
        if rc != OK:
            raise KwgpuError(rc, self.last_error)

    def autotune(self, d_opts: int, n: int, d_prices: int, d_status: int | None,
                 p: Fd1dParams) -> dict:
        """Times the kernel variants per grid-size bucket and keeps the fastest."""
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        rc = lib().kwgpu_fd1d_autotune(self.h, d_opts, n, ctypes.byref(p), d_prices, d_status,
                                       buf, len(buf))
        if rc != OK:
            raise KwgpuError(rc, self.last_error)
        return json.loads(buf.value.decode())

    def tune_save(self, path: str):
        rc = lib().kwgpu_tune_save(self.h, path.encode())
        if rc != OK:
            raise KwgpuError(rc, self.last_error)

    def tune_load(self, path: str):
        rc = lib().kwgpu_tune_load(self.h, path.encode())
        if rc != OK:
            raise KwgpuError(rc, self.last_error)

    def bs_price_device(self, d_opts: int, n: int, d_prices: int, stream: int | None = None):
        rc = lib().kwgpu_bs_price_device(self.h, d_opts, n, d_prices, stream)
        if rc != OK:
            raise KwgpuError(rc, self.last_error)


class Surrogate:
    """An ensemble of premium networks on a Context (see kwgpu_surrogate_* in kwgpu.h)."""

    def __init__(self, ctx: Context, members: int = 4, seed: int = 1, _handle=None):
        self.ctx = ctx
        if _handle is not None:
            self.h = _handle
        else:
            h = ctypes.c_void_p()
            rc = lib().kwgpu_surrogate_create(ctx.h, members, seed, ctypes.byref(h))
            if rc != OK:
                raise KwgpuError(rc, ctx.last_error)
            self.h = h
        self.members = members

    @classmethod
    def load(cls, ctx: Context, path: str) -> "Surrogate":
        h = ctypes.c_void_p()
        rc = lib().kwgpu_surrogate_load(ctx.h, path.encode(), ctypes.byref(h))
        if rc != OK:
            raise KwgpuError(rc, ctx.last_error)
        with open(path, "rb") as f:
            members = int(np.frombuffer(f.read(16), np.uint32)[2])
        return cls(ctx, members, _handle=h)

    def close(self):
        if getattr(self, "h", None):
            lib().kwgpu_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise KwgpuError(rc, self.ctx.last_error)

    def save(self, path: str):
        self._check(lib().kwgpu_surrogate_save(self.ctx.h, self.h, path.encode()))

    @staticmethod
    def _train_params(epochs, batch, lr, lr_final, seed):
        tp = SurrogateTrainParams()
        lib().kwgpu_surrogate_train_params_default(ctypes.byref(tp))
        tp.epochs, tp.batch, tp.lr, tp.lr_final, tp.seed = epochs, batch, lr, lr_final, seed
        return tp

    def train(self, d_opts: int, d_premium: int, n: int, epochs=30, batch=8192, lr=2e-3,
              lr_final=2e-5, seed=7) -> np.ndarray:
        """Trains on (option, premium) rows on the surrogate's single grid."""
        tp = self._train_params(epochs, batch, lr, lr_final, seed)
        loss = np.zeros((epochs, self.members), np.float32)
        self._check(lib().kwgpu_surrogate_train(self.ctx.h, self.h, d_opts, d_premium, n,
                                                ctypes.byref(tp), loss.ctypes.data))
        return loss

    def fit(self, d_opts: int, n: int, fd: Fd1dParams, epochs=30, batch=8192, lr=2e-3,
            lr_final=2e-5, seed=7):
        """Labels each option's two bracketing mesh nodes with the FD pricer configured by
        `fd` and trains on them; returns (losses[epochs, members], options kept)."""
        tp = self._train_params(epochs, batch, lr, lr_final, seed)
        loss = np.zeros((epochs, self.members), np.float32)
        kept = ctypes.c_size_t(0)
        self._check(lib().kwgpu_surrogate_fit(self.ctx.h, self.h, d_opts, n, ctypes.byref(fd),
                                              ctypes.byref(tp), loss.ctypes.data,
                                              ctypes.byref(kept)))
        return loss, kept.value

    def predict(self, d_opts: int, n: int, d_premium: int, d_spread: int, d_in_box: int,
                fd: Fd1dParams | None = None, stream: int | None = None):
        """Premium (surrogate price - BS at the spot), spread and in-box flags; `fd` is the
        emulated configuration (needed for per-option grids)."""
        self._check(lib().kwgpu_surrogate_predict(self.ctx.h, self.h, d_opts, n,
                                                  None if fd is None else ctypes.byref(fd),
                                                  d_premium, d_spread, d_in_box, stream))

    def price(self, d_opts: int, n: int, spread_max: float, fd: Fd1dParams, d_prices: int,
              d_route: int, scale_max: float = float("inf")) -> int:
        """Surrogate-first prices into d_prices; d_route[i] = 0 (surrogate), 1 (FD at
        fd.precision) or 2 (FD in FP64). A negative spread_max serves nothing and skips
        the ensemble; options whose baseline scale exceeds scale_max skip it too.
        fd.x_dims / fd.t_dims may hold device arrays of per-option grid sizes. Returns
        the number of options priced by FD."""
        nfd = ctypes.c_size_t(0)
        self._check(lib().kwgpu_surrogate_price(self.ctx.h, self.h, d_opts, n, spread_max,
                                                scale_max, ctypes.byref(fd), d_prices, d_route,
                                                ctypes.byref(nfd)))
        return nfd.value

    def price_gated(self, d_opts: int, n: int, g: Gate, fd: Fd1dParams, d_prices: int,
                    d_route: int) -> int:
        """price() with per-class bounds (kwgpu_gate); returns the number priced by FD."""
        nfd = ctypes.c_size_t(0)
        self._check(lib().kwgpu_surrogate_price_gated(self.ctx.h, self.h, d_opts, n,
                                                      ctypes.byref(g), ctypes.byref(fd), d_prices,
                                                      d_route, ctypes.byref(nfd)))
        return nfd.value

    def get_params(self, member: int) -> np.ndarray:
        w = np.zeros(SURROGATE_PARAMS, np.float32)
        self._check(lib().kwgpu_surrogate_get_params(self.ctx.h, self.h, member, w.ctypes.data))
        return w

    def set_params(self, member: int, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == SURROGATE_PARAMS
        self._check(lib().kwgpu_surrogate_set_params(self.ctx.h, self.h, member, w.ctypes.data))

    def loss_grad(self, member: int, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = ctypes.c_float()
        g = np.zeros(SURROGATE_PARAMS, np.float32)
        self._check(lib().kwgpu_surrogate_loss_grad(self.ctx.h, self.h, member, x.ctypes.data,
                                                    t.ctypes.data, len(t), ctypes.byref(loss),
                                                    g.ctypes.data))
        return float(loss.value), g


SURROGATE_FEATURES = 8


def surrogate_features(opts, x_dim=1024, t_dim=1024) -> np.ndarray:
    """Normalised network inputs (n x 8) of host options priced on an x_dim x t_dim grid,
    as the kernels compute them."""
    a = as_options(opts)
    out = np.zeros((len(a), SURROGATE_FEATURES), np.float32)
    rc = lib().kwgpu_surrogate_features(None, a.ctypes.data, len(a), x_dim, t_dim,
                                        out.ctypes.data)
    if rc != OK:
        raise KwgpuError(rc, "kwgpu_surrogate_features failed")
    return out


def surrogate_labels(ctx: Context, d_opts: int, n: int, fd: Fd1dParams, d_nodes: int,
                     d_premium: int, d_keep: int):
    """kwgpu_surrogate_labels: node options (2n), their FD premiums (2n) and keep flags
    (n), device pointers."""
    rc = lib().kwgpu_surrogate_labels(ctx.h, d_opts, n, ctypes.byref(fd), d_nodes, d_premium,
                                      d_keep)
    if rc != OK:
        raise KwgpuError(rc, ctx.last_error)


def surrogate_baseline(ctx: Context, opts) -> np.ndarray:
    """The analytic baseline premium (Ju & Zhong's correction of Barone-Adesi & Whaley),
    computed on the device: the gate's scale is this premium + 1e-4 K."""
    a = as_options(opts)
    out = np.full(len(a), np.nan)
    rc = lib().kwgpu_surrogate_baseline(ctx.h, a.ctypes.data, len(a), out.ctypes.data)
    if rc != OK:
        raise KwgpuError(rc, ctx.last_error)
    return out
