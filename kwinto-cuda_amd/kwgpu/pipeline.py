"""Surrogate pipeline: fitting to the GPU FD pricer, held-out evaluation against FD,
gate calibration and surrogate-first serving.

Everything runs on the device through the kwgpu C-ABI; torch provides device memory
and timing events only.

A surrogate emulates one FD configuration. `train_and_evaluate` fits it to the FD pricer
on one grid (1024 x-points x 1024 t-points) or, with `per_option_grids`, on per-option
grids drawn by the serving batch's rule (x-points from {256, 512, 1024, 2048}, t-points
max(32, ceil(365 T))). Labels are the FD pricer's values at the two mesh nodes that
bracket each training option's log(S/K) (kwgpu_surrogate_fit); the held-out check prices
options at their own spots, where the surrogate interpolates between its node values
exactly as the FD pricer interpolates between its own, against the FD price on the same
grid. Workloads (kwgpu.workloads): the training set is `surrogate_train` (K U[40,200],
T U[1/52,5], the plain-method sigma/r/q/w draws, seed 3); the serving batch is `serving`
(the same box with ~10% of options moved outside it, seed 4).
"""
from __future__ import annotations

import time

import numpy as np

from . import FP64, Context, Gate, Surrogate, gate_class, params, surrogate_baseline
from . import gate as make_gate
from . import workloads as W

X_DIM = 1024
T_DIM = 1024
ERR_BAR = 1e-4          # max abs price error allowed for the surrogate-first form
SPREAD_MAX = 5e-5       # ensemble disagreement above which an option goes to FD


def _torch():
    import torch
    return torch


def to_device(opts):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(opts).view(np.uint8).copy()).cuda()


def fd_prices(ctx: Context, d_opts, n: int, precision=FP64, x_dim=X_DIM, t_dim=T_DIM,
              with_status=False):
    """FD prices of device options (device float64 tensor). Raises on any failed option
    unless `with_status`, which returns (prices, status) instead."""
    torch = _torch()
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_s = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.fd1d_price_device(d_opts.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(),
                          params(x_dim, t_dim, precision=precision))
    ctx.synchronize()
    if with_status:
        return d_p, d_s
    bad = int((d_s != 0).sum().item())
    if bad:
        raise RuntimeError(f"{bad} FD solves failed")
    return d_p


def bs_euro(ctx: Context, opts) -> "object":
    """Black-Scholes European prices of the options (device float64 tensor)."""
    torch = _torch()
    e = np.ascontiguousarray(opts).copy()
    e["e"] = 0
    d_e = to_device(e)
    d_b = torch.empty(len(e), dtype=torch.float64, device="cuda")
    ctx.bs_price_device(d_e.data_ptr(), len(e), d_b.data_ptr())
    ctx.synchronize()
    return d_b


def device_dims(n, x_dims=None, t_dims=None):
    """Device int32 tensors of per-option grid sizes (or (None, None))."""
    torch = _torch()
    if x_dims is None and t_dims is None:
        return None, None
    xd = np.full(n, X_DIM, np.int32) if x_dims is None else np.asarray(x_dims, np.int32)
    td = np.full(n, T_DIM, np.int32) if t_dims is None else np.asarray(t_dims, np.int32)
    assert xd.shape == (n,) and td.shape == (n,)
    return torch.from_numpy(xd.copy()).cuda(), torch.from_numpy(td.copy()).cuda()


def fd_config(d_x=None, d_t=None, precision=FP64):
    """The FD configuration: 1024 x 1024, or per-option grids from device tensors."""
    dims = {}
    if d_x is not None:
        dims = {"x_dims": d_x.data_ptr(), "t_dims": d_t.data_ptr()}
    return params(X_DIM, T_DIM, precision=precision, **dims)


def predict(ctx: Context, sur: Surrogate, d_opts, n: int, fd=None):
    torch = _torch()
    d_pm = torch.empty(n, dtype=torch.float64, device="cuda")
    d_sp = torch.empty(n, dtype=torch.float64, device="cuda")
    d_ib = torch.empty(n, dtype=torch.uint8, device="cuda")
    sur.predict(d_opts.data_ptr(), n, d_pm.data_ptr(), d_sp.data_ptr(), d_ib.data_ptr(), fd)
    ctx.synchronize()
    return d_pm, d_sp, d_ib


MIN_CALIBRATED = 1000


def calibrate_spread(err, spread, bar=ERR_BAR, cap=SPREAD_MAX, margin=0.5,
                     min_count=MIN_CALIBRATED):
    """Largest spread bound <= cap such that every held-out option at or under it has
    |error| <= margin * bar, and at least `min_count` held-out options are under it;
    -1 (a bound no spread meets: serve nothing, and the pricer skips the ensemble)
    otherwise. With k exchangeable held-out options under the bound,
    a new option under it exceeds their largest error with probability <= 1/(k+1), so
    the count is what makes the held-out maximum mean anything."""
    order = np.argsort(spread)
    s, e = spread[order], np.abs(err[order])
    ok = np.maximum.accumulate(e) <= margin * bar
    idx = np.nonzero(ok & (s <= cap))[0]
    if not len(idx) or idx[-1] + 1 < min_count:
        return -1.0
    return float(s[idx[-1]])


# 0.15 and 0.2: with the exercise-boundary control variate the larger-premium classes
# can be served up to there (36.0 % -> 37.3 % of the calibration set, the same maximum)
SCALE_CAPS = (0.01, 0.015, 0.02, 0.03, 0.05, 0.1, 0.15, 0.2, float("inf"))


# Serve only when the gate saves at least this share of the FD solves: the serving path's
# fixed costs (two count read-backs, the compactions) eat smaller savings -- a per-option
# gate expected to save 1.2 % measured 0.998x plain FD.
MIN_GAIN = 0.02


def calibrate_gate(err, spread, scale, bar=ERR_BAR, cap=SPREAD_MAX, caps=SCALE_CAPS, cost=0.0,
                   min_gain=0.0):
    """(spread bound, scale bound) of the gate: an option is served when its baseline
    scale (baseline premium + 1e-4 K, the price scale of the network's error) is at most
    the scale bound and its ensemble spread at most the spread bound. For each candidate
    scale cap the spread bound is calibrated (calibrate_spread) on the held-out options
    under the cap. The cap matters because the members share their biases: where the
    baseline premium is large, a small spread does not bound the error.
    Every option under the cap (a candidate) is run through the ensemble, which costs
    `cost` FD solves (inference_cost_ratio); an option served saves one. The cap with the
    largest saving, (served - cost * candidates) / held-out options, wins (the smaller on
    ties). A saving below `min_gain` is not worth the error tail a batch may show beyond
    the held-out maximum: then, as with no saving at all, the gate serves nothing:
    (-1, -1)."""
    b, c, gain = _best_cap(err, spread, scale, len(err), bar, cap, caps, cost, MIN_CALIBRATED)
    return (b, c) if gain >= min_gain and b >= 0 else (-1.0, -1.0)


def _best_cap(err, spread, scale, n_all, bar, cap, caps, cost, min_count):
    """(spread bound, scale cap, saving as a share of n_all) of the cap that saves most,
    or (-1, -1, 0) when none saves anything."""
    best = (-1.0, -1.0, 0.0)
    for c in caps:
        m = scale <= c
        if m.sum() < min_count:
            continue
        b = calibrate_spread(err[m], spread[m], bar=bar, cap=cap, min_count=min_count)
        n = int(np.sum(m & (spread <= b))) if b >= 0 else 0
        gain = (n - cost * float(m.sum())) / n_all
        if n > 0 and gain > best[2]:
            best = (b, float(c), gain)
    return best


# Classes of the gate: put/call x maturity bands x log-moneyness bands (kwgpu_gate)
GATE_T_EDGES = (0.5, 2.0)
GATE_M_EDGES = (0.0,)
MIN_CALIBRATED_CLASS = 10000


def n_classes(t_edges=GATE_T_EDGES, m_edges=GATE_M_EDGES) -> int:
    return 2 * (len(t_edges) + 1) * (len(m_edges) + 1)


def calibrate_classes(err, spread, scale, cls, t_edges=GATE_T_EDGES, m_edges=GATE_M_EDGES,
                      bar=ERR_BAR, cap=SPREAD_MAX, caps=SCALE_CAPS, cost=0.0, min_gain=0.0,
                      min_count=MIN_CALIBRATED_CLASS):
    """The gate by class (kwgpu_gate): each class's (spread bound, scale bound) calibrated
    as calibrate_gate calibrates the one gate, on that class's held-out options alone
    (`cls`: their class indices, kwgpu.gate_class), with at least `min_count` of them
    under its bounds. Each class then keeps the guarantee of its own calibration, and a
    class whose errors are small is not held to the bound of one whose errors are large.
    A class that saves nothing is closed (-1); if all classes together save less than
    `min_gain` of the held-out options' FD solves, all are. Returns a dict of the band
    edges and the per-class bounds (gate(), serve())."""
    k = n_classes(t_edges, m_edges)
    smax, cmax, total = [-1.0] * k, [-1.0] * k, 0.0
    for c in range(k):
        m = cls == c
        if m.sum() < min_count:
            continue
        b, sc, gain = _best_cap(err[m], spread[m], scale[m], len(err), bar, cap, caps, cost,
                                min_count)
        if b >= 0:
            smax[c], cmax[c], total = b, sc, total + gain
    if total < min_gain:
        smax, cmax = [-1.0] * k, [-1.0] * k
    return {"t_edges": list(t_edges), "m_edges": list(m_edges), "spread_max": smax,
            "scale_max": cmax}


def one_class_gate(spread_max: float, scale_max: float = float("inf")):
    """The one-class gate (kwgpu_surrogate_price's two bounds) as a gate dict."""
    return {"t_edges": [], "m_edges": [], "spread_max": [spread_max] * 2,
            "scale_max": [scale_max] * 2}


def gate_mask(gate, opts, spread, scale):
    """Which options a gate dict serves, given their spreads and scales (host)."""
    c = gate_class(opts, gate["t_edges"], gate["m_edges"])
    return (spread <= np.asarray(gate["spread_max"])[c]) & (scale <= np.asarray(gate["scale_max"])[c])


def to_gate(g) -> Gate:
    return make_gate(g["t_edges"], g["m_edges"], g["spread_max"], g["scale_max"])


def inference_cost_ratio(ctx: Context, sur: Surrogate, d_opts, n: int, fd) -> float:
    """Time of one ensemble prediction per option over the time of one FD solve per
    option, on the same n options and FD configuration (each timed after a warm-up
    call): what running an option through the network costs, in FD solves."""
    torch = _torch()
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    st = torch.empty(n, dtype=torch.int32, device="cuda")

    def timed(f):
        f()
        ctx.synchronize()
        t0 = time.perf_counter()
        f()
        ctx.synchronize()
        return time.perf_counter() - t0

    t_fd = timed(lambda: ctx.fd1d_price_device(d_opts.data_ptr(), n, out.data_ptr(),
                                               st.data_ptr(), fd))
    t_nn = timed(lambda: predict(ctx, sur, d_opts, n, fd))
    return t_nn / t_fd


def heldout_errors(ctx: Context, sur: Surrogate, opts, d_x=None, d_t=None):
    """(error, spread, scale, count, options) of the surrogate's prices against FD at the
    options' own spots (same configuration), for the in-box options with a labelled
    bracket whose FD price succeeded; scale = baseline premium + 1e-4 K."""
    torch = _torch()
    n = len(opts)
    d_o = to_device(opts)
    fd = fd_config(d_x, d_t)
    truth_d = torch.empty(n, dtype=torch.float64, device="cuda")
    st_d = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.fd1d_price_device(d_o.data_ptr(), n, truth_d.data_ptr(), st_d.data_ptr(), fd)
    pm, sp, ib = predict(ctx, sur, d_o, n, fd)
    bs = bs_euro(ctx, opts).cpu().numpy()
    truth, st = truth_d.cpu().numpy(), st_d.cpu().numpy()
    inb = (ib.cpu().numpy() == 1) & (st == 0)
    err = (bs + pm.cpu().numpy()) - truth
    scale = surrogate_baseline(ctx, opts) + 1e-4 * opts["k"]
    return err[inb], sp.cpu().numpy()[inb], scale[inb], int(inb.sum()), opts[inb]


def train_and_evaluate(ctx: Context, n_train=2_000_000, n_test=65536, members=4, epochs=160,
                       batch=8192, seed=3, per_option_grids=False, n_calibrate=4 << 20,
                       fitted: Surrogate | None = None, log=print):
    """Fits the ensemble to the FD pricer, reports its error on the held-out options, and
    calibrates the gate on a separate set of `n_calibrate` options (seed + 1000) priced by
    FD. With k held-out options under the bounds, each new option under them exceeds
    their largest error with probability <= 1/(k+1), so a batch serving as many options
    as k expects one such option: the set is four times a 1M serving batch, so that one
    is expected in four such batches, and the margin of half the bar absorbs it (a 100k
    set once admitted a 1.2e-4 error on a 1M batch; a 1M set once admitted 1.8e-4 for
    the per-option ensemble, which saved 0.4 % of the FD work).
    `fitted`: a surrogate already fitted to this configuration (Surrogate.load), which is
    evaluated and calibrated the same way without fitting again."""
    opts = W.surrogate_train(n_train + n_test, seed)
    dims = W.serving_grids(opts, seed + 100) if per_option_grids else (None, None)
    d_x, d_t = device_dims(len(opts), *dims)
    grid = "per-option grids" if per_option_grids else f"{X_DIM}x{T_DIM}"
    if fitted is None:
        d_o = to_device(opts)
        sur = Surrogate(ctx, members=members, seed=seed)
        fd_tr = fd_config(None if d_x is None else d_x[:n_train],
                          None if d_t is None else d_t[:n_train])
        t0 = time.perf_counter()
        loss, pairs = sur.fit(d_o.data_ptr(), n_train, fd_tr, epochs=epochs, batch=batch)
        t_fit = time.perf_counter() - t0
        final_loss = loss[-1].tolist()
        log(f"fit: {pairs} of {n_train} options labelled at their two bracketing nodes "
            f"(FD FP64, {grid}), {members} members x {epochs} epochs in {t_fit:.2f} s, "
            f"final loss {final_loss}")
    else:
        sur, pairs, t_fit, final_loss, epochs = fitted, 0, 0.0, None, 0
        members = sur.members
    # held-out options at their own spots, against the FD price on the same grid
    ho = slice(n_train, n_train + n_test)
    e_in, s_in, c_in, n_in, o_in = heldout_errors(ctx, sur, opts[ho],
                                                  None if d_x is None else d_x[ho],
                                                  None if d_t is None else d_t[ho])
    # the gate, calibrated on its own set
    cal = W.surrogate_train(n_calibrate, seed + 1000)
    cdims = W.serving_grids(cal, seed + 1100) if per_option_grids else (None, None)
    c_x, c_t = device_dims(len(cal), *cdims)
    e_c, s_c, c_c, _, o_c = heldout_errors(ctx, sur, cal, c_x, c_t)
    kappa = inference_cost_ratio(ctx, sur, to_device(cal), len(cal), fd_config(c_x, c_t))
    # the one gate (reported for comparison) and the gate by class (served with)
    thr, scale_max = calibrate_gate(e_c, s_c, c_c, cost=kappa, min_gain=MIN_GAIN)
    one_served = (s_c <= thr) & (c_c <= scale_max)
    cls_c = gate_class(o_c, GATE_T_EDGES, GATE_M_EDGES)
    gate = calibrate_classes(e_c, s_c, c_c, cls_c, cost=kappa, min_gain=MIN_GAIN)
    served = gate_mask(gate, o_c, s_c, c_c)
    served_ho = gate_mask(gate, o_in, s_in, c_in)
    # the one gate at looser error bars, for the accuracy / throughput trade-off (by
    # class, the 1e-2 bar once admitted a 1.07e-2 error on the 1M batch: with a spread cap
    # of half such a bar the classes' tails are far longer than at 1e-4)
    looser = {}
    for bar in (1e-3, 1e-2):
        tb, tc = calibrate_gate(e_c, s_c, c_c, bar=bar, cap=bar / 2, cost=kappa,
                                min_gain=MIN_GAIN)
        gb = one_class_gate(tb, tc)
        sv = gate_mask(gb, o_c, s_c, c_c)
        looser[f"{bar:g}"] = {"gate": gb, "calibration_pass_fraction": float(np.mean(sv)),
                              "calibration_max_err_served":
                                  float(np.max(np.abs(e_c[sv]))) if sv.any() else 0.0}
    out = {
        "grid": grid, "n_train": n_train, "n_labelled_pairs": int(pairs), "n_test": n_test,
        "n_test_in_box": n_in, "members": members, "epochs": epochs,
        "fit_s": t_fit, "final_loss": final_loss,
        "heldout_max_abs_err": float(np.max(np.abs(e_in))),
        "heldout_rms_abs_err": float(np.sqrt(np.mean(e_in * e_in))),
        "heldout_frac_within_bar": float(np.mean(np.abs(e_in) <= ERR_BAR)),
        "n_calibration": int(len(e_c)),
        "inference_cost_ratio": kappa,
        "gate": gate,
        "one_class_gate": {"spread_bound": thr, "scale_bound": scale_max,
                           "calibration_pass_fraction": float(np.mean(one_served))},
        "calibration_candidate_fraction": float(np.mean(
            c_c <= np.asarray(gate["scale_max"])[cls_c])),
        "calibration_pass_fraction": float(np.mean(served)),
        "calibration_max_err_served": float(np.max(np.abs(e_c[served]))) if served.any() else 0.0,
        "heldout_pass_fraction": float(np.mean(served_ho)),
        "heldout_max_err_served":
            float(np.max(np.abs(e_in[served_ho]))) if served_ho.any() else 0.0,
        "looser_bars": looser,
    }
    return sur, out


def serve(ctx: Context, sur: Surrogate, opts, gate, repeats=3, log=print,
          x_dims=None, t_dims=None, precision=FP64):
    """Surrogate-first pricing of `opts` under `gate` (a gate dict: calibrate_classes,
    one_class_gate) against plain FP64 FD of the same batch:
    options/s of both, max/RMS error of the served prices vs FP64 FD. `x_dims`/`t_dims`
    (host int32 arrays, optional) give every option its own grid, for the FD fallback and
    for the FP64 FD reference alike. `precision` is the fallback's; FP64 is the default since
    on MI355X the FP64 form is both exact and the faster one."""
    torch = _torch()
    n = len(opts)
    d_o = to_device(opts)
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_r = torch.empty(n, dtype=torch.int32, device="cuda")
    d_x, d_t = device_dims(n, x_dims, t_dims)
    fd = fd_config(d_x, d_t, precision)
    fd64 = fd_config(d_x, d_t, FP64)
    d_ref = torch.empty(n, dtype=torch.float64, device="cuda")
    d_st = torch.empty(n, dtype=torch.int32, device="cuda")

    def fd_ref():
        ctx.fd1d_price_device(d_o.data_ptr(), n, d_ref.data_ptr(), d_st.data_ptr(), fd64)
        ctx.synchronize()

    g = to_gate(gate)
    sur.price_gated(d_o.data_ptr(), n, g, fd, d_p.data_ptr(), d_r.data_ptr())  # warm-up
    torch.cuda.synchronize()
    ts = []
    nfd = 0
    for _ in range(repeats):
        t0 = time.perf_counter()
        nfd = sur.price_gated(d_o.data_ptr(), n, g, fd, d_p.data_ptr(), d_r.data_ptr())
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t_sur = min(ts)
    served = d_p.cpu().numpy()
    route = d_r.cpu().numpy()
    fd_ref()  # warm-up, as for the surrogate-first path
    torch.cuda.synchronize()
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        fd_ref()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t_fd = min(ts)
    if int((d_st != 0).sum().item()):
        raise RuntimeError("FP64 FD reference: failed options")
    ref = d_ref.cpu().numpy()
    d = np.abs(served - ref)
    out = {
        "n": n, "routed_to_fd": int(nfd), "served_by_surrogate": int(np.sum(route == 0)),
        "surrogate_first_options_per_s": n / t_sur, "fd_fp64_options_per_s": n / t_fd,
        "speedup_vs_fd_fp64": t_fd / t_sur,
        "max_abs_err_vs_fp64_fd": float(d.max()), "rms_abs_err_vs_fp64_fd":
            float(np.sqrt(np.mean(d * d))),
        "max_abs_err_surrogate_served": float(d[route == 0].max()) if np.any(route == 0) else 0.0,
    }
    log(f"serving: {out}")
    return out
