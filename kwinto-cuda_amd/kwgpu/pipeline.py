"""Surrogate pipeline: labels from the GPU FD pricer, ensemble training, held-out
evaluation against FD, gate calibration and surrogate-first serving.

Everything runs on the device through the kwgpu C-ABI; torch provides device memory
and timing events only.

Workloads (kwgpu.workloads): the training set is `surrogate_train` (K U[40,200],
T U[1/52,5], the plain-method sigma/r/q/w draws, seed 3); the serving batch is
`serving` (the same box with ~10% of options moved outside it, seed 4). Labels use one
grid, 1024 x-points x 1024 t-points. Serving uses that grid for the FD fallback too, so
served prices and the FP64 FD prices they are compared with describe the same
discretisation, or per-option grids (`serve(..., x_dims, t_dims)`), against which FP64
FD on those same grids is the reference.
"""
from __future__ import annotations

import time

import numpy as np

from . import FP64, Context, Surrogate, params
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


def premium_labels(ctx: Context, opts, chunk=1 << 18):
    """(device options, device premium = FD FP64 - BS, kept) for a host option array.
    Options the FD pricer rejects (a spot outside its grid: the box's corner K = 200,
    sigma = 0.1, T = 1/52 has log(S/K) at the grid's end) are dropped; `kept` is the
    host index of the options returned."""
    torch = _torch()
    n = len(opts)
    d_o = to_device(opts)
    prem = torch.empty(n, dtype=torch.float64, device="cuda")
    st = torch.empty(n, dtype=torch.int32, device="cuda")
    bs = bs_euro(ctx, opts)
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        prem[c0:c1], st[c0:c1] = fd_prices(ctx, d_o[c0 * 56:c1 * 56], c1 - c0, with_status=True)
    prem -= bs
    ok = (st == 0).cpu().numpy()
    kept = np.nonzero(ok)[0]
    if len(kept) < n:
        sel = torch.from_numpy(kept).cuda()
        d_o = d_o.view(n, 56)[sel].reshape(-1).contiguous()
        prem = prem[sel].contiguous()
    return d_o, prem, kept


def predict(ctx: Context, sur: Surrogate, d_opts, n: int):
    torch = _torch()
    d_pm = torch.empty(n, dtype=torch.float64, device="cuda")
    d_sp = torch.empty(n, dtype=torch.float64, device="cuda")
    d_ib = torch.empty(n, dtype=torch.uint8, device="cuda")
    sur.predict(d_opts.data_ptr(), n, d_pm.data_ptr(), d_sp.data_ptr(), d_ib.data_ptr())
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


def train_and_evaluate(ctx: Context, n_train=2_000_000, n_test=65536, members=4, epochs=30,
                       batch=8192, seed=3, log=print):
    """Trains the ensemble on FD labels and evaluates it on held-out FD labels."""
    t0 = time.perf_counter()
    opts = W.surrogate_train(n_train + n_test, seed)
    d_o, prem, kept = premium_labels(ctx, opts)
    t_label = time.perf_counter() - t0
    dropped = n_train + n_test - len(kept)
    n_train = int(np.sum(kept < n_train))  # kept options keep their order
    n_test = len(kept) - n_train
    log(f"labels: {len(kept)} options by FD FP64 {X_DIM}x{T_DIM} in {t_label:.2f} s "
        f"({dropped} outside their grid dropped)")
    sur = Surrogate(ctx, members=members, seed=seed)
    t0 = time.perf_counter()
    loss = sur.train(d_o.data_ptr(), prem.data_ptr(), n_train, epochs=epochs, batch=batch)
    t_train = time.perf_counter() - t0
    log(f"training: {members} members x {epochs} epochs in {t_train:.2f} s, "
        f"final loss {loss[-1].tolist()}")
    d_te = d_o[n_train * 56:]
    pm, sp, ib = predict(ctx, sur, d_te, n_test)
    truth = prem[n_train:].cpu().numpy()
    err = pm.cpu().numpy() - truth
    spread = sp.cpu().numpy()
    thr = calibrate_spread(err, spread)
    served = spread <= thr
    # the same calibration at looser error bars, for the accuracy / throughput trade-off
    looser = {}
    for bar in (1e-3, 1e-2):
        tb = calibrate_spread(err, spread, bar=bar, cap=bar / 2)
        sv = spread <= tb
        looser[f"{bar:g}"] = {"spread_bound": tb, "heldout_pass_fraction": float(np.mean(sv)),
                              "heldout_max_err_served":
                                  float(np.max(np.abs(err[sv]))) if sv.any() else 0.0}
    out = {
        "n_train": n_train, "n_test": n_test, "dropped_outside_grid": dropped,
        "members": members, "epochs": epochs,
        "label_s": t_label, "train_s": t_train, "final_loss": loss[-1].tolist(),
        "heldout_max_abs_err": float(np.max(np.abs(err))),
        "heldout_rms_abs_err": float(np.sqrt(np.mean(err * err))),
        "label_rms": float(np.sqrt(np.mean((truth - truth.mean()) ** 2))),
        "spread_bound": thr,
        "heldout_pass_fraction": float(np.mean(served)),
        "heldout_max_err_served": float(np.max(np.abs(err[served]))) if served.any() else 0.0,
        "looser_bars": looser,
    }
    return sur, out


def serve(ctx: Context, sur: Surrogate, opts, spread_bound: float, repeats=3, log=print,
          x_dims=None, t_dims=None, precision=FP64):
    """Surrogate-first pricing of `opts` against plain FP64 FD of the same batch:
    options/s of both, max/RMS error of the served prices vs FP64 FD. `x_dims`/`t_dims`
    (host int32 arrays, optional) give every option its own grid, for the FD fallback and
    for the FP64 FD reference alike. `precision` is the fallback's; FP64 is the default since
    on MI355X the FP64 form is both exact and the faster one."""
    torch = _torch()
    n = len(opts)
    d_o = to_device(opts)
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_r = torch.empty(n, dtype=torch.int32, device="cuda")
    dims = {}
    if x_dims is not None or t_dims is not None:
        xd = np.full(n, X_DIM, np.int32) if x_dims is None else np.asarray(x_dims, np.int32)
        td = np.full(n, T_DIM, np.int32) if t_dims is None else np.asarray(t_dims, np.int32)
        assert xd.shape == (n,) and td.shape == (n,)
        d_x = torch.from_numpy(xd.copy()).cuda()
        d_t = torch.from_numpy(td.copy()).cuda()
        dims = {"x_dims": d_x.data_ptr(), "t_dims": d_t.data_ptr()}
    fd = params(X_DIM, T_DIM, precision=precision, **dims)
    fd64 = params(X_DIM, T_DIM, precision=FP64, **dims)
    d_ref = torch.empty(n, dtype=torch.float64, device="cuda")
    d_st = torch.empty(n, dtype=torch.int32, device="cuda")

    def fd_ref():
        ctx.fd1d_price_device(d_o.data_ptr(), n, d_ref.data_ptr(), d_st.data_ptr(), fd64)
        ctx.synchronize()

    sur.price(d_o.data_ptr(), n, spread_bound, fd, d_p.data_ptr(), d_r.data_ptr())  # warm-up
    torch.cuda.synchronize()
    ts = []
    nfd = 0
    for _ in range(repeats):
        t0 = time.perf_counter()
        nfd = sur.price(d_o.data_ptr(), n, spread_bound, fd, d_p.data_ptr(), d_r.data_ptr())
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
