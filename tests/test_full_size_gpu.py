"""The full-size workloads under test, through the C-ABI, at the sizes the benchmark runs:

* the uneven batch: 65536 options (seed 2), x-points from {256, 512, 1024, 2048}, t-points
  max(32, ceil(365 T)) -- up to 1095 -- through the planned, graph-replayed path in FP64
  and FP32-mixed (src/Pricer/kwFd1d.cpp:95-125 per option);
* surrogate training: 2M options labelled by the FD pricer at 1024 x 1024 in FP64 (node
  labels checked against the oracle on a sample), a 4-member ensemble, held-out error
  against FD;
* surrogate-first serving of the 1M-option batch (~10 % outside the training box) with
  the FP64 FD fallback: every FD-routed price equals the FD pricer's bit for bit, every
  out-of-box option is routed, every served price is within 1e-4 of FP64 FD, and the
  surrogate-first path prices the batch faster than plain FP64 FD does: what the trained
  network gains over the method without it.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

import kwgpu
import oracle
from conftest import load_golden
from kwgpu import pipeline as P
from kwgpu import workloads as W

pytestmark = pytest.mark.gpu

EPS = 1e-8
BAR = 1e-4
SUR_EPOCHS = 320  # the benchmark's training length


def _say(*a):
    print(*a, flush=True)


def _classes(o):
    for w, wn in ((-1, "put"), (1, "call")):
        for lo, hi, tn in ((0.0, 1.0, "T<1"), (1.0, np.inf, "T>=1")):
            yield f"{wn} {tn}", (o["w"] == w) & (o["t"] >= lo) & (o["t"] < hi)


@pytest.mark.parametrize("precision", [kwgpu.FP64, kwgpu.FP32_MIXED], ids=["fp64", "mixed"])
def test_uneven_65536_full_size(ctx, precision):
    import torch
    _say("uneven 65536, precision", precision)
    o, xd, td = W.uneven(65536, 2)
    assert td.max() >= 1095
    # the golden subset (priced by the reference): 8 options per x bucket and 4 of the
    # stress slice among the workload's first 4096, in index order
    g, c = load_golden("uneven_small")
    pick = sorted(set(sum((np.nonzero(xd[:4096] == xb)[0][:8].tolist() for xb in W.X_BUCKETS),
                          []) + np.nonzero(np.arange(4096) % 50 == 0)[0][:4].tolist()))
    pick = np.array(pick)
    assert np.array_equal(o[pick], g) and np.array_equal(xd[pick], c["x_dim"])
    assert np.array_equal(td[pick], c["t_dim"])
    n = len(o)
    d_o = torch.from_numpy(o.view(np.uint8).copy()).cuda()
    d_x = torch.from_numpy(xd.copy()).cuda()
    d_t = torch.from_numpy(td.copy()).cuda()
    prm = kwgpu.params(precision=precision, x_dims=d_x.data_ptr(), t_dims=d_t.data_ptr())
    runs = []
    for _ in range(3):  # plan + capture, then two graph replays
        d_p = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
        d_s = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        ctx.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
        ctx.synchronize()
        runs.append((d_p.cpu().numpy(), d_s.cpu().numpy()))
    p, st = runs[0]
    assert np.all(st == kwgpu.STATUS_OK)
    for q, s in runs[1:]:
        assert np.array_equal(q, p) and np.array_equal(s, st)
    d = np.abs(p[pick] - c["ref_fd1d"])
    _say("  golden subset max", d.max())
    if precision == kwgpu.FP64:
        assert d.max() <= EPS
        # a strided sample across every bucket against the oracle
        idx = np.arange(7, n, n // 64)
        ref, rst = oracle.fd1d_price_uneven(o[idx], xd[idx], td[idx])
        assert np.all(rst == 0)
        dd = np.abs(p[idx] - ref)
        _say("  oracle sample max", dd.max())
        assert dd.max() <= EPS
    else:
        assert d.max() <= BAR
        p64 = np.empty(n)
        s64 = np.empty(n, np.int32)
        d_p = torch.empty(n, dtype=torch.float64, device="cuda")
        d_s = torch.empty(n, dtype=torch.int32, device="cuda")
        f64 = kwgpu.params(precision=kwgpu.FP64, x_dims=d_x.data_ptr(), t_dims=d_t.data_ptr())
        ctx.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), f64)
        ctx.synchronize()
        p64[:], s64[:] = d_p.cpu().numpy(), d_s.cpu().numpy()
        assert np.all(s64 == 0)
        err = np.abs(p - p64)
        for name, m in _classes(o):
            _say(f"  {name}: max {err[m].max():.2e}")
            assert err[m].max() <= BAR, name


@pytest.fixture(scope="module")
def fitted(ctx):
    """The training configuration: 2M options (5 % held out) labelled by the FD pricer at
    1024 x 1024 in FP64, a 4-member ensemble."""
    n_all = 2_000_000
    n_test = n_all // 20
    sur, out = P.train_and_evaluate(ctx, n_train=n_all - n_test, n_test=n_test, members=4,
                                    epochs=SUR_EPOCHS, log=_say)
    _say("  evaluation", {k: v for k, v in out.items() if k != "looser_bars"})
    yield sur, out
    sur.close()


def test_training_labels_match_oracle(ctx):
    """kwgpu_surrogate_labels on a sample of the training set: node spots at the sinh mesh
    nodes bracketing log(S/K) (first node >= log(S/K)), premiums = the oracle's FD price
    at those spots minus its BS price, to 1e-8."""
    import torch
    o = W.surrogate_train(256, 3)
    n = len(o)
    d_o = P.to_device(o)
    d_nodes = torch.empty(2 * n * 56, dtype=torch.uint8, device="cuda")
    d_prem = torch.empty(2 * n, dtype=torch.float64, device="cuda")
    d_keep = torch.empty(n, dtype=torch.uint8, device="cuda")
    kwgpu.surrogate_labels(ctx, d_o.data_ptr(), n, kwgpu.params(1024, 1024), d_nodes.data_ptr(),
                           d_prem.data_ptr(), d_keep.data_ptr())
    nodes = d_nodes.cpu().numpy().view(kwgpu.OPTION_DTYPE)
    prem, keep = d_prem.cpu().numpy(), d_keep.cpu().numpy() == 1
    assert keep.mean() > 0.98
    xs = np.log(o["s"] / o["k"])
    xmax = 50.0 * o["z"] * np.sqrt(o["t"])
    ymin, ymax = np.arcsinh(-xmax / 0.25), np.arcsinh(xmax / 0.25)
    u = np.arange(1024)[None, :] / 1023.0
    mesh = 0.25 * np.sinh(ymin[:, None] * (1 - u) + ymax[:, None] * u)
    j = np.argmax(mesh >= xs[:, None], axis=1)
    rows = np.arange(n)
    for half, col in ((0, j - 1), (1, j)):
        got = np.log(nodes["s"][half * n:(half + 1) * n] / o["k"])
        assert np.max(np.abs(got - mesh[rows, col])[keep]) <= 1e-12
    both = np.concatenate([o, o])
    for f in ("t", "k", "z", "r", "q", "e", "w"):
        assert np.array_equal(nodes[f], both[f])
    kk = np.concatenate([keep, keep])
    fd, err = oracle.fd1d_price(nodes[kk], 1024, 1024)
    assert not err
    euro = nodes[kk].copy()
    euro["e"] = 0
    want = fd - oracle.bs_price(euro)
    d = np.abs(prem[kk] - want)
    _say("  labels vs oracle: max", d.max())
    assert d.max() <= EPS


def test_training_heldout_error(fitted):
    sur, out = fitted
    _say("  held-out rms", out["heldout_rms_abs_err"], "max", out["heldout_max_abs_err"],
         "pass", out["heldout_pass_fraction"], "gate", out["gate"])
    assert out["n_labelled_pairs"] >= out["n_train"] - 100
    assert all(np.isfinite(out["final_loss"]))
    assert out["heldout_rms_abs_err"] <= 3e-3
    assert out["heldout_frac_within_bar"] >= 0.3
    # the gate by class, calibrated on its own 4M-option set, serves a share of it within
    # half the bar, and of the held-out set within the bar; every open class's spread bound
    # is within the cap, and the classes serve more than the one gate does
    assert out["n_calibration"] > 4_000_000
    g = out["gate"]
    assert len(g["spread_max"]) == P.n_classes() and max(g["spread_max"]) > 0
    assert all(b <= P.SPREAD_MAX for b in g["spread_max"])
    assert all(c > 0 for b, c in zip(g["spread_max"], g["scale_max"]) if b >= 0)
    assert out["calibration_pass_fraction"] > 0.02
    assert out["calibration_pass_fraction"] >= out["one_class_gate"]["calibration_pass_fraction"]
    assert out["calibration_max_err_served"] <= 0.5 * BAR
    assert out["heldout_pass_fraction"] > 0.02 and out["heldout_max_err_served"] <= BAR


def test_serving_1m_full_size(ctx, fitted):
    import torch
    sur, out = fitted
    o = W.serving(1 << 20, 4)
    n = len(o)
    d_o = P.to_device(o)
    fd = kwgpu.params(1024, 1024)
    d_ref = torch.empty(n, dtype=torch.float64, device="cuda")
    d_st = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.fd1d_price_device(d_o.data_ptr(), n, d_ref.data_ptr(), d_st.data_ptr(), fd)
    ctx.synchronize()
    ref, st = d_ref.cpu().numpy(), d_st.cpu().numpy()
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_r = torch.empty(n, dtype=torch.int32, device="cuda")
    gate = P.to_gate(out["gate"])
    t0 = time.perf_counter()
    nfd = sur.price_gated(d_o.data_ptr(), n, gate, fd, d_p.data_ptr(), d_r.data_ptr())
    t_sur = time.perf_counter() - t0
    p, route = d_p.cpu().numpy(), d_r.cpu().numpy()
    served = route == 0
    _say(f"  served {served.sum()} of {n} ({served.mean():.3f}), surrogate-first {t_sur:.3f} s")
    assert nfd == n - served.sum()
    # the FD fallback is the FD pricer: bit for bit, statuses included
    fdr = ~served
    assert np.array_equal(p[fdr & (st == 0)], ref[fdr & (st == 0)])
    assert np.all(np.isnan(p[fdr & (st != 0)]))
    # every option outside the training box goes to FD
    feats = kwgpu.surrogate_features(o, 1024, 1024)
    outside = ~np.all(np.abs(feats[:, :6]) <= 1 + 1e-9, axis=1)
    assert outside.mean() > 0.08 and np.all(fdr[outside])
    # the served prices: within the bar of FP64 FD, on every one of them
    assert served.sum() > 0
    err = np.abs(p[served] - ref[served])
    _say(f"  served max err {err.max():.3e}, rms {np.sqrt(np.mean(err ** 2)):.3e}")
    assert err.max() <= BAR
    # a strided sample against the oracle
    idx = np.arange(5, n, n // 48)
    oref, ost = oracle.fd1d_price_status(o[idx], 1024, 1024)
    ok = ost == 0
    assert np.all(np.abs(ref[idx][ok] - oref[ok]) <= EPS)
    assert np.all(np.abs(p[idx][ok] - oref[ok]) <= np.where(served[idx][ok], BAR, EPS))
    # the gain: the gate serves a share of the batch and the surrogate-first path is the
    # faster one (best of 3 timed calls each, after the calls above)
    def best(f):
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            f()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)
    t_fd = best(lambda: ctx.fd1d_price_device(d_o.data_ptr(), n, d_ref.data_ptr(),
                                              d_st.data_ptr(), fd))
    t_sf = best(lambda: sur.price_gated(d_o.data_ptr(), n, gate, fd, d_p.data_ptr(),
                                        d_r.data_ptr()))
    _say(f"  plain FP64 FD {t_fd * 1e3:.1f} ms, surrogate-first {t_sf * 1e3:.1f} ms: "
         f"{t_fd / t_sf:.3f}x with {served.mean():.3f} of the batch served")
    # measured: 31 % served, 1.36x (the served share is set by the fit, which is the
    # same on every box; the ratio by both solves on the same box)
    assert served.mean() >= 0.2
    assert t_fd / t_sf >= 1.15
