"""The premium network's BF16-MFMA forward/backward kernels, its training, and the
surrogate-first serving path, through the C-ABI.

Gradient oracle: an explicit float64 restatement of the network's forward and
backward pass with BF16 rounding (round-to-nearest-even) applied at exactly the points
the kernels round (activations entering the 256x256 GEMMs, the BF16 weight copies, the
layer deltas), so the comparison can be tight; plus torch's FP32 autograd without any
BF16 rounding as an independent, looser check.
"""
from __future__ import annotations

import numpy as np
import pytest

import kwgpu
from kwgpu import workloads as W

pytestmark = pytest.mark.gpu

H, F = 256, 8
kW1, kB1 = 0, 256 * 8
kW2 = kB1 + H
kB2 = kW2 + H * H
kW3 = kB2 + H
kB3 = kW3 + H * H
kW4 = kB3 + H
kB4 = kW4 + H * H
kW5 = kB4 + H
kB5 = kW5 + H
NP = kB5 + 1


def _say(*a):
    print(*a, flush=True)


def bf16(a):
    """Round to bfloat16 (round to nearest even) and back, as v_cvt to bf16 does."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w):
    w = w.astype(np.float64)
    return dict(W1=w[kW1:kB1].reshape(H, F), b1=w[kB1:kW2], W2=w[kW2:kB2].reshape(H, H),
                b2=w[kB2:kW3], W3=w[kW3:kB3].reshape(H, H), b3=w[kB3:kW4],
                W4=w[kW4:kB4].reshape(H, H), b4=w[kB4:kW5], W5=w[kW5:kB5], b5=w[kB5])


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sig(z)


def dsilu(z):
    s = sig(z)
    return s * (1 + z * (1 - s))


def oracle_loss_grad(w, x, t):
    p = unpack(w)
    n = len(t)
    x = x.astype(np.float64)
    z = [x @ p["W1"].T + p["b1"]]
    h = [bf16(silu(z[0]))]
    for l in (2, 3, 4):
        z.append(h[-1] @ bf16(p[f"W{l}"]).T + p[f"b{l}"])
        h.append(bf16(silu(z[-1])))
    y = silu(z[3]) @ p["W5"] + p["b5"]
    loss = np.mean((y - t) ** 2)
    dy = 2 * (y - t) / n
    g = {"W5": silu(z[3]).T @ dy, "b5": np.sum(dy)}
    dz = bf16(np.outer(dy, p["W5"]) * dsilu(z[3]))
    for l in (4, 3, 2):
        g[f"W{l}"] = dz.T @ h[l - 2]
        g[f"b{l}"] = dz.sum(0)
        dz = bf16((dz @ bf16(p[f"W{l}"])) * dsilu(z[l - 2]))
    g["W1"] = dz.T @ x
    g["b1"] = dz.sum(0)
    flat = np.concatenate([g["W1"].ravel(), g["b1"], g["W2"].ravel(), g["b2"], g["W3"].ravel(),
                           g["b3"], g["W4"].ravel(), g["b4"], g["W5"], [g["b5"]]])
    return loss, flat


SLICES = {"W1": (kW1, kB1), "b1": (kB1, kW2), "W2": (kW2, kB2), "b2": (kB2, kW3),
          "W3": (kW3, kB3), "b3": (kB3, kW4), "W4": (kW4, kB4), "b4": (kB4, kW5),
          "W5": (kW5, kB5), "b5": (kB5, NP)}


@pytest.fixture(scope="module")
def sur(ctx):
    s = kwgpu.Surrogate(ctx, members=2, seed=11)
    yield s
    s.close()


@pytest.mark.parametrize("n", [64, 333, 4096])
def test_loss_and_gradient_match_oracle(sur, n):
    _say("loss/grad n =", n)
    rng = np.random.default_rng(n)
    w = sur.get_params(0)
    # make the output layer and biases non-trivial so every path carries signal
    w[kW5:kB5] = rng.uniform(-0.3, 0.3, H).astype(np.float32)
    w[kB2:kW3] = rng.uniform(-0.1, 0.1, H).astype(np.float32)
    sur.set_params(0, w)
    x = rng.uniform(-1, 1, (n, F)).astype(np.float32)
    x[:, 5] = np.sign(x[:, 5])
    t = rng.normal(0, 0.5, n).astype(np.float32)
    loss, g = sur.loss_grad(0, x, t)
    rl, rg = oracle_loss_grad(w, x, t)
    assert abs(loss - rl) <= 1e-4 * abs(rl)
    for name, (a, b) in SLICES.items():
        ga, gr = g[a:b].astype(np.float64), rg[a:b]
        err = np.linalg.norm(ga - gr) / max(np.linalg.norm(gr), 1e-30)
        _say(f"  {name}: rel err {err:.2e}")
        assert err < 5e-3, name


def test_gradient_close_to_fp32_autograd(sur):
    """The kernels' gradient against torch's FP32 autograd of the same network, which has
    no BF16 rounding: close over all parameters, and slice by slice (each layer's weights
    and biases), so that a wrong slice -- a bias gradient summed over the wrong axis, a
    missing factor -- fails on its own."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    n = 1024
    w = sur.get_params(1)
    # a non-trivial output layer and biases, so that every slice carries signal
    w[kW5:kB5] = rng.uniform(-0.3, 0.3, H).astype(np.float32)
    for b in (kB1, kB2, kB3, kB4):
        w[b:b + H] = rng.uniform(-0.1, 0.1, H).astype(np.float32)
    sur.set_params(1, w)
    x = rng.uniform(-1, 1, (n, F)).astype(np.float32)
    t = rng.normal(0, 0.5, n).astype(np.float32)
    _, g = sur.loss_grad(1, x, t)
    tw = torch.tensor(w, dtype=torch.float32, requires_grad=True)
    p = {k: tw[a:b] for k, (a, b) in SLICES.items()}
    act = torch.nn.functional.silu
    h = act(torch.tensor(x) @ p["W1"].view(H, F).T + p["b1"])
    for l in (2, 3):
        h = act(h @ p[f"W{l}"].view(H, H).T + p[f"b{l}"])
    z4 = h @ p["W4"].view(H, H).T + p["b4"]
    y = act(z4) @ p["W5"] + p["b5"]
    loss = torch.mean((y - torch.tensor(t)) ** 2)
    loss.backward()
    ref = tw.grad.numpy().astype(np.float64)
    g = g.astype(np.float64)
    cos = np.dot(ref, g) / (np.linalg.norm(ref) * np.linalg.norm(g))
    assert cos > 0.995, cos

    def slice_ok(gg, name):
        a, b = SLICES[name]
        ga, gr = gg[a:b], ref[a:b]
        c = np.dot(ga, gr) / (np.linalg.norm(ga) * np.linalg.norm(gr))
        ratio = np.linalg.norm(ga) / np.linalg.norm(gr)
        return c, ratio, c > 0.999 and 0.98 < ratio < 1.02

    for name in SLICES:
        c, ratio, ok = slice_ok(g, name)
        _say(f"  {name}: cos {c:.5f}, norm ratio {ratio:.4f}")
        assert ok, name
    # the check rejects the mistakes it is there for: a bias gradient twice too large,
    # and another layer's bias gradient in its place
    bad = g.copy()
    bad[kB2:kW3] *= 2.0
    assert not slice_ok(bad, "b2")[2]
    bad = g.copy()
    bad[kB3:kW4] = g[kB4:kW5]
    assert not slice_ok(bad, "b3")[2]


def test_baseline_matches_numpy_restatement(ctx):
    """The device's baseline premium (Ju & Zhong's correction of Barone-Adesi & Whaley)
    against tests/baw_ref.py, including the no-early-exercise cases (puts at r = 0,
    calls at q = 0)."""
    from baw_ref import baseline_premium
    o = W.surrogate_train(4096, 12)
    o["r"][:64] = 0.0     # puts: no premium; calls with q > 0: the r -> 0 limit
    o["q"][64:128] = 0.0  # calls: no premium
    got = kwgpu.surrogate_baseline(ctx, o)
    ref = baseline_premium(o)
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - ref) / o["k"]) < 1e-9
    assert np.all(got[:64][o["w"][:64] < 0] == 0) and np.all(got[64:128][o["w"][64:128] > 0] == 0)
    assert np.all(got[:64][(o["w"][:64] > 0) & (o["q"][:64] > 0)] >= 0)


def _fd_premium(ctx, opts, x_dim, t_dim, precision=kwgpu.FP64):
    """FD American price minus BS European price, both on the GPU."""
    import torch
    n = len(opts)
    d_o = torch.from_numpy(opts.view(np.uint8).copy()).cuda()
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_s = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(),
                          kwgpu.params(x_dim, t_dim, precision=precision))
    e = opts.copy()
    e["e"] = 0
    d_e = torch.from_numpy(e.view(np.uint8).copy()).cuda()
    d_b = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.bs_price_device(d_e.data_ptr(), n, d_b.data_ptr())
    ctx.synchronize()
    assert np.all(d_s.cpu().numpy() == 0)
    return d_o, d_p - d_b


def _device(opts):
    import torch
    return torch.from_numpy(np.ascontiguousarray(opts).view(np.uint8).copy()).cuda()


def _predict(s, d_o, n, fd):
    import torch
    d_pm = torch.empty(n, dtype=torch.float64, device="cuda")
    d_sp = torch.empty(n, dtype=torch.float64, device="cuda")
    d_ib = torch.empty(n, dtype=torch.uint8, device="cuda")
    s.predict(d_o.data_ptr(), n, d_pm.data_ptr(), d_sp.data_ptr(), d_ib.data_ptr(), fd)
    s.ctx.synchronize()
    return d_pm.cpu().numpy(), d_sp.cpu().numpy(), d_ib.cpu().numpy()


def test_training_learns_the_premium(ctx):
    """A short fit to the FD pricer on a 256 x 256 grid (labels at the two nodes
    bracketing each option's log(S/K)) lowers the loss, and the surrogate prices held-out
    options at their own spots far better than the premiums' spread. With as many steps
    as x-points the strike's kink is damped (n / mu >= 10.9) for every option of the box."""
    _say("training")
    o = W.surrogate_train(40000, 31)
    fd = kwgpu.params(256, 256)
    n_tr = 32768
    d_o = _device(o)
    s = kwgpu.Surrogate(ctx, members=2, seed=3)
    loss, pairs = s.fit(d_o.data_ptr(), n_tr, fd, epochs=12, batch=2048, lr=2e-3)
    _say("  loss per epoch", loss[:, 0], "pairs", pairs)
    assert pairs >= n_tr - 16  # the box's far corner can put a spot in the first cell
    assert np.all(np.isfinite(loss))
    assert loss[-1].max() < loss[0].min() / 4
    te = o[n_tr:]
    n_te = len(te)
    pm, sp, ib = _predict(s, d_o[n_tr * 56:], n_te, fd)
    fdp, st, rc, _ = ctx.fd1d_price(te, 256, 256)
    e = te.copy()
    e["e"] = 0
    bs = ctx.bs_price(e)
    ok = (ib == 1) & (st == 0)
    assert ok.mean() > 0.99
    truth = fdp - bs
    rms = np.sqrt(np.mean((pm[ok] - truth[ok]) ** 2))
    base = np.sqrt(np.mean((truth[ok] - truth[ok].mean()) ** 2))
    _say(f"  held-out rms {rms:.3e} vs spread of the premiums {base:.3e}")
    assert rms < base / 10
    assert np.all(sp[ok] >= 0)
    # another configuration is refused
    with pytest.raises(kwgpu.KwgpuError):
        _predict(s, d_o[n_tr * 56:], n_te, kwgpu.params(512, 256))
    s.close()


def test_node_form_matches_fd_interpolation(ctx):
    """The surrogate forms a price from its two node values as the FD pricer does
    (src/Math/kwFd1d.cpp:139-158). With the network's weights zeroed except the output
    bias, the premium at every node is base + (base + 1e-4 K) b5 (base: the control
    variate, the premium from the exercise boundary), so the price is BS plus that,
    interpolated with the reference's weights at the bracketing nodes (computed here
    from the sinh mesh) -- to rounding."""
    from baw_ref import boundary_premium
    from node_ref import DAMP_MIN, bracket, kink_damping, nodes
    o = W.surrogate_train(2000, 77)
    fd = kwgpu.params(512, 256)
    s = kwgpu.Surrogate(ctx, members=1, seed=1)
    d_o = _device(o)
    s.fit(d_o.data_ptr(), 2000, fd, epochs=1, batch=1024)  # records the configuration
    w = np.zeros(kwgpu.SURROGATE_PARAMS, np.float32)
    b5 = 0.37
    w[-1] = b5
    s.set_params(0, w)
    pm, sp, ib = _predict(s, d_o, len(o), fd)
    ok = ib == 1
    # bracketing nodes on the 512-point sinh mesh (first node >= log(S/K))
    xs = np.log(o["s"] / o["k"])
    g = nodes(o, 512)
    jj = bracket(g, xs)
    # in the box: a labelled bracket and a damped kink (at 512 x 256 the long, volatile
    # options are not: n / mu falls to ~2.7 at sigma = 0.6, T = 5)
    damp = kink_damping(o, 512, 256)
    want_ok = (jj >= 2) & (jj < 512) & (damp >= DAMP_MIN)
    assert 0.3 < want_ok.mean() < 0.99
    near = np.abs(damp / DAMP_MIN - 1) < 1e-9  # rounding may decide these either way
    assert np.all((ok == want_ok) | near) and np.all(sp[ok] == 0)
    x0, x1 = g[np.arange(len(o)), jj - 1], g[np.arange(len(o)), np.minimum(jj, 511)]
    a, b = o.copy(), o.copy()
    a["s"], b["s"] = o["k"] * np.exp(x0), o["k"] * np.exp(x1)
    ea, eb, es = a.copy(), b.copy(), o.copy()
    for q in (ea, eb, es):
        q["e"] = 0
    pa, pb = boundary_premium(a), boundary_premium(b)
    va = ctx.bs_price(ea) + pa + (pa + 1e-4 * o["k"]) * b5
    vb = ctx.bs_price(eb) + pb + (pb + 1e-4 * o["k"]) * b5
    want = ((x1 - xs) * va + (xs - x0) * vb) / (x1 - x0) - ctx.bs_price(es)
    d = np.abs(pm[ok] - want[ok])
    _say("  node form vs restatement: max", d.max())
    # the device's control variate and its restatement agree to rounding, scaled here by
    # 1 + b5; a wrong weight or node would be off by 1e-4 to 1e-2
    assert d.max() <= 2e-9 * (1 + b5) * o["k"].max()
    s.close()


def oracle_forward(w, x):
    """The network's output y for features x: oracle_loss_grad's forward pass."""
    p = unpack(w)
    h = bf16(silu(x.astype(np.float64) @ p["W1"].T + p["b1"]))
    for l in (2, 3):
        h = bf16(silu(h @ bf16(p[f"W{l}"]).T + p[f"b{l}"]))
    return silu(h @ bf16(p["W4"]).T + p["b4"]) @ p["W5"] + p["b5"]


def test_prediction_matches_forward_oracle(ctx):
    """Prediction runs the whole network in one kernel per 128-row tile and member. Its
    premium and spread against the float64 forward restatement (BF16 rounding where
    the kernels round) at each option's two bracketing nodes, combined as the FD pricer
    interpolates: two members with different random weights, and 2 x 1500 node rows, so
    the last tile is partial. A wrong member or layer offset, a transposed weight or a
    dropped bias moves the network's share of the premium by far more than 3e-3 of
    itself. FP32 accumulation against float64 decides some BF16 rounding ties the
    other way, and those ties compound over three layers: measured up to 8e-4."""
    from baw_ref import boundary_premium
    from node_ref import bracket, nodes
    o = W.surrogate_train(1500, 79)
    fd = kwgpu.params(512, 256)
    s = kwgpu.Surrogate(ctx, members=2, seed=5)
    d_o = _device(o)
    s.fit(d_o.data_ptr(), len(o), fd, epochs=1, batch=1024)  # records the configuration
    rng = np.random.default_rng(21)
    ws = []
    for m in range(2):
        w = s.get_params(m)
        w[kW5:kB5] = rng.uniform(-0.3, 0.3, H).astype(np.float32)
        w[kB5] = rng.uniform(-0.5, 0.5)
        for b0, b1 in ((kB1, kW2), (kB2, kW3), (kB3, kW4), (kB4, kW5)):
            w[b0:b1] = rng.uniform(-0.1, 0.1, b1 - b0).astype(np.float32)
        s.set_params(m, w)
        ws.append(w)
    pm, sp, ib = _predict(s, d_o, len(o), fd)
    ok = ib == 1
    assert ok.mean() > 0.3
    xs = np.log(o["s"] / o["k"])
    g = nodes(o, 512)
    jj = np.minimum(bracket(g, xs), 511)
    r = np.arange(len(o))
    x0, x1 = g[r, jj - 1], g[r, jj]
    a, b = o.copy(), o.copy()
    a["s"], b["s"] = o["k"] * np.exp(x0), o["k"] * np.exp(x1)
    ea, eb, es = a.copy(), b.copy(), o.copy()
    for q in (ea, eb, es):
        q["e"] = 0
    ba, bb = boundary_premium(a), boundary_premium(b)
    sa, sb = ba + 1e-4 * o["k"], bb + 1e-4 * o["k"]
    fa, fb = kwgpu.surrogate_features(a, 512, 256), kwgpu.surrogate_features(b, 512, 256)
    v, net = [], np.zeros(len(o))
    for w in ws:
        ya, yb = oracle_forward(w, fa), oracle_forward(w, fb)
        va = ctx.bs_price(ea) + ba + sa * ya
        vb = ctx.bs_price(eb) + bb + sb * yb
        v.append(((x1 - xs) * va + (xs - x0) * vb) / (x1 - x0))
        net += np.abs(sa * ya) + np.abs(sb * yb)
    want_pm = 0.5 * (v[0] + v[1]) - ctx.bs_price(es)
    want_sp = np.abs(v[0] - v[1])
    tol = 3e-3 * net + 1e-9 * o["k"]
    d_pm, d_sp = np.abs(pm - want_pm)[ok], np.abs(sp - want_sp)[ok]
    _say(f"  fused prediction vs oracle: premium {np.max(d_pm / tol[ok]):.2e} of the "
         f"tolerance, spread {np.max(d_sp / tol[ok]):.2e}; spread rms {np.sqrt(np.mean(want_sp[ok] ** 2)):.2e}")
    assert np.all(d_pm <= tol[ok]) and np.all(d_sp <= tol[ok])
    # the members differ, so the spread is not the trivial 0
    assert np.median(want_sp[ok]) > 30 * np.median(tol[ok])
    s.close()


def test_serving_routes_and_fallback(ctx, tmp_path):
    import torch
    _say("serving")
    o = W.serving(4096, 41)
    n = len(o)
    d_o = torch.from_numpy(o.view(np.uint8).copy()).cuda()
    s = kwgpu.Surrogate(ctx, members=2, seed=4)
    fd = kwgpu.params(256, 64)
    s.fit(d_o.data_ptr(), n, fd, epochs=1, batch=1024)  # the configuration it emulates
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_r = torch.empty(n, dtype=torch.int32, device="cuda")
    # nothing passes a negative spread bound: every option is priced by FD
    nfd = s.price(d_o.data_ptr(), n, -1.0, fd, d_p.data_ptr(), d_r.data_ptr())
    assert nfd == n
    route = d_r.cpu().numpy()
    assert np.all(route >= 1)
    p_all_fd = d_p.cpu().numpy()
    ref, _, rc, _ = ctx.fd1d_price(o, 256, 64, precision=kwgpu.FP64)
    assert rc == 0
    # an FP64 fallback is the FP64 pricer itself
    assert np.max(np.abs(p_all_fd - ref)) <= 1e-8
    # an FP32-mixed fallback stays within its 1e-4 bar, on a smaller batch that reuses
    # the buffers the first call grew
    half = n // 2
    fd32 = kwgpu.params(256, 64, precision=kwgpu.FP32_MIXED)
    nfd = s.price(d_o.data_ptr(), half, -1.0, fd32, d_p.data_ptr(), d_r.data_ptr())
    assert nfd == half
    assert np.max(np.abs(d_p.cpu().numpy()[:half] - ref[:half])) <= 1e-4
    # everything in the box passes an infinite bound; out-of-box options still go to FD
    from node_ref import DAMP_MIN, kink_damping
    feats = kwgpu.surrogate_features(o, 256, 64)
    _, _, ib = _predict(s, d_o, n, fd)
    in_box = ib == 1
    # the option's own box, and a damped kink at the strike (64 steps on 256 points leave
    # the longer, more volatile options undamped)
    want = np.all(np.abs(feats[:, :6]) <= 1 + 1e-9, axis=1) & (kink_damping(o, 256, 64) >= DAMP_MIN)
    assert np.all(want[in_box]) and in_box.sum() >= want.sum() - 4 and 0 < in_box.sum() < n
    nfd = s.price(d_o.data_ptr(), n, 1e300, fd, d_p.data_ptr(), d_r.data_ptr())
    route = d_r.cpu().numpy()
    assert nfd == int((~in_box).sum())
    assert np.array_equal(route == 0, in_box)
    p = d_p.cpu().numpy()
    assert np.max(np.abs(p[~in_box] - ref[~in_box])) <= 1e-4
    # save / load gives the same predictions
    path = str(tmp_path / "sur.bin")
    s.save(path)
    s2 = kwgpu.Surrogate.load(ctx, path)
    d_p2 = torch.empty(n, dtype=torch.float64, device="cuda")
    s2.price(d_o.data_ptr(), n, 1e300, fd, d_p2.data_ptr(), d_r.data_ptr())
    assert np.array_equal(d_p2.cpu().numpy()[in_box], p[in_box])
    # another FD configuration than the emulated one: nothing is served
    fd2 = kwgpu.params(512, 64)
    assert s.price(d_o.data_ptr(), n, 1e300, fd2, d_p.data_ptr(), d_r.data_ptr()) == n
    s.close()
    s2.close()


def test_gated_serving_by_class(ctx):
    """kwgpu_surrogate_price_gated: with one class it is kwgpu_surrogate_price bit for bit;
    with classes, an in-box option is served exactly when its class is open (infinite
    bounds there, closed elsewhere), classes assigned as kwgpu.gate_class assigns them; a
    malformed gate is refused."""
    import torch
    _say("serving by class")
    o = W.serving(4096, 47)
    n = len(o)
    d_o = _device(o)
    s = kwgpu.Surrogate(ctx, members=2, seed=6)
    fd = kwgpu.params(256, 256)
    s.fit(d_o.data_ptr(), n, fd, epochs=1, batch=1024)  # the configuration it emulates
    bufs = [(torch.empty(n, dtype=torch.float64, device="cuda"),
             torch.empty(n, dtype=torch.int32, device="cuda")) for _ in range(3)]
    (p1, r1), (p2, r2), (p3, r3) = bufs
    nfd1 = s.price(d_o.data_ptr(), n, 1e300, fd, p1.data_ptr(), r1.data_ptr(), 1e300)
    g1 = kwgpu.gate([], [], [1e300] * 2, [1e300] * 2)
    nfd2 = s.price_gated(d_o.data_ptr(), n, g1, fd, p2.data_ptr(), r2.data_ptr())
    a1, a2 = p1.cpu().numpy(), p2.cpu().numpy()
    assert nfd1 == nfd2 and np.array_equal(a1, a2, equal_nan=True)
    assert np.array_equal(r1.cpu().numpy(), r2.cpu().numpy())
    te, me = [0.5, 2.0], [0.0]
    open_ = np.arange(12) % 2 == 1  # calls or puts, every T band, S >= K only
    g = kwgpu.gate(te, me, np.where(open_, 1e300, -1.0), np.full(12, 1e300))
    _, _, ib = _predict(s, d_o, n, fd)
    want = (ib == 1) & open_[kwgpu.gate_class(o, te, me)]
    assert 0 < want.sum() < (ib == 1).sum()
    nfd = s.price_gated(d_o.data_ptr(), n, g, fd, p3.data_ptr(), r3.data_ptr())
    route = r3.cpu().numpy()
    assert np.array_equal(route == 0, want) and nfd == n - int(want.sum())
    assert np.array_equal(p3.cpu().numpy()[want], a1[want])
    bad = kwgpu.gate([2.0, 0.5], me, [1.0] * 12, [1.0] * 12)  # edges not increasing
    with pytest.raises(kwgpu.KwgpuError):
        s.price_gated(d_o.data_ptr(), n, bad, fd, p3.data_ptr(), r3.data_ptr())
    bad = kwgpu.gate([], [], [1.0] * 2, [1.0] * 2)
    bad.n_t = kwgpu.GATE_MAX_T + 1
    with pytest.raises(kwgpu.KwgpuError):
        s.price_gated(d_o.data_ptr(), n, bad, fd, p3.data_ptr(), r3.data_ptr())
    s.close()


def test_serving_per_option_grids(ctx):
    """The FD fallback solves each routed option on its own grid: the compacted subset
    carries its options' sizes, as the uneven-batch pricer takes them."""
    import torch
    _say("serving, per-option grids")
    o = W.serving(3000, 43)
    n = len(o)
    rng = np.random.default_rng(43)
    xd = rng.choice(np.array([64, 128, 256, 512], np.int32), n)
    td = np.minimum(W.t_points_for(o["t"]), 160).astype(np.int32)
    d_o = torch.from_numpy(o.view(np.uint8).copy()).cuda()
    d_x = torch.from_numpy(xd.copy()).cuda()
    d_t = torch.from_numpy(td.copy()).cuda()
    s = kwgpu.Surrogate(ctx, members=2, seed=5)
    s.fit(d_o.data_ptr(), n, kwgpu.params(512, 512, x_dims=d_x.data_ptr(), t_dims=d_t.data_ptr()),
          epochs=1, batch=1024)  # a surrogate of per-option grids
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_r = torch.empty(n, dtype=torch.int32, device="cuda")
    ref, st, rc, _ = ctx.fd1d_price(o, 512, 512, x_dims=xd, t_dims=td, precision=kwgpu.FP64)
    assert rc == 0 and np.all(st == 0)
    # a uniform grid of any of the sizes differs from the per-option prices
    uni, _, _, _ = ctx.fd1d_price(o, 256, 160, precision=kwgpu.FP64)
    assert np.max(np.abs(uni - ref)) > 1e-3
    for prec, tol in ((kwgpu.FP64, 1e-8), (kwgpu.FP32_MIXED, 1e-4)):
        fd = kwgpu.params(512, 512, precision=prec, x_dims=d_x.data_ptr(), t_dims=d_t.data_ptr())
        # every option to FD
        assert s.price(d_o.data_ptr(), n, -1.0, fd, d_p.data_ptr(), d_r.data_ptr()) == n
        assert np.max(np.abs(d_p.cpu().numpy() - ref)) <= tol
        # only the out-of-box options to FD: a scattered subset keeps its own sizes
        _, _, ib = _predict(s, d_o, n, fd)
        out = ib == 0
        assert 0 < out.sum() < n
        nfd = s.price(d_o.data_ptr(), n, 1e300, fd, d_p.data_ptr(), d_r.data_ptr())
        assert nfd == int(out.sum())
        assert np.array_equal(d_r.cpu().numpy() != 0, out)
        p1 = d_p.cpu().numpy()
        assert np.max(np.abs(p1[out] - ref[out])) <= tol
        # the same routing gives the same prices again (the subset is compacted in order)
        assert s.price(d_o.data_ptr(), n, 1e300, fd, d_p.data_ptr(), d_r.data_ptr()) == nfd
        assert np.array_equal(d_p.cpu().numpy(), p1, equal_nan=True)
    s.close()
