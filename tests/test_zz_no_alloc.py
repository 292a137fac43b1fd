"""kwgpu_fd1d_price_device_noalloc (the entry point that allocates nothing, its batch state
in the library's fixed device scratch) against kwgpu_fd1d_price_device on the same
inputs: bit for bit, on two calls in a row with the second smaller, for per-option and
uniform grids in both precisions, plus an oracle sample."""
from __future__ import annotations

import numpy as np
import pytest

import kwgpu
import oracle
from kwgpu import workloads as W

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).cuda()


def _run(ctx, fn, d_o, n, prm):
    import torch
    d_p = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    d_s = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    fn(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
    ctx.synchronize()
    return d_p.cpu().numpy(), d_s.cpu().numpy()


@pytest.mark.parametrize("precision", [kwgpu.FP64, kwgpu.FP32_MIXED], ids=["fp64", "mixed"])
def test_noalloc_matches_device_entry_point(ctx, precision):
    import torch
    print("noalloc, precision", precision, flush=True)
    o, xd, td = W.uneven(12000, 61)
    td = np.minimum(td, 240).astype(np.int32)
    d_o = _dev(o)
    d_x = torch.from_numpy(xd.copy()).cuda()
    d_t = torch.from_numpy(td.copy()).cuda()
    n = len(o)
    ref, rst = _run(ctx, ctx.fd1d_price_device, d_o, n,
                    kwgpu.params(precision=precision, x_dims=d_x.data_ptr(), t_dims=d_t.data_ptr()))
    assert np.all(rst == 0)
    for m in (n, 5000):  # two calls in a row, the second smaller
        hx, ht = np.ascontiguousarray(xd[:m]), np.ascontiguousarray(td[:m])
        p, st = _run(ctx, ctx.fd1d_price_device_noalloc, d_o, m,
                     kwgpu.params(precision=precision, x_dims=hx.ctypes.data, t_dims=ht.ctypes.data))
        assert np.array_equal(st, rst[:m])
        assert np.array_equal(p, ref[:m])
    if precision == kwgpu.FP64:
        idx = np.arange(3, n, n // 24)
        want, wst = oracle.fd1d_price_uneven(o[idx], xd[idx], td[idx])
        assert np.all(wst == 0) and np.max(np.abs(p[idx[idx < 5000]] - want[idx < 5000])) <= 1e-8
    # one grid for all
    u = W.plain(3000, 62)
    d_u = _dev(u)
    prm = kwgpu.params(1024, 64, precision=precision)
    ref_u, rst_u = _run(ctx, ctx.fd1d_price_device, d_u, 3000, prm)
    for m in (3000, 1000):
        p, st = _run(ctx, ctx.fd1d_price_device_noalloc, d_u, m, prm)
        assert np.array_equal(st, rst_u[:m]) and np.array_equal(p, ref_u[:m])


def test_noalloc_refuses_what_it_cannot_hold(ctx):
    o = W.plain(16, 63)
    d_o = _dev(o)
    with pytest.raises(kwgpu.KwgpuError) as e:
        _run(ctx, ctx.fd1d_price_device_noalloc, d_o, 16, kwgpu.params(3000, 32))
    assert e.value.code == kwgpu.ERR_UNSUPPORTED
    with pytest.raises(kwgpu.KwgpuError) as e:
        _run(ctx, ctx.fd1d_price_device_noalloc, d_o, (1 << 20) + 1, kwgpu.params(512, 32))
    assert e.value.code == kwgpu.ERR_INVALID
