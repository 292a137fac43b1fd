"""Parity of the HIP FD pricer (through the C-ABI) with the reference and the oracle.

FP64 bar: max abs price difference <= 1e-8 against the reference CPU pricer's
golden vectors and against the oracle on the same seeded inputs. Full-size runs
(32768 options at 1024 x 1024) are checked through the golden head and through
properties that hold at any size: American >= European >= 0, American >= intrinsic,
determinism, every status OK.
"""
from __future__ import annotations

import numpy as np
import pytest

import kwgpu
import oracle
from conftest import golden_meta, load_golden
from kwgpu import workloads as W

pytestmark = pytest.mark.gpu

EPS = 1e-8
LAYOUTS = [kwgpu.LAYOUT_WAVE_REG, kwgpu.LAYOUT_LANE_SOA]
LAYOUT_IDS = ["wave", "lane"]


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
def test_plain_512_golden(ctx, layout):
    _say("plain_512 layout", layout)
    o, c = load_golden("plain_512")
    p, st, rc, err = ctx.fd1d_price(o, 512, 512, layout=layout)
    assert rc == kwgpu.OK, err
    assert np.all(st == kwgpu.STATUS_OK)
    assert np.max(np.abs(p - c["ref_fd1d"])) <= EPS
    e = o.copy()
    e["e"] = 0
    p, st, rc, err = ctx.fd1d_price(e, 512, 512, layout=layout)
    assert rc == kwgpu.OK, err
    assert np.max(np.abs(p - c["ref_fd1d_euro"])) <= EPS


@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
def test_plain_1024_head_golden(ctx, layout):
    _say("plain_1024_head256 layout", layout)
    o, c = load_golden("plain_1024_head256")
    p, st, rc, err = ctx.fd1d_price(o, 1024, 1024, layout=layout)
    assert rc == kwgpu.OK, err
    assert np.max(np.abs(p - c["ref_fd1d"])) <= EPS


def test_pricer_test_12(ctx):
    """The three pricers of test/kwPricer_test.cpp:37-108 through the GPU."""
    _say("pricer_test_12")
    o, c = load_golden("pricer_test_12")
    fd, st, rc, err = ctx.fd1d_price(o)
    assert rc == kwgpu.OK, err
    assert np.max(np.abs(fd - c["ref_fd1d"])) <= EPS
    assert np.max(np.abs(fd - c["want"])) <= 1.3e-3
    bs = ctx.bs_price(o)
    ok = ~np.isnan(c["ref_bs"])
    assert np.array_equal(np.isnan(bs), ~ok)
    assert np.max(np.abs(bs[ok] - c["ref_bs"][ok])) <= 1e-12
    fdbs, st, rc, err = ctx.fd1d_bs_price(o)
    assert rc == kwgpu.OK, err
    assert np.max(np.abs(fdbs - c["ref_fd1d_bs"])) <= EPS
    assert np.max(np.abs(fdbs - c["want"])) <= 1.3e-3


def test_fd1d_bs_plain_512(ctx):
    _say("fd1d_bs plain_512")
    o, c = load_golden("plain_512")
    p, st, rc, err = ctx.fd1d_bs_price(o, 512, 512)
    assert rc == kwgpu.OK, err
    assert np.max(np.abs(p - c["ref_fd1d_bs"])) <= EPS
    e = o.copy()
    e["e"] = 0
    assert np.max(np.abs(ctx.bs_price(e) - c["ref_bs_euro"])) <= 1e-12


def test_fd1d_bs_uneven_grids(ctx):
    """FD1D-BS with per-option grid sizes: both FD legs go through one solve of 2n
    options, so each option's sizes must reach its European copy too; the oracle prices
    each (x, t) group as its own FD1D-BS call (src/Pricer/kwFd1d_BlackScholes.cpp:15-43)."""
    _say("fd1d_bs uneven")
    o = W.plain(90, 23)
    xd = np.array([256, 300, 512], np.int32)[np.arange(len(o)) % 3]
    td = np.array([40, 64], np.int32)[np.arange(len(o)) % 2]
    p, st, rc, err = ctx.fd1d_bs_price(o, x_dims=xd, t_dims=td)
    assert rc == kwgpu.OK, err
    assert np.all(st == 0)
    want = np.full(len(o), np.nan)
    for x in np.unique(xd):
        for t in np.unique(td):
            sel = (xd == x) & (td == t)
            if sel.any():
                want[sel], e2 = oracle.fd1d_bs_price(o[sel], int(x), int(t))
                assert e2 == ""
    assert np.max(np.abs(p - want)) <= EPS


@pytest.mark.parametrize("precision", ["fp64", "uneven"])
def test_greeks_match_oracle(ctx, precision):
    """Bump Greeks (test/portfolio.py:35-56) through the C-ABI against the oracle's: each
    Greek is a fixed linear combination of nine prices, so its tolerance is that
    combination applied to the price tolerance 1e-8 max(1, |V|)."""
    _say("greeks", precision)
    o = W.plain(96, 17)
    b = 0.01
    if precision == "uneven":
        xd = np.where(np.arange(len(o)) % 2, 256, 512).astype(np.int32)
        td = np.where(np.arange(len(o)) % 3, 128, 200).astype(np.int32)
        g, st, rc, err = ctx.fd1d_greeks(o, 0, 0, b, x_dims=xd, t_dims=td)
        sets = oracle.bump_sets(o, b)
        v = np.stack([oracle.fd1d_price_uneven(s_, xd, td)[0] for s_ in sets])
        ref = oracle.greeks_from_prices(o, v, b)
    else:
        g, st, rc, err = ctx.fd1d_greeks(o, 512, 512, b)
        ref, _, v = oracle.fd1d_greeks(o, 512, 512, b)
    assert rc == kwgpu.OK, err
    assert np.all(st == 0)
    ptol = EPS * np.maximum(1.0, np.abs(v).max(axis=0))
    tol = {"price": ptol,
           "delta": 2 * ptol / (2 * b * o["s"]),
           "gamma": 4 * ptol / (b * o["s"]) ** 2,
           "vega": 2 * ptol / (2 * b * o["z"]) / 100,
           "theta": 2 * ptol / (2 * b * o["t"]) / 360,
           "rho": 2 * ptol / (2 * b * o["r"]) / 100}
    for f, tl in tol.items():
        assert np.all(np.abs(g[f] - ref[f]) <= tl), f


def test_greeks_bad_bump_and_range(ctx):
    o = W.plain(4, 18)
    _, _, rc, err = ctx.fd1d_greeks(o, 256, 64, 0.0)
    assert rc == kwgpu.ERR_INVALID and "bump" in err
    # a spot at the grid's upper end: the up-bump leaves the grid
    o["s"] = o["k"] * np.exp(50 * o["z"] * np.sqrt(o["t"])) * 0.999
    _, st, rc, err = ctx.fd1d_greeks(o, 256, 64, 0.01)
    assert rc == kwgpu.ERR_RANGE and "not in range" in err
    assert np.all(st == kwgpu.STATUS_OUT_OF_GRID)


def test_qdfp_portfolio(ctx):
    """test/kwPortfolio_test.cpp:29-58 through the GPU: 6000 puts (600 PDEs after the
    reference's compression) within 0.005 of QuantLib, and at parity with the
    reference FD1D output."""
    _say("qdfp portfolio")
    o, c = load_golden("portfolio_qdfp")
    p, st, rc, err = ctx.fd1d_price(o, 512, 512)
    assert rc == kwgpu.OK, err
    assert np.max(np.abs(p - c["want"])) <= 0.005
    assert np.max(np.abs(p - c["ref_fd1d"])) <= EPS


def test_uneven_small_golden(ctx):
    _say("uneven_small")
    o, c = load_golden("uneven_small")
    p, st, rc, err = ctx.fd1d_price(o, x_dims=c["x_dim"], t_dims=c["t_dim"])
    assert rc == kwgpu.OK, err
    assert np.max(np.abs(p - c["ref_fd1d"])) <= EPS


@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
@pytest.mark.parametrize("x_dim,t_dim", [(3, 2), (17, 5), (64, 37), (100, 64), (129, 33),
                                         (256, 40), (300, 16), (511, 24), (1000, 12),
                                         (1024, 20)])
def test_sizes_vs_oracle(ctx, layout, x_dim, t_dim):
    _say("sizes", layout, x_dim, t_dim)
    o = W.plain(37, 1000 + x_dim)
    o["e"][::4] = 0
    o["s"] = np.linspace(80, 125, 37)
    ref, rst = oracle.fd1d_price_status(o, x_dim, t_dim)
    p, st, rc, err = ctx.fd1d_price(o, x_dim, t_dim, layout=layout)
    assert np.array_equal(st, rst)
    ok = rst == 0
    # absolute 1e-8; grids as coarse as 3 points price calls at ~1e16 (e^x at the far
    # node), where the same bar applies relative to the price
    tol = EPS * np.maximum(1.0, np.abs(ref[ok]))
    assert np.all(np.abs(p[ok] - ref[ok]) <= tol)
    assert np.all(np.isnan(p[~ok]))


def test_range_error_like_reference(ctx):
    _say("range error")
    m = golden_meta()["range_error"]
    o = W.plain(4, m["seed"])
    o["s"] = m["spots"]
    p, st, rc, err = ctx.fd1d_price(o, m["x_dim"], m["t_dim"], density=m["density"],
                                    scale=m["scale"])
    assert rc == kwgpu.ERR_RANGE
    assert err == m["message"]
    ref, rst = oracle.fd1d_price_status(o, m["x_dim"], m["t_dim"], m["density"], m["scale"])
    assert np.array_equal(st, rst)
    ok = rst == 0
    assert np.max(np.abs(p[ok] - ref[ok]), initial=0.0) <= EPS


def test_bad_inputs_and_params(ctx):
    _say("bad inputs")
    o = W.plain(8, 31)
    o["z"][2] = -0.2
    o["t"][5] = 0.0
    p, st, rc, err = ctx.fd1d_price(o, 128, 32)
    assert rc == kwgpu.ERR_INVALID
    assert st[2] == kwgpu.STATUS_BAD_INPUT and st[5] == kwgpu.STATUS_BAD_INPUT
    good = np.ones(8, bool)
    good[[2, 5]] = False
    ref, _ = oracle.fd1d_price_status(o[good], 128, 32)
    assert np.max(np.abs(p[good] - ref)) <= EPS
    for bad in (dict(x_dim=2), dict(t_dim=1), dict(density=0.0), dict(theta=0.0)):
        kw = dict(x_dim=128, t_dim=32)
        kw.update(bad)
        _, _, rc, err = ctx.fd1d_price(o, **kw)
        assert rc == kwgpu.ERR_INVALID and err
    _, _, rc, _ = ctx.fd1d_price(np.zeros(0, W.OPTION_DTYPE))
    assert rc == kwgpu.OK


@pytest.mark.parametrize("x_dim", [100, 300, 500, 1000, 2000])
def test_neighbours_of_bad_options_unaffected(ctx, x_dim):
    """Options sharing a wave with an invalid option, with one whose march overflows
    (its far call payoff is inf) and with an empty segment (odd count) are priced as if
    alone: nothing crosses from one option's lanes into another's."""
    for n in (7, 8):  # the longest-dated (overflowing) option last: with 8 it shares a
        _say("neighbours", x_dim, n)  # wave with a valid option, with 7 with nothing
        o = W.plain(n, 61)
        o["s"] = 100
        o["z"][1] = -0.2                                      # invalid
        o["w"][-1], o["t"][-1], o["z"][-1] = 1, 30.0, 3.0     # e^(50 sigma sqrt T) = inf
        ref, _ = oracle.fd1d_price_status(o, x_dim, 20)
        good = np.ones(n, bool)
        good[[1, n - 1]] = False
        for lay in LAYOUTS:
            p, st, rc, err = ctx.fd1d_price(o, x_dim, 20, layout=lay)
            assert st[1] == kwgpu.STATUS_BAD_INPUT
            assert st[-1] == kwgpu.STATUS_NONFINITE or not np.isfinite(p[-1])
            assert np.all(st[good] == 0)
            assert np.max(np.abs(p[good] - ref[good])) <= EPS, lay


def _extreme_options():
    """Every combination of extreme but valid inputs: vol 1 % to 300 %, maturity
    0.002 to 30 years, negative to 25 % rates, dividend yields to 30 %, spot from a
    quarter to four times the strike, puts and calls, American and European."""
    import itertools
    rows = list(itertools.product([0.01, 0.05, 1.5, 3.0], [0.002, 0.05, 10.0, 30.0],
                                  [-0.03, 0.0, 0.25], [0.0, 0.3], [0.25, 0.9, 1.0, 1.1, 4.0],
                                  [-1, 1], [0, 1]))
    o = np.zeros(len(rows), dtype=kwgpu.OPTION_DTYPE)
    for f, col in zip(["z", "t", "r", "q", "s", "w", "e"], zip(*rows)):
        o[f] = col
    o["k"] = 100.0
    o["s"] *= 100.0
    return o


@pytest.mark.parametrize("x_dim", [64, 300, 1024, 2048])
def test_extreme_inputs_match_oracle(ctx, x_dim):
    """Extreme inputs through each kernel geometry (4 options per wave at 64 points, 2 at
    300, one per wave at 1024, two waves per option at 2048): where the reference's value
    lookup fails (spot outside the grid) the status says so; where the reference's price
    is finite it is matched to 1e-8 relative; where it is not, the status or the price
    says so too."""
    _say("extreme inputs", x_dim)
    o = _extreme_options()
    ref, ref_st = oracle.fd1d_price_status(o, x_dim, 40)
    p, st, rc, err = ctx.fd1d_price(o, x_dim, 40, layout=kwgpu.LAYOUT_WAVE_REG)
    out = ref_st == 1
    assert 0 < out.sum() < len(o)  # the set reaches both sides of the grid's edge
    assert np.all(st[out] == kwgpu.STATUS_OUT_OF_GRID)
    fin = ~out & np.isfinite(ref)
    assert np.all(st[fin] == kwgpu.STATUS_OK)
    assert np.all(np.abs(p[fin] - ref[fin]) <= EPS * np.maximum(1.0, np.abs(ref[fin])))
    bad = ~out & ~np.isfinite(ref)
    assert np.all((st[bad] == kwgpu.STATUS_NONFINITE) | ~np.isfinite(p[bad]))
    _say("   out of grid", int(out.sum()), "finite", int(fin.sum()), "non-finite", int(bad.sum()))


def test_chain_compression(ctx):
    """Options sharing (t, r, q, z, e, w) solve one PDE and interpolate per spot/strike
    (src/Pricer/kwFd1d.cpp:28-65, :146-157)."""
    _say("chain")
    base = W.plain(5, 41)
    o = np.repeat(base, 13)
    o["s"] = np.tile(np.linspace(70, 140, 13), 5)
    o["k"] = np.tile(np.linspace(90, 110, 13)[::-1], 5)
    ref, err = oracle.fd1d_price(o, 300, 50)
    assert err == ""
    for lay in LAYOUTS:
        p, st, rc, err = ctx.fd1d_price(o, 300, 50, layout=lay)
        assert rc == kwgpu.OK, err
        assert np.max(np.abs(p - ref)) <= EPS


def test_theta_general_matches_reference_at_half(ctx):
    """theta = 1/2 is the reference's scheme; theta = 1 (implicit Euler) is first order in
    time and converges to the same prices as the time grid is refined."""
    o = W.plain(16, 51)
    p1, _, rc, _ = ctx.fd1d_price(o, 256, 64, theta=0.5)
    assert rc == 0
    ref, _ = oracle.fd1d_price(o, 256, 64)
    assert np.max(np.abs(p1 - ref)) <= EPS
    fine, _ = oracle.fd1d_price(o, 256, 2048)
    e64 = np.max(np.abs(ctx.fd1d_price(o, 256, 64, theta=1.0)[0] - fine))
    e1024 = np.max(np.abs(ctx.fd1d_price(o, 256, 1024, theta=1.0)[0] - fine))
    assert e1024 < e64 / 6, (e64, e1024)


@pytest.mark.parametrize("x_dim", [256, 1024, 2048])
def test_general_theta_kernel_agrees_at_half(ctx, x_dim):
    """The FP64 march is built as a theta = 1/2 kernel and a general-theta kernel (chosen
    by the host on theta != 1/2). theta = 1/2 + 1e-12 runs the general kernel, whose
    prices must match the theta = 1/2 kernel's to the FP64 bar: this covers the
    multi-option (256), one-wave (1024) and two-wave (2048) general kernels."""
    _say("general theta", x_dim)
    o = W.plain(48, 61)
    half, st, rc, err = ctx.fd1d_price(o, x_dim, 64)
    assert rc == kwgpu.OK, err
    gen, st2, rc, err = ctx.fd1d_price(o, x_dim, 64, theta=0.5 + 1e-12)
    assert rc == kwgpu.OK, err
    assert np.array_equal(st, st2)
    assert np.max(np.abs(gen - half)) <= EPS


def test_device_path_full_size_properties(ctx):
    """32768 options, 1024 x 1024, device-resident: head equals the golden vectors,
    American >= European >= 0, American >= K x the payoff interpolated between the nodes
    around log(S/K), bitwise deterministic."""
    import torch
    _say("full size 32768 x 1024 x 1024")
    n = 32768
    o = W.plain(n, 1)
    d_o = torch.from_numpy(o.view(np.uint8).copy()).cuda()
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_s = torch.empty(n, dtype=torch.int32, device="cuda")
    prm = kwgpu.params(1024, 1024, layout=kwgpu.LAYOUT_WAVE_REG)
    ctx.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
    ctx.synchronize()
    p = d_p.cpu().numpy().copy()
    assert np.all(d_s.cpu().numpy() == 0)
    _, c = load_golden("plain_1024_head256")
    assert np.max(np.abs(p[:256] - c["ref_fd1d"])) <= EPS
    ctx.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
    ctx.synchronize()
    assert np.array_equal(p, d_p.cpu().numpy())
    e = o.copy()
    e["e"] = 0
    d_e = torch.from_numpy(e.view(np.uint8).copy()).cuda()
    ctx.fd1d_price_device(d_e.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
    ctx.synchronize()
    pe = d_p.cpu().numpy()
    assert np.all(pe >= -1e-12)
    assert np.all(p >= pe - 1e-12)
    # projection holds at the nodes j <= N-2 and the price interpolates linearly between
    # the two nodes around log(S/K), so it is bounded below by K times the payoff
    # interpolated the same way (not by the intrinsic value: the put payoff 1 - e^x is
    # concave, so the chord lies below it)
    lo = K_times_interpolated_payoff(o, 1024)
    assert np.all(p >= lo - 1e-10 * o["k"])
    intrinsic = np.maximum(0, o["w"] * (o["s"] - o["k"]))
    assert np.all(p >= intrinsic - 1e-3 * o["k"])


def K_times_interpolated_payoff(o, x_dim, density=0.25, scale=50.0):
    """K * (linear interpolation of the payoff between the grid nodes around log(S/K)),
    with the grid x_j = density sinh(y_min (1 - j/(N-1)) + y_max j/(N-1)),
    y = asinh(-+scale sigma sqrt(T) / density) and the first node x_j >= log(S/K)
    (src/Pricer/kwFd1d.cpp:105-139, src/Math/kwFd1d.cpp:139-158)."""
    xs = np.log(o["s"] / o["k"])
    half = scale * o["z"] * np.sqrt(o["t"])
    y0, y1 = np.arcsinh(-half / density), np.arcsinh(half / density)

    def node(j):
        u = j / (x_dim - 1.0)
        return density * np.sinh(y0 * (1.0 - u) + y1 * u)

    j = np.ceil((np.arcsinh(xs / density) - y0) / (y1 - y0) * (x_dim - 1)).astype(np.int64)
    j = np.clip(j, 1, x_dim - 1)
    for _ in range(2):  # settle rounding at a node
        j = np.where((j > 1) & (node(j - 1) >= xs), j - 1, j)
        j = np.where((j < x_dim - 1) & (node(j) < xs), j + 1, j)
    xl, xr = node(j - 1), node(j)
    assert np.all((xl < xs) & (xs <= xr))

    def payoff(x):
        return np.maximum(0.0, o["w"] * (np.exp(x) - 1.0))

    f = (xs - xl) / (xr - xl)
    return o["k"] * (payoff(xl) + (payoff(xr) - payoff(xl)) * f)


# ---- FP32-mixed: within 1e-4 of the FP64 reference -------------------------------

MIXED_TOL = 1e-4


def test_mixed_plain_512_golden(ctx):
    _say("mixed plain_512")
    o, c = load_golden("plain_512")
    p, st, rc, err = ctx.fd1d_price(o, 512, 512, precision=kwgpu.FP32_MIXED)
    assert rc == kwgpu.OK, err
    d = np.abs(p - c["ref_fd1d"])
    _say("  max", d.max(), "rms", np.sqrt(np.mean(d * d)))
    assert d.max() <= MIXED_TOL


def test_mixed_plain_1024_head_golden(ctx):
    _say("mixed plain_1024_head256")
    o, c = load_golden("plain_1024_head256")
    p, st, rc, err = ctx.fd1d_price(o, 1024, 1024, precision=kwgpu.FP32_MIXED)
    assert rc == kwgpu.OK, err
    d = np.abs(p - c["ref_fd1d"])
    _say("  max", d.max(), "rms", np.sqrt(np.mean(d * d)))
    assert d.max() <= MIXED_TOL


def test_mixed_uneven_small_golden(ctx):
    _say("mixed uneven_small")
    o, c = load_golden("uneven_small")
    p, st, rc, err = ctx.fd1d_price(o, x_dims=c["x_dim"], t_dims=c["t_dim"],
                                    precision=kwgpu.FP32_MIXED)
    assert rc == kwgpu.OK, err
    assert np.max(np.abs(p - c["ref_fd1d"])) <= MIXED_TOL


@pytest.mark.parametrize("x_dim,t_dim", [(64, 37), (300, 50), (1000, 40), (2048, 30)])
def test_mixed_sizes_vs_oracle(ctx, x_dim, t_dim):
    _say("mixed sizes", x_dim, t_dim)
    o = W.plain(24, 2000 + x_dim)
    o["e"][::5] = 0
    ref, rst = oracle.fd1d_price_status(o, x_dim, t_dim)
    p, st, rc, err = ctx.fd1d_price(o, x_dim, t_dim, precision=kwgpu.FP32_MIXED)
    assert np.array_equal(st, rst)
    ok = rst == 0
    assert np.all(np.abs(p[ok] - ref[ok]) <= MIXED_TOL)  # absolute, as the bar is stated


def test_mixed_validated_domain(ctx):
    """FP32-mixed over a wide random box at 1024 x 1024 (vol 2 %-120 %, maturity
    0.002-15 years, rates -3 %-20 %, yields 0-25 %, spot 0.4-2.5x strike), plus the corner
    beyond the drift limit (vol 1-2 %, 10-30 years, yields 25-30 %). Device entry point
    (one solve per option, as asked): within 1e-4 *relative* to max(1, |FP64 price|)
    wherever both limits hold (prices here reach the hundreds; measured: at most 3.2e-8
    relative); above the overflow limit, calls whose FP32 increments overflow are flagged
    non-finite; past the drift limit every option is flagged KWGPU_STATUS_PRECISION with a
    NaN price; so no price is silently outside the bar. Host entry point: the PDEs beyond
    either limit are solved in FP64, so every option is priced."""
    import torch
    _say("mixed validated domain")
    rng = np.random.default_rng(7)
    n, nc = 16384, 1024
    o = np.zeros(n + nc, dtype=kwgpu.OPTION_DTYPE)
    o["z"][:n] = rng.uniform(0.02, 1.2, n)
    o["t"][:n] = np.exp(rng.uniform(np.log(0.002), np.log(15.0), n))
    o["r"][:n] = rng.uniform(-0.03, 0.2, n)
    o["q"][:n] = rng.uniform(0.0, 0.25, n)
    o["z"][n:] = rng.uniform(0.01, 0.02, nc)
    o["t"][n:] = rng.uniform(10.0, 30.0, nc)
    o["r"][n:] = rng.uniform(-0.03, 0.0, nc)
    o["q"][n:] = rng.uniform(0.25, 0.3, nc)
    n += nc
    o["k"] = 100.0
    o["s"] = 100.0 * np.exp(rng.uniform(np.log(0.4), np.log(2.5), n))
    o["s"][-nc:] = 100.0 * np.exp(rng.uniform(np.log(0.8), np.log(1.25), nc))
    o["w"] = np.where(rng.random(n) < 0.5, -1, 1)
    o["e"] = 1
    p64, s64, _, _ = ctx.fd1d_price(o, 1024, 1024)
    ok64 = (s64 == kwgpu.STATUS_OK) & np.isfinite(p64)
    assert ok64.mean() > 0.9
    half = 50.0 * o["z"] * np.sqrt(o["t"])
    over = ok64 & ~(half <= 80.0)
    drifted = ok64 & ~over & ~(np.abs(o["r"] - o["q"] - 0.5 * o["z"] ** 2) * o["t"] <= half)
    inside = ok64 & ~over & ~drifted
    assert drifted[-nc:].mean() > 0.5 and not drifted[:-nc].any()  # the box is within it

    d_o = torch.from_numpy(o.view(np.uint8).copy()).cuda()
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_s = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(),
                          kwgpu.params(1024, 1024, precision=kwgpu.FP32_MIXED))
    ctx.synchronize()
    p32, s32 = d_p.cpu().numpy(), d_s.cpu().numpy()
    ok32 = (s32 == kwgpu.STATUS_OK) & np.isfinite(p32)
    rel = np.abs(p32 - p64) / np.maximum(1.0, np.abs(p64))
    assert np.all(ok32[inside])
    _say("   inside", int(inside.sum()), "max rel", float(rel[inside].max()), "drifted",
         int(drifted.sum()))
    assert rel[inside].max() <= 1e-4
    flagged = over & ~ok32
    assert flagged.any() and np.all(o["w"][flagged] == 1)
    assert np.all(s32[flagged] == kwgpu.STATUS_NONFINITE)
    assert np.all(s32[drifted] == kwgpu.STATUS_PRECISION) and np.all(np.isnan(p32[drifted]))
    assert np.all(rel[ok64 & ok32] <= 1e-4)

    h32, hs32, rc, err = ctx.fd1d_price(o, 1024, 1024, precision=kwgpu.FP32_MIXED)
    assert np.all(hs32[ok64] == kwgpu.STATUS_OK)
    hrel = np.abs(h32 - p64) / np.maximum(1.0, np.abs(p64))
    assert hrel[inside].max() <= 1e-4
    assert np.all(hrel[over | drifted] <= EPS)  # the FP64 form itself


def test_mixed_uneven_with_fp64_pdes(ctx):
    """An uneven FP32-mixed batch through the host entry point with overflow-prone PDEs
    (scale sigma sqrt(T) > 80) among the others: one plan holds buckets of both
    precisions, those PDEs match the FP64 pricer, the rest stay within 1e-4, and a
    repeated call (a graph replay) gives the same prices."""
    _say("mixed uneven with FP64 PDEs")
    o, xd, td = W.uneven(1500, 83)
    td = np.minimum(td, 60)
    risky = np.arange(len(o)) % 5 == 0
    o["z"][risky] = 1.5
    o["t"][risky] = 10.0
    assert np.all(50.0 * o["z"][risky] * np.sqrt(o["t"][risky]) > 80.0)
    p64, s64, rc, err = ctx.fd1d_price(o, x_dims=xd, t_dims=td)
    assert rc == kwgpu.OK, err
    p32, s32, rc, err = ctx.fd1d_price(o, x_dims=xd, t_dims=td, precision=kwgpu.FP32_MIXED)
    assert rc == kwgpu.OK, err
    assert np.all(s32 == kwgpu.STATUS_OK)
    scale = np.maximum(1.0, np.abs(p64))
    assert np.all(np.abs(p32 - p64)[risky] <= EPS * scale[risky])
    assert np.abs(p32 - p64)[~risky].max() <= MIXED_TOL  # absolute
    again, _, rc, _ = ctx.fd1d_price(o, x_dims=xd, t_dims=td, precision=kwgpu.FP32_MIXED)
    assert np.array_equal(again, p32)


def test_mixed_full_size_per_class(ctx):
    """32768 options at 1024 x 1024: FP32-mixed against the FP64 GPU result (itself at
    parity with the reference), per class put/call x T < 1 / T >= 1."""
    import torch
    _say("mixed full size")
    n = 32768
    o = W.plain(n, 1)
    d_o = torch.from_numpy(o.view(np.uint8).copy()).cuda()
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_s = torch.empty(n, dtype=torch.int32, device="cuda")
    out = {}
    for prec in (kwgpu.FP64, kwgpu.FP32_MIXED):
        prm = kwgpu.params(1024, 1024, precision=prec)
        ctx.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
        ctx.synchronize()
        assert np.all(d_s.cpu().numpy() == 0)
        out[prec] = d_p.cpu().numpy().copy()
    d = np.abs(out[kwgpu.FP32_MIXED] - out[kwgpu.FP64])
    for w in (-1, 1):
        for lo, hi in ((0, 1), (1, 9)):
            m = (o["w"] == w) & (o["t"] >= lo) & (o["t"] < hi)
            _say(f"  w={w:+d} T in [{lo},{hi}): n={m.sum()} max={d[m].max():.3e} "
                 f"rms={np.sqrt(np.mean(d[m] ** 2)):.3e}")
    assert d.max() <= MIXED_TOL


# ---- uneven batches: per-bucket kernels, persistent grids, hipGraph replay -------------

def _uneven_device(o, xd, td):
    import torch
    d_o = torch.from_numpy(o.view(np.uint8).copy()).cuda()
    d_x = torch.from_numpy(xd.astype(np.int32)).cuda()
    d_t = torch.from_numpy(td.astype(np.int32)).cuda()
    return d_o, d_x, d_t


@pytest.mark.parametrize("precision", [kwgpu.FP64, kwgpu.FP32_MIXED], ids=["fp64", "mixed"])
def test_uneven_device_graph_replay(ctx, precision, monkeypatch):
    """Device path with per-option grids: matches the oracle (one reference call per grid
    size), replays its captured graph bitwise, and equals direct launches."""
    import torch
    _say("uneven device", precision)
    o, xd, td = W.uneven(3000, 77)
    td = np.minimum(td, 64)  # keep the oracle check quick; sizes still span all buckets
    d_o, d_x, d_t = _uneven_device(o, xd, td)
    n = len(o)
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_s = torch.empty(n, dtype=torch.int32, device="cuda")
    prm = kwgpu.params(precision=precision, x_dims=d_x.data_ptr(), t_dims=d_t.data_ptr())
    runs = []
    for _ in range(3):
        d_p.fill_(-1.0)
        ctx.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
        ctx.synchronize()
        runs.append(d_p.cpu().numpy().copy())
    assert np.all(d_s.cpu().numpy() == 0)
    assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[1], runs[2])
    sub = np.arange(0, n, 7)
    ref, st = oracle.fd1d_price_uneven(o[sub], xd[sub], td[sub])
    tol = EPS if precision == kwgpu.FP64 else 1e-4
    assert np.all(st == 0)
    assert np.max(np.abs(runs[0][sub] - ref)) <= tol
    monkeypatch.setenv("KWGPU_NO_GRAPH", "1")
    direct = kwgpu.Context(0)
    direct.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
    direct.synchronize()
    assert np.array_equal(d_p.cpu().numpy(), runs[0])
    direct.close()


def test_uneven_plans_alternate(ctx):
    """Several uneven batches in turn on one context: each call equals a fresh context's
    first (unplanned) call on that batch, whether it replays the cached plan (a batch
    repeated) or re-plans (a batch alternated with another)."""
    import torch
    _say("uneven, alternating batches")
    batches = []
    for seed in range(80, 86):  # six size patterns
        o, xd, td = W.uneven(700 + 50 * (seed - 80), seed)
        d_o, d_x, d_t = _uneven_device(o, xd, np.minimum(td, 40))
        n = len(o)
        prm = kwgpu.params(x_dims=d_x.data_ptr(), t_dims=d_t.data_ptr())
        batches.append((d_o, d_x, d_t, n, prm))

    def price(c, k):
        d_o, _, _, n, prm = batches[k]
        d_p = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
        d_s = torch.empty(n, dtype=torch.int32, device="cuda")
        c.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
        c.synchronize()
        assert np.all(d_s.cpu().numpy() == 0)
        return d_p.cpu().numpy()

    first = []
    for k in range(len(batches)):
        fresh = kwgpu.Context(0)
        first.append(price(fresh, k))
        fresh.close()
    c = kwgpu.Context(0)
    for k in [0, 1, 0, 1, 2, 3, 0, 1, 2, 3, 4, 5, 0, 5, 4, 1]:
        assert np.array_equal(price(c, k), first[k]), k
    c.close()


def test_autotune_keeps_results(ctx):
    import torch
    _say("autotune")
    o, xd, td = W.uneven(6000, 78)
    td = np.minimum(td, 48)
    d_o, d_x, d_t = _uneven_device(o, xd, td)
    n = len(o)
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_s = torch.empty(n, dtype=torch.int32, device="cuda")
    c2 = kwgpu.Context(0)
    prm = kwgpu.params(x_dims=d_x.data_ptr(), t_dims=d_t.data_ptr())
    c2.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
    c2.synchronize()
    before = d_p.cpu().numpy().copy()
    rep = c2.autotune(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
    _say("  ", rep)
    assert {b["bucket"] for b in rep["buckets"]} == {256, 512, 1024, 2048}
    for b in rep["buckets"]:
        assert all(c["ms"] > 0 for c in b["candidates"])
        # a lane candidate left untimed carries a lower bound no faster than the best
        # wave variant, so pruning never discards a winner
        best_wave = min(c["ms"] for c in b["candidates"] if c["layout"] == "wave")
        for c in b["candidates"]:
            if c["layout"] == "lane":
                assert isinstance(c["pruned"], bool)
                if c["pruned"]:
                    assert c["ms"] >= best_wave
        assert b["candidates"][b["chosen"]]["layout"] == "wave" or \
            not b["candidates"][b["chosen"]].get("pruned", False)
    c2.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
    c2.synchronize()
    after = d_p.cpu().numpy()
    assert np.all(d_s.cpu().numpy() == 0)
    assert np.max(np.abs(after - before)) <= EPS
    assert rep["schedule"]["chosen"] in ("serial", "concurrent")
    c2.close()


def test_bucket_schedules_agree(ctx, tmp_path):
    """An uneven batch's buckets launched one after another on one stream and launched
    concurrently on the side streams give bitwise the same prices."""
    import torch
    _say("bucket schedules")
    o, xd, td = W.uneven(3000, 79)
    td = np.minimum(td, 40)
    d_o, d_x, d_t = _uneven_device(o, xd, td)
    n = len(o)
    prm = kwgpu.params(x_dims=d_x.data_ptr(), t_dims=d_t.data_ptr())
    out = {}
    for serial in (0, 1):
        c2 = kwgpu.Context(0)
        path = str(tmp_path / f"tune{serial}.txt")
        with open(path, "w") as f:
            f.write(f"-1 {kwgpu.FP64} {serial} 0\n")
        c2.tune_load(path)
        d_p = torch.empty(n, dtype=torch.float64, device="cuda")
        d_s = torch.empty(n, dtype=torch.int32, device="cuda")
        for _ in range(2):  # the second call replays the captured graph
            c2.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
            c2.synchronize()
        assert np.all(d_s.cpu().numpy() == 0)
        out[serial] = d_p.cpu().numpy().copy()
        c2.close()
    assert np.array_equal(out[0], out[1])
    ref, _ = oracle.fd1d_price_uneven(o[:200], xd[:200], td[:200])
    assert np.max(np.abs(out[1][:200] - ref)) <= EPS
