"""Pins the CPU oracle (oracle/kw_oracle.cpp) before it is trusted as the checker.

- against the golden vectors the compiled reference wrote (tests/golden/), to 1e-12;
- against the known answers the reference's own tests hold (test/kwPricer_test.cpp
  values at 1.3e-3, the QuantLib fixture test/portfolio_qdfp.csv.zst at 0.005 as in
  test/kwPortfolio_test.cpp:52-57);
- against mathematics: LAPACK on the tridiagonal solve, put-call parity and the
  closed form for BS, second-order convergence of the European FD price to BS,
  grid end points and symmetry, American >= European, V >= payoff, monotonicity;
- and by planting plausible mistakes in a rebuilt oracle and showing that the
  golden-vector check rejects each of them.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import golden_meta, load_golden
from kwgpu import workloads as W

EPS_GOLDEN = 1e-12


def test_plain_512_matches_reference_vectors():
    o, c = load_golden("plain_512")
    p, err = oracle.fd1d_price(o, 512, 512)
    assert err == ""
    assert np.max(np.abs(p - c["ref_fd1d"])) <= EPS_GOLDEN
    p, err = oracle.fd1d_bs_price(o, 512, 512)
    assert err == ""
    assert np.max(np.abs(p - c["ref_fd1d_bs"])) <= EPS_GOLDEN
    oe = o.copy()
    oe["e"] = 0
    assert np.max(np.abs(oracle.bs_price(oe) - c["ref_bs_euro"])) <= EPS_GOLDEN
    p, _ = oracle.fd1d_price(oe, 512, 512)
    assert np.max(np.abs(p - c["ref_fd1d_euro"])) <= EPS_GOLDEN


def test_pricer_test_12_known_answers():
    """test/kwPricer_test.cpp:37-108: all three pricers within 1.3e-3 of the stated
    values (BS NaNs skipped), and equal to the reference's own outputs."""
    o, c = load_golden("pricer_test_12")
    bs = oracle.bs_price(o)
    fd, e1 = oracle.fd1d_price(o)
    fdbs, e2 = oracle.fd1d_bs_price(o)
    assert e1 == "" and e2 == ""
    ok = ~np.isnan(bs)
    assert ok.sum() == 6 and np.all(np.isnan(c["ref_bs"][~ok]))
    assert np.max(np.abs(bs[ok] - c["want"][ok])) <= 1.3e-3
    assert np.max(np.abs(fd - c["want"])) <= 1.3e-3
    assert np.max(np.abs(fdbs - c["want"])) <= 1.3e-3
    assert np.max(np.abs(bs[ok] - c["ref_bs"][ok])) <= EPS_GOLDEN
    assert np.max(np.abs(fd - c["ref_fd1d"])) <= EPS_GOLDEN
    assert np.max(np.abs(fdbs - c["ref_fd1d_bs"])) <= EPS_GOLDEN


def test_qdfp_fixture():
    """test/kwPortfolio_test.cpp:29-58: 6000 American puts priced at 512 x 512 within
    0.005 of QuantLib's QdFp prices; and equal to the reference FD1D output."""
    o, c = load_golden("portfolio_qdfp")
    assert len(o) == 6000
    p, err = oracle.fd1d_price(o, 512, 512)
    assert err == ""
    assert np.max(np.abs(p - c["want"])) <= 0.005
    assert np.max(np.abs(p - c["ref_fd1d"])) <= EPS_GOLDEN


def test_plain_1024_head():
    o, c = load_golden("plain_1024_head256")
    p, err = oracle.fd1d_price(o, 1024, 1024)
    assert err == ""
    assert np.max(np.abs(p - c["ref_fd1d"])) <= EPS_GOLDEN


def test_uneven_small():
    o, c = load_golden("uneven_small")
    p, st = oracle.fd1d_price_uneven(o, c["x_dim"], c["t_dim"])
    assert np.all(st == 0)
    assert np.max(np.abs(p - c["ref_fd1d"])) <= EPS_GOLDEN


def test_range_error_message():
    m = golden_meta()["range_error"]
    o = W.plain(4, m["seed"])
    o["s"] = m["spots"]
    _, err = oracle.fd1d_price(o, m["x_dim"], m["t_dim"], m["density"], m["scale"])
    assert err == m["message"] and err.startswith("Fd1d_Pricer::price Fd1d::value: x=")
    _, st = oracle.fd1d_price_status(o, m["x_dim"], m["t_dim"], m["density"], m["scale"])
    assert np.any(st == 1)


def test_empty_portfolio():
    p, err = oracle.fd1d_price(np.zeros(0, W.OPTION_DTYPE))
    assert err == "" and len(p) == 0


@pytest.mark.skipif(not oracle.ref_available(), reason="reference checker not built here")
@pytest.mark.parametrize("seed,x,t", [(11, 100, 37), (12, 257, 64), (13, 64, 300)])
def test_against_compiled_reference(seed, x, t):
    o = W.plain(48, seed)
    o["e"][::3] = 0
    p, e1 = oracle.fd1d_price(o, x, t)
    r, e2 = oracle.ref_price("FD1D", o, x, t)
    assert e1 == e2
    assert np.array_equal(p, r)


# ---- mathematics ----------------------------------------------------------------

def test_tridiagonal_matches_lapack():
    rng = np.random.default_rng(0)
    for n in (3, 7, 64, 513):
        al = rng.uniform(-1, 0, n)
        au = rng.uniform(-1, 0, n)
        a = 2.5 + rng.uniform(0, 1, n)
        al[0] = 0
        au[-1] = 0
        y = rng.normal(size=n)
        x, rc = oracle.solve_tridiagonal(al, a, au, y)
        assert rc == 0
        A = np.diag(a) + np.diag(al[1:], -1) + np.diag(au[:-1], 1)
        assert np.max(np.abs(x - np.linalg.solve(A, y))) < 1e-13
        assert np.max(np.abs(A @ x - y)) < 1e-13
    # error codes of src/Math/kwMath.cpp:19-24
    assert oracle.solve_tridiagonal([0, 0, 0], [0, 1, 1], [0, 0, 0], [1, 1, 1])[1] == 1
    assert oracle.solve_tridiagonal([0, 0], [1, 1], [0, 0], [1, 1])[1] == 3


def _bs_closed_form(o):
    """Textbook Black-Scholes-Merton with continuous dividend yield."""
    out = []
    for a in o:
        t, k, z, r, q, s, w = (float(a[c]) for c in ("t", "k", "z", "r", "q", "s", "w"))
        d1 = (math.log(s / k) + (r - q + 0.5 * z * z) * t) / (z * math.sqrt(t))
        d2 = d1 - z * math.sqrt(t)
        N = lambda v: 0.5 * math.erfc(-v / math.sqrt(2))  # noqa: E731
        if w > 0:
            out.append(s * math.exp(-q * t) * N(d1) - k * math.exp(-r * t) * N(d2))
        else:
            out.append(k * math.exp(-r * t) * N(-d2) - s * math.exp(-q * t) * N(-d1))
    return np.array(out)


def test_bs_closed_form_and_parity():
    o = W.plain(64, 21)
    o["e"] = 0
    o["s"] = np.linspace(60, 160, 64)
    p = oracle.bs_price(o)
    assert np.max(np.abs(p - _bs_closed_form(o))) < 1e-10
    c = o.copy()
    c["w"] = 1
    pu = o.copy()
    pu["w"] = -1
    lhs = oracle.bs_price(c) - oracle.bs_price(pu)
    rhs = c["s"] * np.exp(-c["q"] * c["t"]) - c["k"] * np.exp(-c["r"] * c["t"])
    assert np.max(np.abs(lhs - rhs)) < 1e-10
    am = o.copy()
    am["e"] = 1
    assert np.all(np.isnan(oracle.bs_price(am)))


def test_european_fd_converges_to_bs_second_order():
    o = W.plain(16, 22)
    o["e"] = 0
    bs = oracle.bs_price(o)
    errs = []
    for n in (128, 256, 512):
        p, err = oracle.fd1d_price(o, n, n)
        assert err == ""
        errs.append(np.max(np.abs(p - bs)))
    assert errs[0] / errs[1] > 2.8 and errs[1] / errs[2] > 2.8, errs


def test_grid_endpoints_symmetry_and_payoff_floor():
    o = W.plain(8, 23)
    for a in o:
        x, v = oracle.solve_one(np.array([a]), 257, 64)
        half = 50 * a["z"] * math.sqrt(a["t"])
        assert abs(x[0] + half) < 1e-12 * half and abs(x[-1] - half) < 1e-12 * half
        assert np.all(np.diff(x) > 0)
        assert np.max(np.abs(x + x[::-1])) < 1e-12 * half
        pay = np.maximum(0, np.exp(x) - 1) if a["w"] > 0 else np.maximum(0, 1 - np.exp(x))
        assert np.all(v[:-1] >= pay[:-1])


def test_american_ge_european_and_monotone_in_spot():
    o = np.repeat(W.plain(6, 24), 9)
    o["s"] = np.tile(np.linspace(70, 130, 9), 6)
    am, _ = oracle.fd1d_price(o, 256, 256)
    e = o.copy()
    e["e"] = 0
    eu, _ = oracle.fd1d_price(e, 256, 256)
    assert np.all(am >= eu - 1e-12)
    for i in range(6):
        blk = am[9 * i: 9 * i + 9]
        d = np.diff(blk)
        assert np.all(d >= -1e-12) if o["w"][9 * i] > 0 else np.all(d <= 1e-12)


def _bs_greeks(o):
    """Closed-form Black-Scholes Greeks in the bump script's units: vega and rho per
    1%, theta as dV/dT per day of 360 (test/portfolio.py:35-56)."""
    from scipy.special import ndtr
    s, k, t, z, r, q = (o[f] for f in ("s", "k", "t", "z", "r", "q"))
    w = o["w"].astype(np.float64)
    sq = z * np.sqrt(t)
    d1 = (np.log(s / k) + (r - q + 0.5 * z * z) * t) / sq
    d2 = d1 - sq
    pdf = np.exp(-0.5 * d1 * d1) / math.sqrt(2 * math.pi)
    eq, er = np.exp(-q * t), np.exp(-r * t)
    return {
        "price": w * (s * eq * ndtr(w * d1) - k * er * ndtr(w * d2)),
        "delta": w * eq * ndtr(w * d1),
        "gamma": eq * pdf / (s * sq),
        "vega": s * eq * pdf * np.sqrt(t) / 100,
        "theta": (s * eq * pdf * z / (2 * np.sqrt(t)) + w * r * k * er * ndtr(w * d2)
                  - w * q * s * eq * ndtr(w * d1)) / 360,
        "rho": w * k * t * er * ndtr(w * d2) / 100,
    }


def test_greeks_of_european_fd_match_black_scholes():
    """The bump Greeks of the FD pricer on European options against the closed forms:
    agreement to the grid's discretisation pins signs, factors and bump bookkeeping."""
    rng = np.random.default_rng(0)
    n = 12
    o = np.zeros(n, oracle.OPTION_DTYPE)
    o["s"] = 100
    o["k"] = rng.uniform(80, 120, n)
    o["t"] = rng.uniform(0.25, 2, n)
    o["z"] = rng.uniform(0.15, 0.45, n)
    o["r"] = rng.uniform(0.01, 0.08, n)
    o["q"] = rng.uniform(0, 0.05, n)
    o["w"] = np.where(np.arange(n) % 2, 1, -1)
    g, st, _ = oracle.fd1d_greeks(o, 1024, 1024)
    assert np.all(st == 0)
    ref = _bs_greeks(o)
    tol = {"price": 5e-4, "delta": 5e-4, "gamma": 5e-4, "vega": 5e-4, "theta": 5e-5,
           "rho": 5e-5}
    for f, tl in tol.items():
        assert np.max(np.abs(g[f] - ref[f])) < tl, f


def test_american_put_greeks_signs():
    o = np.repeat(W.plain(4, 25), 5)
    o["w"] = -1
    o["s"] = np.tile([60.0, 80.0, 100.0, 120.0, 140.0], 4)
    g, st, _ = oracle.fd1d_greeks(o, 512, 512)
    assert np.all(st == 0)
    # deep in the exercise region the bumped prices lie on chords of the concave payoff,
    # which leaves gamma, vega and theta slightly negative (a few 1e-4 at 512 points)
    assert np.all((g["delta"] >= -1 - 1e-6) & (g["delta"] <= 1e-12))
    assert np.all(g["gamma"] >= -1e-4)
    assert np.all(g["vega"] >= -1e-3) and np.all(g["theta"] >= -1e-4)
    assert np.all(g["rho"] <= 1e-9)
    live = o["s"] >= 80  # outside the exercise region the signs are strict
    assert np.all(g["gamma"][live] > 0) and np.all(g["vega"][live] > 0)
    assert np.all(g["theta"][live] > 0) and np.all(g["rho"][live] < 0)


def test_baw_baseline_tracks_the_fd_american_price():
    """The surrogate's analytic baseline: BS + BAW premium within the approximation's
    known accuracy of the FD American price (a wrong sign or root gives errors of whole
    price units), value matching at the critical spot, and no premium where early
    exercise never pays."""
    from baw_ref import baw_premium, critical_spot, european
    rng = np.random.default_rng(7)
    n = 48
    o = np.zeros(n, oracle.OPTION_DTYPE)
    o["s"] = 100
    o["k"] = rng.uniform(85, 115, n)
    o["t"] = rng.uniform(0.25, 1.0, n)
    o["z"] = rng.uniform(0.2, 0.4, n)
    o["r"] = rng.uniform(0.03, 0.08, n)
    o["q"] = rng.uniform(0.02, 0.08, n)
    o["w"] = np.where(np.arange(n) % 2, 1, -1)
    o["e"] = 1
    fd, err = oracle.fd1d_price(o, 1024, 1024)
    assert err == ""
    e = o.copy()
    e["e"] = 0
    eu = oracle.bs_price(e)
    prem, left = fd - eu, fd - eu - baw_premium(o)
    assert np.max(np.abs(left)) < 0.15 and np.median(np.abs(left)) < 0.02
    # it removes most of the premium (here about 80%) from what the network must learn
    assert np.sqrt(np.mean(left ** 2)) < 0.3 * np.sqrt(np.mean(prem ** 2))
    assert np.all(baw_premium(o) >= 0)
    # at S = S* the continuation formula meets the exercise value
    ss, _ = critical_spot(o)
    at = o.copy()
    at["s"] = ss * (1 + 1e-12)
    eur, _ = european(at)
    assert np.allclose(baw_premium(at) + eur, np.where(o["w"] < 0, o["k"] - ss, ss - o["k"]),
                       rtol=0, atol=1e-6 * o["k"])
    z = o.copy()
    z["q"] = np.where(z["w"] > 0, 0.0, z["q"])
    z["r"] = np.where(z["w"] < 0, 0.0, z["r"])
    assert np.all(baw_premium(z) == 0)


def test_boundary_premium_is_the_fd_solvers_limit():
    """The surrogate's control variate, the premium from the exercise boundary's fixed
    point, against the FD American price as the grid is refined: the gap halves with
    each doubling of the grid (the FD scheme converges at first order under the
    early-exercise projection), so the boundary premium is the limit the FD prices
    approach. A wrong term in the boundary equation or the premium integral leaves a
    gap that does not shrink. It also pays nothing where early exercise never pays,
    prices calls as the put with S and K, r and q exchanged, and pays the exercise
    value minus the European price below the critical spot."""
    from baw_ref import baseline_premium, boundary_premium, european
    o = W.surrogate_train(96, seed=11)
    o["s"] = o["k"] * np.exp(np.random.default_rng(3).uniform(-0.3, 0.3, len(o)))
    bp = boundary_premium(o)
    eu, _ = european(o)
    rms = []
    for nx in (512, 1024, 2048):
        fd, err = oracle.fd1d_price(o, nx, nx)
        assert err == ""
        rms.append(float(np.sqrt(np.mean(((fd - eu) - bp) ** 2))))
    # measured: 5.1e-3, 2.4e-3, 1.25e-3
    assert rms[1] < 0.6 * rms[0] and rms[2] < 0.6 * rms[1], rms
    fd, _ = oracle.fd1d_price(o, 1024, 1024)
    assert rms[1] < 0.1 * float(np.sqrt(np.mean(((fd - eu) - baseline_premium(o)) ** 2)))
    assert np.all(bp >= -1e-12)
    z = o.copy()
    z["q"] = np.where(z["w"] > 0, 0.0, z["q"])
    z["r"] = np.where(z["w"] < 0, 0.0, z["r"])
    assert np.all(boundary_premium(z) == 0)
    c = o[o["w"] > 0][:8].copy()
    p = c.copy()
    p["w"], p["s"], p["k"], p["r"], p["q"] = -1, c["k"], c["s"], c["q"], c["r"]
    assert np.allclose(boundary_premium(c), boundary_premium(p), rtol=1e-12, atol=1e-12)
    deep = o[(o["w"] < 0) & (o["r"] > o["q"] + 0.01)][:8].copy()  # exercise region: b(T)
    deep["s"] = deep["k"] * 0.05                                    # is near K when r > q
    ed, _ = european(deep)
    assert np.allclose(boundary_premium(deep), (deep["k"] - deep["s"]) - ed, rtol=0, atol=1e-12)


def test_ju_zhong_baseline_improves_on_baw():
    """The baseline the surrogate uses, Ju & Zhong's correction of BAW, against the FD
    American price over the training box's maturities (to 5 years, where BAW's error is
    largest): a wrong sign or factor in either coefficient of chi makes the correction
    worse than BAW, not several times better. It also meets BAW at S* (chi = 0 there),
    pays nothing where early exercise never pays, and is continuous in r at r = 0."""
    from baw_ref import baseline_premium, baw_premium, critical_spot
    o = W.surrogate_train(160, seed=11)
    o["s"] = o["k"] * np.exp(np.random.default_rng(3).uniform(-0.3, 0.3, len(o)))
    fd, err = oracle.fd1d_price(o, 1024, 1024)
    assert err == ""
    e = o.copy()
    e["e"] = 0
    prem = fd - oracle.bs_price(e)
    jz, baw = baseline_premium(o) - prem, baw_premium(o) - prem
    rms = lambda v: float(np.sqrt(np.mean(v ** 2)))
    # measured: rms 0.059 against 0.50, median 0.017 against 0.092
    assert rms(jz) < 0.35 * rms(baw)
    assert np.median(np.abs(jz)) < 0.5 * np.median(np.abs(baw))
    assert np.all(baseline_premium(o) >= 0)
    ss, _ = critical_spot(o)
    at = o.copy()
    at["s"] = ss * (1 + 1e-12)
    assert np.allclose(baseline_premium(at), baw_premium(at), rtol=1e-9, atol=1e-9 * o["k"])
    z = o.copy()
    z["q"] = np.where(z["w"] > 0, 0.0, z["q"])
    z["r"] = np.where(z["w"] < 0, 0.0, z["r"])
    assert np.all(baseline_premium(z) == 0)
    calls = o[(o["w"] > 0) & (o["q"] > 0.01)].copy()
    calls["r"] = 0.0
    near = calls.copy()
    near["r"] = 1e-9
    assert np.allclose(baseline_premium(calls), baseline_premium(near), rtol=1e-6, atol=1e-9)


# ---- planted mistakes -----------------------------------------------------------

PLANTS = {
    1: "theta 0.55 instead of the Crank-Nicolson 0.5",
    2: "swap the lower and upper second-difference weights of the implicit rows",
    3: "t-grid with T / tDim instead of T / (tDim - 1)",
    4: "interpolate from the node after the bracket",
    5: "payoff of a call for every option",
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_golden_check_catches_planted_mistake(plant, tmp_path):
    so = tmp_path / f"liboracle_plant{plant}.so"
    subprocess.check_call(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", f"-DKWO_PLANT={plant}",
                           "-o", str(so), os.path.join(oracle.HERE, "kw_oracle.cpp"),
                           "-lpthread"])
    L = ctypes.CDLL(str(so))
    P, SZ, D, I64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_int64
    L.kwo_fd1d_price.argtypes = [P, SZ, D, D, I64, I64, ctypes.c_int, P, ctypes.c_char_p, SZ]
    o, c = load_golden("plain_512")
    out = np.zeros(len(o))
    err = ctypes.create_string_buffer(256)
    L.kwo_fd1d_price(np.ascontiguousarray(o).ctypes.data, len(o), 0.25, 50.0, 512, 512, 0,
                     out.ctypes.data, err, 256)
    assert np.max(np.abs(out - c["ref_fd1d"])) > 1e-8, PLANTS[plant]
