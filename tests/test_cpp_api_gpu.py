"""Runs tests/cpp/kw_api_test (the reference's gtest cases written against the kw::
C++ API of this build) on the GPU."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, load_golden

pytestmark = pytest.mark.gpu

BIN = os.path.join(PKG, "_build", "kw_api_test")


def _write_reference_csv(path, opts, want):
    """The reference's portfolio CSV layout (test/portfolio.py:97)."""
    with open(path, "w") as f:
        f.write("expiry,spot,strike,volatility,interest_rate,dividend_rate,parity,exercise,price\n")
        for a, p in zip(opts, want):
            t, s, k, z, r, q = (repr(float(a[c])) for c in ("t", "s", "k", "z", "r", "q"))
            f.write(f"{t},{s},{k},{z},{r},{q},{'c' if a['w'] > 0 else 'p'},"
                    f"{'a' if a['e'] else 'e'},{float(p)!r}\n")


def test_reference_test_cases_through_cpp_api(tmp_path):
    assert os.path.exists(BIN), "build() first"
    o, c = load_golden("portfolio_qdfp")
    csv = tmp_path / "portfolio_qdfp.csv"
    _write_reference_csv(csv, o, [float(x) for x in c["want"]])
    print("running", BIN, flush=True)
    r = subprocess.run([BIN, str(csv)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "0 failed" in r.stdout


CLI = os.path.join(PKG, "bin", "kwinto")


def test_cli_prices_the_qdfp_portfolio(tmp_path):
    """`kwinto price` (src/kwinto.cpp:40-83) on the QuantLib fixture: FD1D at 512 x 512
    within the reference test's 0.005 of QuantLib, statistics in the reference's format."""
    assert os.path.exists(CLI), "build() first"
    o, c = load_golden("portfolio_qdfp")
    csv = tmp_path / "portfolio_qdfp.csv"
    _write_reference_csv(csv, o, [float(x) for x in c["want"]])
    r = subprocess.run([CLI, "price", "-p", "FD1D", "-t", "512", "-x", "512", str(csv)],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "Assets : 6000" in r.stdout and "Price Statistics" in r.stdout
    mae = float(r.stdout.split("MAE :")[1].split()[0])
    assert mae <= 0.005
    # the statistics skip reference prices below -e (0.5), as the reference does
    n_stats = int(np.sum(np.asarray(c["want"], np.float64) >= 0.5))
    assert f"total : {n_stats} options" in r.stdout
