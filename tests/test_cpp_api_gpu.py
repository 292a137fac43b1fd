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
