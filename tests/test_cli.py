"""`kwinto price` argument handling (src/kwinto.cpp:15-85); these paths end before any
GPU call."""
from __future__ import annotations

import os
import subprocess

from conftest import PKG

CLI = os.path.join(PKG, "bin", "kwinto")


def _run(*args):
    assert os.path.exists(CLI), "build() first"
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_usage_and_version():
    r = _run("--help")
    assert r.returncode == 0 and "kwinto price [options] <path>" in r.stdout
    assert "[default: FD1D-BS]" in r.stdout
    assert _run("--version").returncode == 0
    r = _run()
    assert r.returncode == 1 and "Usage:" in r.stderr
    assert _run("price").returncode == 1
    assert _run("bench", "x.csv").returncode == 1


def test_parse_errors_use_the_reference_text():
    r = _run("price", "--density", "abc", "x.csv")
    assert r.returncode == 1
    assert "cmdBench: Fail to parse '--density <num>'" in r.stderr
    r = _run("price", "-x", "12q", "x.csv")
    assert r.returncode == 1 and "cmdPrice: Fail to parse '-x <num>'" in r.stderr


def test_missing_portfolio():
    r = _run("price", "-v", "/nonexistent/portfolio.csv")
    assert r.returncode == 1
    assert "cmdPrice: Portfolio::load : Failed to open /nonexistent/portfolio.csv" in r.stderr
    assert "Command-Line Arguments" in r.stdout
