"""Shared pytest configuration: the `gpu` marker, import paths, golden fixtures."""
from __future__ import annotations

import gzip
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "kwinto-cuda_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

from kwgpu.workloads import OPTION_DTYPE  # noqa: E402

OPT_COLS = ["t", "k", "z", "r", "q", "s", "e", "w"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def load_golden(name):
    """(options, {column: array}) from tests/golden/<name>.csv.gz."""
    with gzip.open(os.path.join(GOLDEN, name + ".csv.gz"), "rt") as f:
        lines = [ln for ln in f.read().splitlines() if ln and not ln.startswith("#")]
    hdr = lines[0].split(",")
    rows = [ln.split(",") for ln in lines[1:]]
    opts = np.zeros(len(rows), OPTION_DTYPE)
    cols = {}
    for j, name_ in enumerate(hdr):
        vals = [r[j] for r in rows]
        if name_ in OPT_COLS:
            opts[name_] = np.array(vals, dtype=np.float64).astype(opts.dtype[name_])
        elif name_ in ("x_dim", "t_dim"):
            cols[name_] = np.array(vals, dtype=np.int32)
        else:
            cols[name_] = np.array(vals, dtype=np.float64)
    return opts, cols


def golden_meta():
    with open(os.path.join(GOLDEN, "golden_meta.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ctx():
    import kwgpu
    c = kwgpu.Context(0)
    yield c
    c.close()
