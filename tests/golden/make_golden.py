"""Writes the golden vectors under tests/golden/ from the compiled reference.

Run in a container where the reference tree is mounted (it is built into
oracle/_ref/libkwref.so by `make -C oracle ref`); the GPU box never runs this.
Every vector is the reference's own Pricer output (PricerFactory::create +
Pricer::price, src/Pricer/kwPricerFactory.h:15-41) on the stated inputs, written
with 17 significant digits. Each batch is priced twice and must agree: the
reference's thread-pool join can return early (src/kwThreadPool.cpp:59-72).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import gzip
import io
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "kwinto-cuda_amd"))

import oracle  # noqa: E402
from kwgpu import workloads as W  # noqa: E402

COLS = ["t", "k", "z", "r", "q", "s", "e", "w"]
REFERENCE = os.environ.get("KW_REFERENCE", "/root/reference")


def ref_price(pricer, opts, x_dim, t_dim, density=0.25, scale=50.0):
    a, e1 = oracle.ref_price(pricer, opts, x_dim, t_dim, density, scale)
    b, e2 = oracle.ref_price(pricer, opts, x_dim, t_dim, density, scale)
    for _ in range(5):
        if np.array_equal(a, b, equal_nan=True) and e1 == e2:
            return a, e1
        a, e1 = b, e2
        b, e2 = oracle.ref_price(pricer, opts, x_dim, t_dim, density, scale)
    raise RuntimeError("reference pricer not reproducible")


def write_csv(name, opts, cols: dict, header_note=""):
    buf = io.StringIO()
    names = COLS + list(cols)
    buf.write("# " + header_note + "\n" if header_note else "")
    buf.write(",".join(names) + "\n")
    for i in range(len(opts)):
        row = [repr(float(opts[c][i])) if c not in ("e", "w") else str(int(opts[c][i]))
               for c in COLS]
        for c, v in cols.items():
            x = v[i]
            if isinstance(x, (np.integer, int)):
                row.append(str(int(x)))
            else:
                row.append("%.17g" % float(x))
        buf.write(",".join(row) + "\n")
    with gzip.GzipFile(os.path.join(HERE, name), "wb", mtime=0) as f:
        f.write(buf.getvalue().encode())


def read_zst_csv(path):
    """Decompress a .zst fixture of the reference with the system libzstd."""
    z = ctypes.CDLL("libzstd.so.1")
    z.ZSTD_getFrameContentSize.restype = ctypes.c_ulonglong
    z.ZSTD_decompress.restype = ctypes.c_size_t
    src = open(path, "rb").read()
    n = z.ZSTD_getFrameContentSize(src, len(src))
    dst = ctypes.create_string_buffer(n)
    r = z.ZSTD_decompress(dst, n, src, len(src))
    return dst.raw[:r].decode()


def portfolio_from_csv(text):
    """Header-mapped columns like Portfolio::load (src/Utils/kwPortfolio.cpp:13-84)."""
    lines = text.strip().splitlines()
    hdr = lines[0].split(",")
    ix = {c: hdr.index(c) for c in hdr}
    opts = np.zeros(len(lines) - 1, W.OPTION_DTYPE)
    want = np.zeros(len(lines) - 1)
    for i, ln in enumerate(lines[1:]):
        v = ln.split(",")
        opts[i]["k"] = float(v[ix["strike"]])
        opts[i]["q"] = float(v[ix["dividend_rate"]])
        opts[i]["r"] = float(v[ix["interest_rate"]])
        opts[i]["t"] = float(v[ix["expiry"]])
        opts[i]["s"] = float(v[ix["spot"]])
        opts[i]["z"] = float(v[ix["volatility"]])
        opts[i]["e"] = 1 if v[ix["exercise"]] == "a" else 0
        opts[i]["w"] = 1 if v[ix["parity"]] == "c" else -1
        want[i] = float(v[ix["price"]])
    return opts, want


def main():
    if not oracle.ref_available():
        oracle.build(REFERENCE)
    meta = {}

    # 1. The 12 options of kwPricerTest (test/kwPricer_test.cpp:16-28) with the
    #    expected values the reference's tests assert at 1.3e-3, and the reference's
    #    own outputs for BS / FD1D / FD1D-BS at the default 512 x 512 config.
    rows = [
        (1.0, 100., 0.2, 0.06, 0.02, 90., 0, -1, 10.627),
        (1.0, 100., 0.2, 0.06, 0.02, 100., 0, -1, 5.885),
        (1.0, 100., 0.2, 0.06, 0.02, 110., 0, -1, 2.987),
        (1.0, 100., 0.2, 0.06, 0.02, 90., 0, +1, 4.668),
        (1.0, 100., 0.2, 0.06, 0.02, 100., 0, +1, 9.729),
        (1.0, 100., 0.2, 0.06, 0.02, 110., 0, +1, 16.633),
        (1.0, 100., 0.2, 0.06, 0.08, 90., 1, -1, 13.988121682),
        (1.0, 100., 0.2, 0.06, 0.08, 100., 1, -1, 8.409190396),
        (1.0, 100., 0.2, 0.06, 0.08, 110., 1, -1, 4.6592955111),
        (1.0, 100., 0.2, 0.06, 0.08, 90., 1, +1, 2.9472036256),
        (1.0, 100., 0.2, 0.06, 0.08, 100., 1, +1, 6.8422540642),
        (1.0, 100., 0.2, 0.06, 0.08, 110., 1, +1, 12.7940107108),
    ]
    o = np.zeros(len(rows), W.OPTION_DTYPE)
    for i, r in enumerate(rows):
        for c, v in zip(COLS, r[:8]):
            o[c][i] = v
    want = np.array([r[8] for r in rows])
    bs, _ = ref_price("BS", o, 512, 512)
    fd, _ = ref_price("FD1D", o, 512, 512)
    fdbs, _ = ref_price("FD1D-BS", o, 512, 512)
    write_csv("pricer_test_12.csv.gz", o,
              {"want": want, "ref_bs": bs, "ref_fd1d": fd, "ref_fd1d_bs": fdbs},
              "kwPricerTest options; want = test/kwPricer_test.cpp values; ref_* = reference "
              "output at 512x512")
    meta["pricer_test_12"] = {"n": len(o), "x_dim": 512, "t_dim": 512}

    # 2. Plain method, 128 options, 512 x 512 (seed 0): FD1D, FD1D-BS and BS(euro).
    o = W.plain(128, 0)
    fd, e = ref_price("FD1D", o, 512, 512)
    assert not e, e
    fdbs, e = ref_price("FD1D-BS", o, 512, 512)
    assert not e, e
    oe = o.copy()
    oe["e"] = 0
    bs, _ = ref_price("BS", oe, 512, 512)
    fde, _ = ref_price("FD1D", oe, 512, 512)
    write_csv("plain_512.csv.gz", o,
              {"ref_fd1d": fd, "ref_fd1d_bs": fdbs, "ref_bs_euro": bs, "ref_fd1d_euro": fde},
              "128 seeded American options, 512 x-points x 512 t-points")
    meta["plain_512"] = {"n": 128, "seed": 0, "x_dim": 512, "t_dim": 512}

    # 3. First 256 options of the 32768-option 1024 x 1024 workload (seed 1).
    o = W.plain(32768, 1)[:256]
    fd, e = ref_price("FD1D", o, 1024, 1024)
    assert not e, e
    write_csv("plain_1024_head256.csv.gz", o, {"ref_fd1d": fd},
              "options 0..255 of the 32768-option seed-1 workload, 1024 x 1024")
    meta["plain_1024_head256"] = {"n": 256, "seed": 1, "x_dim": 1024, "t_dim": 1024}

    # 4. Uneven grids: 8 options from each x bucket of the seed-2 uneven workload,
    #    each priced by the reference at its own (x, t) grid.
    o_all, xd_all, td_all = W.uneven(4096, 2)
    pick = []
    for xb in W.X_BUCKETS:
        idx = np.nonzero(xd_all == xb)[0][:8]
        pick.extend(idx.tolist())
    stress = np.nonzero(np.arange(4096) % 50 == 0)[0][:4]
    pick.extend(stress.tolist())
    pick = np.array(sorted(set(pick)))
    o = o_all[pick]
    xd = xd_all[pick]
    td = td_all[pick]
    fd = np.zeros(len(o))
    for i in range(len(o)):
        fd[i], e = ref_price("FD1D", o[i:i + 1], int(xd[i]), int(td[i]))
        fd[i] = fd[i] if not e else np.nan
    write_csv("uneven_small.csv.gz", o, {"x_dim": xd, "t_dim": td, "ref_fd1d": fd},
              "uneven-grid options (seed 2) each at its own x_dim x t_dim")
    meta["uneven_small"] = {"n": int(len(o)), "seed": 2}

    # 5. The reference's QuantLib fixture (test/portfolio_qdfp.csv.zst: 6000 American
    #    puts), its QdFp prices, and the reference FD1D price at 512 x 512 as asserted by
    #    test/kwPortfolio_test.cpp:29-58 (tolerance 0.005).
    txt = read_zst_csv(os.path.join(REFERENCE, "test", "portfolio_qdfp.csv.zst"))
    o, want = portfolio_from_csv(txt)
    fd, e = ref_price("FD1D", o, 512, 512)
    assert not e, e
    write_csv("portfolio_qdfp.csv.gz", o, {"want": want, "ref_fd1d": fd},
              "test/portfolio_qdfp.csv.zst rows; want = QuantLib QdFp price; ref_fd1d = "
              "reference FD1D at 512x512")
    meta["portfolio_qdfp"] = {"n": int(len(o)), "x_dim": 512, "t_dim": 512}

    # 6. Error behaviour: log(S/K) outside the grid (tiny FD1D.SCALE) -- the reference's
    #    message (src/Math/kwFd1d.cpp:151-153, src/Pricer/kwFd1d.cpp:154-155).
    o = W.plain(4, 5)
    o["s"] = [100.0, 180.0, 100.0, 60.0]
    _, e = ref_price("FD1D", o, 64, 32, 0.25, 0.05)
    meta["range_error"] = {"x_dim": 64, "t_dim": 32, "density": 0.25, "scale": 0.05,
                           "message": e, "seed": 5, "spots": [100.0, 180.0, 100.0, 60.0]}

    with open(os.path.join(HERE, "golden_meta.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
