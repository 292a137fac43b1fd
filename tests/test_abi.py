"""C-ABI checks that need no GPU: the library loads, exports every function the
header declares, lays out kwgpu_option exactly like kw::Option, fills the reference
defaults, and rejects bad arguments before touching the device."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import kwgpu
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "kwgpu.h")


def declared_functions(path=HEADER):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kwgpu_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = kwgpu.lib()
    names = declared_functions()
    assert len(names) >= 11
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(kwgpu.EXPORTED_SYMBOLS) <= set(names)


def test_option_layout_matches_reference(tmp_path):
    """kw::Option (src/Core/kwAsset.h:12-21): t@0 k@8 z@16 r@24 q@32 s@40 e@48 w@49, 56 B."""
    src = tmp_path / "layout.cpp"
    src.write_text(
        '#include <cstddef>\n#include "kwgpu.h"\n'
        "static_assert(sizeof(kwgpu_option) == 56);\n"
        "static_assert(offsetof(kwgpu_option, t) == 0 && offsetof(kwgpu_option, k) == 8);\n"
        "static_assert(offsetof(kwgpu_option, z) == 16 && offsetof(kwgpu_option, r) == 24);\n"
        "static_assert(offsetof(kwgpu_option, q) == 32 && offsetof(kwgpu_option, s) == 40);\n"
        "static_assert(offsetof(kwgpu_option, e) == 48 && offsetof(kwgpu_option, w) == 49);\n"
        "int main() { return 0; }\n")
    subprocess.check_call(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                           str(src)])
    d = kwgpu.OPTION_DTYPE
    assert d.itemsize == 56
    assert [d.fields[c][1] for c in ("t", "k", "z", "r", "q", "s", "e", "w")] == \
        [0, 8, 16, 24, 32, 40, 48, 49]


def test_params_defaults_are_the_reference_defaults():
    p = kwgpu.Fd1dParams()
    kwgpu.lib().kwgpu_fd1d_params_default(ctypes.byref(p))
    # src/Pricer/kwFd1d.cpp:12-16, theta src/Math/kwFd1d.h:31
    assert (p.density, p.scale, p.theta, p.x_dim, p.t_dim) == (0.25, 50.0, 0.5, 512, 512)
    assert p.precision == kwgpu.FP64 and p.layout == kwgpu.LAYOUT_AUTO
    assert p.x_dims is None and p.t_dims is None


def test_null_context_is_rejected():
    lib = kwgpu.lib()
    p = kwgpu.params()
    assert lib.kwgpu_fd1d_price(None, None, 0, ctypes.byref(p), None, None) == kwgpu.ERR_INVALID
    assert lib.kwgpu_bs_price(None, None, 0, None) == kwgpu.ERR_INVALID
    assert lib.kwgpu_synchronize(None) == kwgpu.ERR_INVALID
    assert lib.kwgpu_last_error(None) == b"kwgpu: null context"


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(kwgpu, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(kwgpu, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        kwgpu.lib()


def test_cpp_api_headers_compile(tmp_path):
    """include/kw/*.h (the reference's kw:: API over the C-ABI) compile on their own."""
    src = tmp_path / "api.cpp"
    src.write_text('#include "kw/kwPricerFactory.h"\n#include "kw/kwPortfolio.h"\nint main(){}\n')
    subprocess.check_call(["g++", "-std=c++20", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                           str(src)])


def test_gate_layout_matches_header(tmp_path):
    """The ctypes Gate has kwgpu_gate's size and offsets, and gate_class assigns classes
    in the header's order: (c * n_t + a) * n_m + b, an edge value in the band above."""
    g = kwgpu.Gate
    offs = {f: getattr(g, f).offset for f, _ in g._fields_}
    src = tmp_path / "gate.cpp"
    src.write_text(
        '#include <cstddef>\n#include "kwgpu.h"\n'
        f"static_assert(sizeof(kwgpu_gate) == {ctypes.sizeof(g)});\n" +
        "".join(f"static_assert(offsetof(kwgpu_gate, {f}) == {o});\n" for f, o in offs.items()) +
        "int main() { return 0; }\n")
    subprocess.check_call(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                           str(src)])
    o = np.zeros(4, dtype=kwgpu.OPTION_DTYPE)
    o["k"] = 100.0
    o["s"] = [100.0, 90.0, 110.0, 100.0]
    o["t"] = [0.5, 0.1, 3.0, 1.0]
    o["w"] = [-1, 1, 1, -1]
    # T bands [0, 0.5), [0.5, 2), [2, inf); log(S/K) bands [<0), [>=0)
    got = kwgpu.gate_class(o, [0.5, 2.0], [0.0])
    assert got.tolist() == [(0 * 3 + 1) * 2 + 1, (1 * 3 + 0) * 2 + 0, (1 * 3 + 2) * 2 + 1,
                            (0 * 3 + 1) * 2 + 1]
    with pytest.raises(ValueError):
        kwgpu.gate([0.5] * 8, [], [0.0] * 18, [0.0] * 18)
