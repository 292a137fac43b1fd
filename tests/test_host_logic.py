"""The library's host code built for the CPU against a stand-in HIP runtime
(tests/hostmock/hip_mock.cpp) and run under AddressSanitizer: every C-ABI entry point,
uneven-batch planning and graph replay, PDE grouping, the FD1D-BS and Greeks batches,
the lane workspace, the autotuner and its file, error returns, and the surrogate's
training, prediction and gated serving (tests/hostmock/drive_host.cpp). No kernel runs:
this checks memory safety, handle and allocation lifetimes and return codes of the host
side, with no GPU."""
from __future__ import annotations

import os
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
SRC = ROOT / "kwinto-cuda_amd" / "csrc"
MOCK = pathlib.Path(__file__).resolve().parent / "hostmock"
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"
UNITS = ["kwgpu_api", "fd1d_wave", "fd1d_lane", "bs", "surrogate"]


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)
    assert r.returncode == 0, f"{' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr}"
    return r


def test_host_logic_under_asan(tmp_path):
    if not (os.path.exists(HIPCC) and os.path.exists(CLANG)):
        pytest.skip("ROCm compilers not installed")
    # host code only, sanitizer on the host pass (there is no device pass)
    host = [HIPCC, "-O1", "-g", "-std=c++17", "-fPIC", "--offload-host-only", "-Xarch_host",
            "-fsanitize=address", f"-I{ROOT / 'include'}", f"-I{SRC}"]
    jobs = [(subprocess.Popen(host + ["-c", str(SRC / f"{u}.hip"), "-o", str(tmp_path / f"{u}.o")],
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True), u)
            for u in UNITS]
    for p, u in jobs:
        out, err = p.communicate(timeout=600)
        assert p.returncode == 0, f"{u}: {err}"
    _run(host + ["-c", str(MOCK / "hip_mock.cpp"), "-o", str(tmp_path / "hip_mock.o")])
    _run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address", f"-I{ROOT / 'include'}", "-c",
          str(MOCK / "drive_host.cpp"), "-o", str(tmp_path / "drive_host.o")])
    # each unit refers to its (absent) device code object by a generated name
    names = set()
    for u in UNITS:
        r = _run(["nm", "-u", str(tmp_path / f"{u}.o")])
        names |= {ln.split()[-1] for ln in r.stdout.splitlines() if "__hip_fatbin_" in ln}
    (tmp_path / "fatbins.cpp").write_text(
        "".join(f'extern "C" {{ char {n}[16]; }}\n' for n in sorted(names)))
    _run(["g++", "-c", str(tmp_path / "fatbins.cpp"), "-o", str(tmp_path / "fatbins.o")])
    objs = [str(tmp_path / f"{u}.o") for u in UNITS + ["hip_mock", "drive_host", "fatbins"]]
    _run([CLANG, "-fsanitize=address", *objs, "-o", str(tmp_path / "drive")])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    r = subprocess.run([str(tmp_path / "drive"), str(tmp_path)], capture_output=True, text=True,
                       timeout=600, env=env)
    assert "AddressSanitizer" not in r.stderr and "LeakSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and "host logic: ok" in r.stdout, r.stdout + r.stderr[-4000:]
