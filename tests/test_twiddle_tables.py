"""The FFT twiddle tables that the kernels copy into LDS are compile-time
constants (msosm_gpu_amd/csrc/twiddle_tables.h).  tests/cpp/twiddle_check.cpp
evaluates the same constexpr builders with the host compiler and compares every
entry of every transform size's two-level table (2^1 .. 2^14) and of the
2^14-point middle pass's row table with the float nearest to the exact
exp(-2 pi i e/N) (long-double libm inside an exact quadrant reduction).  No GPU
calls."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "msosm_gpu_amd", "csrc")
SRC = os.path.join(ROOT, "tests", "cpp", "twiddle_check.cpp")


def _build(tmp_path, include_dir):
    exe = tmp_path / "twiddle_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", str(include_dir), SRC, "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs the host C++ compiler")
def test_tables_are_correctly_rounded(tmp_path):
    r = subprocess.run([str(_build(tmp_path, CSRC))], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout
    # every table was checked, each with all of its entries
    assert r.stdout.count(" 0 wrong") == 15


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs the host C++ compiler")
def test_check_catches_a_wrong_octant(tmp_path):
    # plant a mistake in a copy of the header: the third octant's complement
    # argument replaced by the direct one
    with open(os.path.join(CSRC, "twiddle_tables.h")) as fh:
        text = fh.read()
    good = "case 3: return {-cos_oct(y), sin_oct(y)};"
    assert good in text
    bad_dir = tmp_path / "planted"
    bad_dir.mkdir()
    (bad_dir / "twiddle_tables.h").write_text(text.replace(good, "case 3: return {-cos_oct(x), sin_oct(x)};"))
    r = subprocess.run([str(_build(tmp_path, bad_dir))], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "FAIL" in r.stdout
