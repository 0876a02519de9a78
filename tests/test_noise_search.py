"""The device interpreters' noise searches, checked on the CPU.

`hazard_upper_bound` (next firing site of the gap sampler, src/clifft/svm/
svm.h:229-243) and `pick_channel` (channel of a firing site,
src/clifft/svm/svm_kernels.inl:1894-1909) in clifft_amd/csrc/svm_device.h
use a bucket guide, one-round guesses and bracketed binary searches.  Their
text is compiled for the host with g++ and compared with std::upper_bound
and the reference's linear running-sum scan on random tables
(tests/cpp/noise_search_check.cpp).  A planted off-by-one in the guess's
acceptance test must be caught.
"""

from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECKER = os.path.join(ROOT, "tests", "cpp", "noise_search_check.cpp")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")


def _helpers() -> str:
    with open(os.path.join(ROOT, "clifft_amd", "csrc", "svm_device.h")) as f:
        src = f.read()
    return src[src.index("// Noise-site search"):src.index("// svm_math.h:58-61")]


def _run(tmp_path, helpers: str) -> tuple[int, int]:
    (tmp_path / "noise_search.inc").write_text(helpers)
    exe = tmp_path / "check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", str(tmp_path), CHECKER, "-o", str(exe)], check=True,
                   timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120).stdout
    last = out.strip().splitlines()[-1].split()  # "tests N bad M"
    return int(last[1]), int(last[3])


def test_searches_match_upper_bound_and_linear_scan(tmp_path):
    tests, bad = _run(tmp_path, _helpers())
    assert tests > 100_000
    assert bad == 0


def test_planted_off_by_one_is_caught(tmp_path):
    good = "if (g0 >= n || h0 > target) return g0;"
    helpers = _helpers()
    assert good in helpers
    tests, bad = _run(tmp_path, helpers.replace(good, "if (g0 >= n || h0 >= target) return g0;"))
    assert bad > 0
