"""The double-word form's conversions (mpfft_amd/csrc/dd_conv.h), compiled
for the host, against MPFR: the P-bit rounding of a sum of two binary64
numbers (the last pass's output) equals mpfr_add at P bits on structured
random pairs, and the exact split of a P-bit value into two words (the
first pass's input) rounds back to the same value."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dd_conversions_match_mpfr(tmp_path):
    if subprocess.run(["which", "g++"], capture_output=True).returncode:
        pytest.skip("g++ not available")
    exe = tmp_path / "cpu_dd_check"
    cmd = ["g++", "-O2", "-std=c++17", "-Wno-unknown-pragmas", "-I/opt/conda/include", "-o",
           str(exe), os.path.join(ROOT, "tests", "cpu_dd_check.cpp"),
           "/lib/x86_64-linux-gnu/libmpfr.so.6", "/lib/x86_64-linux-gnu/libgmp.so.10"]
    subprocess.run(cmd, check=True)
    r = subprocess.run([str(exe), "40000", "2024"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("PASS")
