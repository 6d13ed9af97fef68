"""The device limb arithmetic (mpfft_amd/csrc/mp_arith.h), compiled for the
host, against MPFR's correctly rounded add/sub/mul on structured random
operands (ties, cancellation, limb-boundary alignments, signed zeros)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_limb_arithmetic_matches_mpfr(tmp_path):
    exe = tmp_path / "cpu_arith_check"
    cmd = ["g++", "-O2", "-std=c++17", "-I/opt/conda/include", "-o", str(exe),
           os.path.join(ROOT, "tests", "cpu_arith_check.cpp"),
           "/lib/x86_64-linux-gnu/libmpfr.so.6", "/lib/x86_64-linux-gnu/libgmp.so.10"]
    if subprocess.run(["which", "g++"], capture_output=True).returncode:
        pytest.skip("g++ not available")
    subprocess.run(cmd, check=True)
    r = subprocess.run([str(exe), "30000", "777"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("PASS")
