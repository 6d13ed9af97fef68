"""The fixed-point arithmetic of the fixed FFT form
(mpfft_amd/csrc/fx_arith.h), compiled for the host, against exact integer
arithmetic (GMP) and MPFR's rounding: exact sums, products within one unit
of 2^-F of the truncated exact product, exact float -> fixed conversion and
correctly rounded fixed -> float conversion."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fixed_point_arithmetic(tmp_path):
    if subprocess.run(["which", "g++"], capture_output=True).returncode:
        pytest.skip("g++ not available")
    exe = tmp_path / "cpu_fx_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I/opt/conda/include", "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpu_fx_check.cpp"),
                    "/lib/x86_64-linux-gnu/libmpfr.so.6", "/lib/x86_64-linux-gnu/libgmp.so.10"],
                   check=True)
    r = subprocess.run([str(exe), "4000", "31"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("PASS")
