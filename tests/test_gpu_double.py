"""GPU parity of the double conjugate-pair FFT with the reference's double
codes (CODE 5, and CODE 6/7 behind d_sr24_cp_fft / _v2).  Bar: bit-exact
(the kernels are built with -ffp-contract=off, like gcc -O3 on x86-64)."""
import ctypes
import hashlib

import numpy as np
import pytest

import mpfft_amd as M
from mpfft_amd import inputs as I
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 4, 8, 16, 32, 64, 1024]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("code", [5, 6])
def test_plan_matches_reference(n, code, d_golden, digests):
    x = I.d_random(n, seed=digests["seed"])
    plan = M.Plan(n, 0, code)
    y = plan.execute_host(x)
    assert np.array_equal(bits(y), bits(d_golden[f"out{code}_N{n}"]))


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("entry", ["v1", "v2"])
def test_legacy_entry_points(n, entry, d_golden, digests):
    x = I.d_random(n, seed=digests["seed"])
    inr, ini = x[:n].copy(), x[n:].copy()
    outr, outi = np.zeros(n), np.zeros(n)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib = M.lib()
    if entry == "v1":
        lib.d_sr24_cp_fft(n, 1, 0, p(inr), p(ini), p(outr), p(outi))
    else:
        lib.d_sr24_cp_fft_v2(n, 0, 1, 0, p(inr), p(ini), p(outr), p(outi))
    assert lib.mpfft_last_error() == 0
    ref = d_golden[f"out{6 if entry == 'v1' else 7}_N{n}"]
    assert np.array_equal(bits(np.concatenate([outr, outi])), bits(ref))


@pytest.mark.parametrize("code", [5, 6])
def test_reference_in2_input(code, d_golden):
    x = d_golden["in2_in_N1024"]
    y = M.Plan(1024, 0, code).execute_host(x)
    assert np.array_equal(bits(y), bits(d_golden[f"in2_out{code}_N1024"]))


@pytest.mark.parametrize("code", [5, 6, 7])
def test_large_global_memory_path(code, digests):
    n = 1 << 16
    x = I.d_random(n, seed=digests["seed"])
    y = M.Plan(n, 0, 5 if code == 5 else 6).execute_host(x)
    assert sha(y) == digests["d_large"][f"code{code}_N{n}"]["out_sha"]


@pytest.mark.parametrize("n", [2048, 8192])
def test_lds_and_global_paths_agree_with_oracle(n):
    x = I.d_random(n, seed=n)
    for code in (5, 6):
        y = M.Plan(n, 0, code).execute_host(x)
        assert np.array_equal(bits(y), bits(O.d_fft(code, n, x)))


def test_batch():
    n, b = 1024, 4
    xs = np.concatenate([I.d_random(n, seed=10 + i) for i in range(b)])
    y = M.Plan(n, 0, 6, b).execute_host(xs)
    for i in range(b):
        assert np.array_equal(bits(np.split(y, b)[i]), bits(O.d_fft(6, n, np.split(xs, b)[i])))
