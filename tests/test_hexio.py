"""Exact text form of SoA-limb data (host code in libmpfft, no GPU):
mpfft_limbs_to_hex prints what MPFR's printf prints for "%Ra", and
mpfft_hex_to_limbs reads such text back, rounding as mpfr_set_str does
(to nearest, ties to even).  Checked against the system libmpfr."""
import ctypes

import numpy as np
import pytest

import mpfft_amd as M
from mpfft_amd import inputs as I
from mpfft_amd import mpfr_arrays as MA

ZERO = np.int32(-2**31)


def mpfr():
    m = MA.MpfrArray.mpfr()
    m.mpfr_set_str.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    return m


def mpfr_ra(arr, k):
    s = ctypes.c_char_p()
    mpfr().mpfr_asprintf(ctypes.byref(s), b"%Ra", arr.ptr(k))
    return s.value.decode()


def sample(n, P, seed):
    """random values with a spread of exponents, signs and signed zeros"""
    L = I.limbs_for(P)
    x = I.mp_random(n, P, seed=seed)
    h, lim = I.split(x, n, L)
    rng = np.random.default_rng(seed)
    e = rng.integers(-3000, 3000, size=2 * n)
    e[:8] = [0, 1, -1, 2, 3, 4, 5, -4]
    h[:] = ((e << 1) | (h & 1)).astype(np.int32)
    h[8] = ZERO
    h[9] = ZERO | 1
    lim[:, 8:10] = 0
    return x


@pytest.mark.parametrize("P", [2, 53, 64, 100, 256, 1024])
def test_to_hex_matches_mpfr_printf(P):
    n = 64
    L = I.limbs_for(P)
    x = sample(n, P, seed=P)
    got = M.to_hex(x, 2 * n, L)
    re, im = MA.pair_from_limbs(n, P, x, L)
    ref = [mpfr_ra(re, k) for k in range(n)] + [mpfr_ra(im, k) for k in range(n)]
    assert got == ref


@pytest.mark.parametrize("P", [53, 64, 256, 1024])
def test_hex_round_trip_is_exact(P):
    n = 128
    L = I.limbs_for(P)
    x = sample(n, P, seed=P + 1)
    assert np.array_equal(M.from_hex(M.to_hex(x, 2 * n, L), L, P), x)


@pytest.mark.parametrize("P", [2, 24, 53, 100, 256])
def test_from_hex_rounds_like_mpfr(P):
    rng = np.random.default_rng(P)
    L = I.limbs_for(P)
    strings = []
    for k in range(200):
        digits = "".join(rng.choice(list("0123456789abcdef"), size=int(rng.integers(1, 90))))
        if k % 7 == 0:                           # exact halfway cases
            digits = "1" + "0" * (P // 4) + "8"
        sign = "-" if k % 3 == 0 else ""
        point = int(rng.integers(0, len(digits) + 1))
        frac = digits[:point] + ("." + digits[point:] if point < len(digits) else "")
        strings.append(f"{sign}0x{frac or '0'}p{int(rng.integers(-900, 900)):+d}")
    got = M.from_hex(strings, L, P)
    n = len(strings)
    arr = MA.MpfrArray(n, P)
    for k, s in enumerate(strings):
        mpfr().mpfr_set_str(arr.ptr(k).value, s.encode(), 16, 0)
    # the converter takes (re, im) pairs: use the strings as n re values and
    # n copies of zero as im, then compare the re half
    zeros = MA.MpfrArray(n, P)                  # mpfr_init2 leaves NaN: set +0
    for k in range(n):
        mpfr().mpfr_set_ui(zeros.ptr(k), ctypes.c_ulong(0), 0)
    ref = MA.limbs_from_pair(n, L, arr, zeros)
    h, lim = I.split(ref, n, L)
    ref_re = np.concatenate([h[:n].view(np.uint32), lim[:, :n].ravel()])
    assert np.array_equal(got, ref_re)


def test_hex_errors():
    L, P = 2, 64
    for bad in ["0x", "1.5p+0", "0x1.g", "0x1p", "0x1.8p+0z", "0x1p+999999999"]:
        with pytest.raises(M.MpfftError) as e:
            M.from_hex([bad], L, P)
        assert e.value.code == -3, bad
    with pytest.raises(M.MpfftError):
        M.from_hex(["0x1p+0"], 1, 64)              # prec beyond the limbs
