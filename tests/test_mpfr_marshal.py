"""mpfr_t <-> SoA-limb conversion through the C ABI helpers
(mpfft_mpfr_to_limbs, mpfft_limbs_to_mpfr; no GPU needed): a round trip
returns the limbs bit for bit, at precisions that are and are not a multiple
of 32 and 64 bits, with signed zeros and sizes that do not fill the
conversion's blocks of 32 points; the mpfr_t values agree with MPFR's own
reading of the same numbers (sign, exponent, significand)."""
import ctypes

import numpy as np
import pytest

from mpfft_amd import inputs as I
from mpfft_amd import mpfr_arrays as MA
from oracle import oracle as O


@pytest.mark.parametrize("n,P", [(1, 53), (7, 100), (31, 64), (33, 1500), (64, 1024),
                                 (1000, 2048)])
def test_mpfr_limbs_round_trip(n, P):
    L = O.limbs(P)
    x = I.mp_random(n, P, seed=n + P)
    h, lm = I.split(x, n, L)
    h, lm = h.copy(), lm.copy()
    for i, neg in ((0, 1), (2 * n - 1, 0)):          # signed zeros at both ends
        h[i] = np.int32(-2**31) | neg
        lm[:, i] = 0
    x = I.join(h, lm)
    re, im = MA.pair_from_limbs(n, P, x, L)
    assert np.array_equal(MA.limbs_from_pair(n, L, re, im), x)
    # MPFR reads the same values: sign and exponent of a few points
    mp = MA.MpfrArray.mpfr()
    mp.mpfr_get_exp.restype = ctypes.c_long
    mp.mpfr_get_exp.argtypes = [ctypes.c_void_p]
    mp.mpfr_signbit.restype = ctypes.c_int
    mp.mpfr_signbit.argtypes = [ctypes.c_void_p]
    for k in range(min(n, 5)):
        for arr, pos in ((re, k), (im, n + k)):
            hk = int(h[pos])
            if (hk & ~1) == -2**31:
                continue
            assert mp.mpfr_get_exp(arr.ptr(k)) == hk >> 1
            assert (mp.mpfr_signbit(arr.ptr(k)) != 0) == bool(hk & 1)
