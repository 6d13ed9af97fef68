"""mpfr_t arrays from Python, for callers of the reference-style entry
points (mp_sr24_cp_fft and mpfft_execute_mpfr), through the system
libmpfr.  Values move in and out through the library's SoA-limb converters
(mpfft_limbs_to_mpfr / mpfft_mpfr_to_limbs)."""
import ctypes

import numpy as np

import mpfft_amd as M


class MpfrArray:
    """n mpfr_t at precision P (x86-64 layout, 32 bytes each)."""
    SIZE = 32

    _mpfr = None

    @classmethod
    def mpfr(cls):
        if cls._mpfr is None:
            cls._mpfr = ctypes.CDLL("libmpfr.so.6")
            cls._mpfr.mpfr_init2.argtypes = [ctypes.c_void_p, ctypes.c_long]
            cls._mpfr.mpfr_clear.argtypes = [ctypes.c_void_p]
        return cls._mpfr

    def __init__(self, n, P):
        self.n = n
        self.buf = (ctypes.c_char * (self.SIZE * max(n, 1)))()
        self.base = ctypes.addressof(self.buf)
        init = self.mpfr().mpfr_init2
        for k in range(n):
            init(self.base + self.SIZE * k, P)

    def ptr(self, k=0):
        return ctypes.c_void_p(self.base + self.SIZE * k)

    def __del__(self):
        clear = self.mpfr().mpfr_clear
        for k in range(self.n):
            clear(self.base + self.SIZE * k)


def pair_from_limbs(n, P, x, L):
    """(re, im) MpfrArrays at P bits from one SoA-limb transform at L limbs"""
    re, im = MpfrArray(n, P), MpfrArray(n, P)
    xx = np.ascontiguousarray(x, dtype=np.uint32)
    M.check(M.lib().mpfft_limbs_to_mpfr(n, L, ctypes.c_void_p(xx.ctypes.data),
                                         re.ptr(), im.ptr()), "limbs_to_mpfr")
    return re, im


def limbs_from_pair(n, L, re, im):
    """one SoA-limb transform at L limbs from (re, im) MpfrArrays"""
    y = np.zeros(2 * n * (L + 1), dtype=np.uint32)
    M.check(M.lib().mpfft_mpfr_to_limbs(n, L, re.ptr(), im.ptr(),
                                         ctypes.c_void_p(y.ctypes.data)), "mpfr_to_limbs")
    return y
