"""The C-ABI library loads and exports every function include/mpfft.h
declares; argument validation that happens before any GPU work returns
the documented status codes.  No compute calls: runs without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import mpfft_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mpfft.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", src)
    keywords = {"if", "sizeof", "defined"}
    return sorted(set(n for n in names if n not in keywords))


def test_header_and_python_export_list_agree():
    assert sorted(M.EXPORTS) == declared_functions()


def test_library_exports_every_declared_symbol():
    lib = M.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_limbs_for_precision():
    assert M.limbs_for_precision(53) == 2
    assert M.limbs_for_precision(256) == 8
    assert M.limbs_for_precision(1024) == 32
    assert M.limbs_for_precision(160) == 6      # padded to a compiled count
    assert M.limbs_for_precision(1) == -1
    assert M.limbs_for_precision(1025) == 64     # the panel kernels' 66-word reals
    assert M.limbs_for_precision(2048) == 64
    assert M.limbs_for_precision(2049) == -1


def test_strerror():
    for code in list(M.ERRORS) + [0]:
        assert M.lib().mpfft_strerror(code)


def test_subtree_call_is_rejected_without_touching_the_gpu():
    lib = M.lib()
    buf = (ctypes.c_double * 8)()
    lib.d_sr24_cp_fft(4, 2, 0, buf, buf, buf, buf)
    assert lib.mpfft_last_error() == -4
    lib.d_sr24_cp_fft_v2(4, 1, 1, 0, buf, buf, buf, buf)
    assert lib.mpfft_last_error() == -4
    lib.d_sr24_cp_fft(3, 1, 0, buf, buf, buf, buf)
    assert lib.mpfft_last_error() == -1
    lib.mp_sr24_cp_fft(8, 1, 0, None, None, None, None)
    assert lib.mpfft_last_error() == -6


def test_plan_rejects_bad_size_and_form():
    h = ctypes.c_void_p()
    assert M.lib().mpfft_plan_create(ctypes.byref(h), 12, 256, 0, 1) == -1
    assert M.lib().mpfft_plan_create(ctypes.byref(h), 16, 256, 3, 1) == -8
    assert M.lib().mpfft_plan_create(None, 16, 256, 0, 1) == -6
