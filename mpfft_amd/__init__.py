"""mpfft_amd: multiprecision split-radix conjugate-pair FFT on AMD MI355X.

The product is the C-ABI library libmpfft.so (include/mpfft.h), built from
mpfft_amd/csrc by `python -m mpfft_amd.build`.  This module only loads it
through ctypes for tests and benchmarks; there is no Python or CPU
fallback: if the library is missing, loading raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpfft.so")

FORM_MP, FORM_D5, FORM_D6 = 0, 5, 6
FORM_MP_C1, FORM_MP_C2 = 1, 2          # the reference's CODE 1 / CODE 2 op orders
FORM_FIXED = 8                          # fixed-point radix-4 form (same limb I/O)
FORM_FIXED_LANE = 9                     # the fixed form on its per-lane kernels only
FORM_FIXED_PANEL = 10                   # the fixed form on panels (any batch, n >= 16)
FORM_DD = 11                            # double-word FP64 form, 53 <= P <= 106
FORM_QD = 12                            # quad-word FP64 form, 53 <= P <= 212
FORM_AUTO = 13                          # the chosen form (resolved at plan creation)
FIXED_FORMS = (FORM_FIXED, FORM_FIXED_LANE, FORM_FIXED_PANEL)
MP_FORMS = (FORM_MP, FORM_MP_C1, FORM_MP_C2, FORM_DD, FORM_QD, FORM_AUTO) + FIXED_FORMS
OK = 0
ERRORS = {
    -1: "MPFFT_ERR_SIZE", -2: "MPFFT_ERR_PREC", -3: "MPFFT_ERR_VALUE",
    -4: "MPFFT_ERR_SUBTREE", -5: "MPFFT_ERR_HIP", -6: "MPFFT_ERR_ARG",
    -7: "MPFFT_ERR_NODEV", -8: "MPFFT_ERR_FORM", -9: "MPFFT_ERR_SCRATCH",
}

# every symbol include/mpfft.h declares
EXPORTS = [
    "mp_sr24_cp_fft", "mp_sr24_cp_fft_v2", "d_sr24_cp_fft", "d_sr24_cp_fft_v2",
    "mpfft_last_error", "mpfft_strerror", "mpfft_plan_create", "mpfft_plan_destroy",
    "mpfft_plan_info", "mpfft_plan_form", "mpfft_execute_mpfr", "mpfft_execute_limbs",
    "mpfft_execute_limbs_host", "mpfft_execute_limbs_range", "mpfft_plan_stages",
    "mpfft_plan_panel_plan",
    "mpfft_execute_f64", "mpfft_execute_f64_host",
    "mpfft_plan_sync", "mpfft_mpfr_to_limbs", "mpfft_limbs_to_mpfr",
    "mpfft_mp_binop_host", "mpfft_limbs_for_precision",
    "mpfft_execute_inverse_limbs", "mpfft_roundtrip_stats", "mpfft_roundtrip_host",
    "mpfft_roundtrip_stats_noalloc",
    "mpfft_plan_last_timing", "mpfft_limbs_to_hex", "mpfft_hex_to_limbs",
    "mpfft_plan_check", "mpfft_batch_create", "mpfft_batch_destroy", "mpfft_batch_words",
    "mpfft_batch_info", "mpfft_batch_tune", "mpfft_batch_schedule",
    "mpfft_batch_execute_limbs", "mpfft_batch_check",
]

_lib = None


class MpfftError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: {ERRORS.get(code, code)}")


def lib():
    """Load libmpfft.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} not built: run python -m mpfft_amd.build")
    try:
        # torch bundles its own libamdhip64 under the same soname; loading it
        # first makes libmpfft bind to that runtime instead of pulling a
        # second HIP runtime into the process
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P, U, I, L_, VP = (ctypes.c_void_p, ctypes.c_ulong, ctypes.c_int, ctypes.c_long,
                       ctypes.c_void_p)
    L.mpfft_plan_create.argtypes = [ctypes.POINTER(VP), U, L_, I, ctypes.c_uint]
    L.mpfft_plan_destroy.argtypes = [VP]
    L.mpfft_plan_info.argtypes = [VP, ctypes.POINTER(L_), ctypes.POINTER(ctypes.c_size_t)]
    L.mpfft_execute_limbs.argtypes = [VP, P, P, P]
    L.mpfft_execute_limbs_host.argtypes = [VP, P, P]
    L.mpfft_execute_limbs_range.argtypes = [VP, P, P, P, I, I]
    L.mpfft_plan_stages.argtypes = [VP]
    L.mpfft_plan_panel_plan.argtypes = [VP]
    L.mpfft_plan_form.argtypes = [VP]
    L.mpfft_execute_f64.argtypes = [VP, P, P, P]
    L.mpfft_execute_f64_host.argtypes = [VP, P, P]
    L.mpfft_plan_sync.argtypes = [VP]
    L.mpfft_execute_inverse_limbs.argtypes = [VP, P, P, P]
    L.mpfft_roundtrip_stats.argtypes = [VP, P, P, P, P]
    L.mpfft_roundtrip_stats_noalloc.argtypes = [VP, P, P, P, P]
    L.mpfft_roundtrip_host.argtypes = [VP, P, P]
    L.mpfft_plan_last_timing.argtypes = [VP, P]
    L.mpfft_plan_check.argtypes = [VP]
    L.mpfft_limbs_to_hex.argtypes = [U, L_, P, P, ctypes.c_size_t,
                                     ctypes.POINTER(ctypes.c_size_t)]
    L.mpfft_hex_to_limbs.argtypes = [U, L_, L_, ctypes.c_char_p, P]
    L.mpfft_mp_binop_host.argtypes = [I, U, L_, P, P, P]
    L.mpfft_batch_create.argtypes = [ctypes.POINTER(VP), ctypes.c_uint, P, P, I]
    L.mpfft_batch_destroy.argtypes = [VP]
    L.mpfft_batch_words.argtypes = [VP, ctypes.POINTER(ctypes.c_size_t)]
    L.mpfft_batch_info.argtypes = [VP, ctypes.c_uint, ctypes.POINTER(ctypes.c_size_t),
                                   ctypes.POINTER(L_), ctypes.POINTER(I)]
    L.mpfft_batch_tune.argtypes = [VP, P, P]
    L.mpfft_batch_schedule.argtypes = [VP, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_double)]
    L.mpfft_batch_execute_limbs.argtypes = [VP, P, P, P]
    L.mpfft_batch_check.argtypes = [VP]
    L.mpfft_limbs_for_precision.argtypes = [L_]
    L.mpfft_limbs_for_precision.restype = L_
    L.mpfft_strerror.restype = ctypes.c_char_p
    L.mp_sr24_cp_fft.argtypes = [U, U, L_, P, P, P, P]
    L.mp_sr24_cp_fft_v2.argtypes = [U, U, U, L_, P, P, P, P]
    L.d_sr24_cp_fft.argtypes = [U, U, L_, P, P, P, P]
    L.d_sr24_cp_fft_v2.argtypes = [U, U, U, L_, P, P, P, P]
    _lib = L
    return L


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def check(rc, what):
    if rc != OK:
        raise MpfftError(rc, what)


def limbs_for_precision(P):
    return int(lib().mpfft_limbs_for_precision(P))


class Plan:
    """A plan of `batch` transforms of size n (mp form at P bits, or a
    double form)."""

    def __init__(self, n, P=0, form=FORM_MP, batch=1):
        self.n, self.P, self.form, self.batch = n, P, form, batch
        h = ctypes.c_void_p()
        check(lib().mpfft_plan_create(ctypes.byref(h), n, P, form, batch), "plan_create")
        self.h = h
        Lc, W = ctypes.c_long(), ctypes.c_size_t()
        check(lib().mpfft_plan_info(h, ctypes.byref(Lc), ctypes.byref(W)), "plan_info")
        self.L, self.words = Lc.value, W.value

    def close(self):
        if self.h:
            lib().mpfft_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def execute_host(self, x):
        """x: host array of batch*words (uint32 for mp, float64 for double)."""
        if self.form in MP_FORMS:
            x = np.ascontiguousarray(x, dtype=np.uint32)
            y = np.empty_like(x)
            assert x.size == self.words * self.batch
            check(lib().mpfft_execute_limbs_host(self.h, _ptr(x), _ptr(y)), "execute_limbs_host")
        else:
            x = np.ascontiguousarray(x, dtype=np.float64)
            y = np.empty_like(x)
            assert x.size == self.words * self.batch
            check(lib().mpfft_execute_f64_host(self.h, _ptr(x), _ptr(y)), "execute_f64_host")
        return y

    def execute_device(self, d_in, d_out, stream=None):
        """d_in, d_out: device pointers (ints), e.g. torch tensor data_ptr()."""
        s = ctypes.c_void_p(stream) if stream else None
        if self.form in MP_FORMS:
            rc = lib().mpfft_execute_limbs(self.h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), s)
        else:
            rc = lib().mpfft_execute_f64(self.h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), s)
        check(rc, "execute_device")

    def stages(self):
        return int(lib().mpfft_plan_stages(self.h))

    def panel_plan(self):
        """2: the four-step panel plan (conversions inside the first and
        last passes), 1: panel passes between conversion stages, 0: none"""
        return int(lib().mpfft_plan_panel_plan(self.h))

    def resolved_form(self):
        """the form the plan runs (FORM_AUTO resolved, FORM_FIXED as its
        kernel family)"""
        rc = int(lib().mpfft_plan_form(self.h))
        check(rc if rc < 0 else 0, "plan_form")
        return rc

    def execute_range(self, d_in, d_out, s0, s1, stream=None):
        s = ctypes.c_void_p(stream) if stream else None
        check(lib().mpfft_execute_limbs_range(self.h, ctypes.c_void_p(d_in),
                                              ctypes.c_void_p(d_out), s, s0, s1),
              "execute_range")

    def sync(self):
        check(lib().mpfft_plan_sync(self.h), "plan_sync")

    def check(self):
        """wait, then raise if a fixed-form input was out of range"""
        check(lib().mpfft_plan_check(self.h), "plan_check")

    def execute_inverse(self, d_in, d_out, stream=None):
        """inverse transform by the reference's swap trick (device pointers)"""
        s = ctypes.c_void_p(stream) if stream else None
        check(lib().mpfft_execute_inverse_limbs(self.h, ctypes.c_void_p(d_in),
                                                ctypes.c_void_p(d_out), s), "execute_inverse")

    def roundtrip_stats(self, d_x, d_y, stream=None):
        """error statistics of y against x (device pointers): array (batch, 8)
        of rRMSE re/im, max|err| re/im and the log2 of the four"""
        out = np.zeros(8 * self.batch)
        s = ctypes.c_void_p(stream) if stream else None
        check(lib().mpfft_roundtrip_stats(self.h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_y),
                                          s, _ptr(out)), "roundtrip_stats")
        return out.reshape(self.batch, 8)

    def roundtrip_host(self, x):
        """the reference's test(): forward, inverse, statistics; (batch, 8)"""
        x = np.ascontiguousarray(x, dtype=np.uint32)
        assert x.size == self.words * self.batch
        out = np.zeros(8 * self.batch)
        check(lib().mpfft_roundtrip_host(self.h, _ptr(x), _ptr(out)), "roundtrip_host")
        return out.reshape(self.batch, 8)


class Batch:
    """Transforms of uneven size n[i] and precision P[i] in one graph launch
    (mpfft_batch_*).  Device buffers use the packed layout: transform i at
    word offsets[i], limbs[i] limbs; `words` words in all."""

    def __init__(self, ns, Ps, form=FORM_FIXED):
        ns = np.ascontiguousarray(ns, dtype=np.uint64)
        Ps = np.ascontiguousarray(Ps, dtype=np.int64)
        assert ns.size == Ps.size and ns.size > 0
        self.ns, self.Ps, self.form, self.count = ns, Ps, form, int(ns.size)
        h = ctypes.c_void_p()
        check(lib().mpfft_batch_create(ctypes.byref(h), self.count, _ptr(ns), _ptr(Ps), form),
              "batch_create")
        self.h = h
        w = ctypes.c_size_t()
        check(lib().mpfft_batch_words(h, ctypes.byref(w)), "batch_words")
        self.words = w.value
        self.offsets, self.limbs = self._layout()

    def _layout(self):
        off, L = np.empty(self.count, dtype=np.int64), np.empty(self.count, dtype=np.int64)
        o, l_ = ctypes.c_size_t(), ctypes.c_long()
        for i in range(self.count):
            check(lib().mpfft_batch_info(self.h, i, ctypes.byref(o), ctypes.byref(l_), None),
                  "batch_info")
            off[i], L[i] = o.value, l_.value
        return off, L

    def forms(self):
        """the form that runs each transform (after tuning: lane or panel)"""
        f, out = ctypes.c_int(), np.empty(self.count, dtype=np.int64)
        for i in range(self.count):
            check(lib().mpfft_batch_info(self.h, i, None, None, ctypes.byref(f)), "batch_info")
            out[i] = f.value
        return out

    def close(self):
        if self.h:
            lib().mpfft_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pack(self, xs):
        """host packed buffer from per-transform SoA-limb arrays"""
        buf = np.zeros(self.words, dtype=np.uint32)
        for i, x in enumerate(xs):
            w = 2 * int(self.ns[i]) * (int(self.limbs[i]) + 1)
            assert x.size == w
            buf[self.offsets[i]:self.offsets[i] + w] = x
        return buf

    def unpack(self, buf):
        out = []
        for i in range(self.count):
            w = 2 * int(self.ns[i]) * (int(self.limbs[i]) + 1)
            out.append(np.array(buf[self.offsets[i]:self.offsets[i] + w], dtype=np.uint32))
        return out

    def tune(self, d_in, d_out):
        check(lib().mpfft_batch_tune(self.h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out)),
              "batch_tune")

    def schedule(self):
        """(streams, measured ms of the graph)"""
        s, ms = ctypes.c_int(), ctypes.c_double()
        check(lib().mpfft_batch_schedule(self.h, ctypes.byref(s), ctypes.byref(ms)),
              "batch_schedule")
        return s.value, ms.value

    def execute_device(self, d_in, d_out, stream=None):
        s = ctypes.c_void_p(stream) if stream else None
        check(lib().mpfft_batch_execute_limbs(self.h, ctypes.c_void_p(d_in),
                                              ctypes.c_void_p(d_out), s), "batch_execute")

    def check(self):
        check(lib().mpfft_batch_check(self.h), "batch_check")


def to_hex(x, count, L):
    """the reals of a SoA-limb array (count reals, L limbs) as MPFR "%Ra" text"""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    need = ctypes.c_size_t(0)
    lib().mpfft_limbs_to_hex(count, L, _ptr(x), None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value)
    check(lib().mpfft_limbs_to_hex(count, L, _ptr(x), buf, need.value, ctypes.byref(need)),
          "limbs_to_hex")
    return buf.value.decode().split()


def from_hex(strings, L, P):
    """SoA-limb array (L limbs) of hexadecimal floats rounded to P bits"""
    count = len(strings)
    y = np.zeros(count * (L + 1), dtype=np.uint32)
    check(lib().mpfft_hex_to_limbs(count, L, P, " ".join(strings).encode(), _ptr(y)),
          "hex_to_limbs")
    return y


def mp_binop(op, x, y, count, P):
    """element-wise correctly rounded op (0 add, 1 sub, 2 mul) on the GPU;
    x, y: SoA-limb arrays of count reals at limbs_for_precision(P) limbs."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    y = np.ascontiguousarray(y, dtype=np.uint32)
    z = np.empty_like(x)
    check(lib().mpfft_mp_binop_host(op, count, P, _ptr(x), _ptr(y), _ptr(z)), "mp_binop")
    return z
