"""ctypes front end of the oracle (oracle/liboracle.so) and of the reference
builds under oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Build with `make -C oracle`.
All multiprecision data crosses this boundary in the SoA-limb layout of
include/mpfft.h (one transform: hdr int32[2N], then limbs uint32[L][2N]).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref")

_lib = None
_refs = {}


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_ORACLE):
            raise OSError(f"{_ORACLE} missing: run make -C oracle")
        L = ctypes.CDLL(_ORACLE)
        U, Lg, I, P = ctypes.c_ulong, ctypes.c_long, ctypes.c_int, ctypes.c_void_p
        L.orc_mp_fft.argtypes = [U, Lg, Lg, P, P, P, P, I]
        L.orc_mp_fft_code.argtypes = [I, U, Lg, Lg, P, P, P, P, I]
        L.orc_mp_fft_time.argtypes = [U, Lg, Lg, P, P, I, I]
        L.orc_mp_fft_time.restype = ctypes.c_double
        L.orc_mp_lut.argtypes = [U, Lg, P, P]
        L.orc_mp_binop.argtypes = [I, Lg, Lg, P, P, P, P, P, P]
        L.orc_mp_input_in2.argtypes = [U, Lg, P, P]
        L.orc_err_stats.argtypes = [U, Lg, Lg, Lg, P, P, P, P, P]
        L.orc_err_excess.argtypes = [U, Lg, Lg, Lg, P, P, P, P, Lg, P]
        L.orc_mp_roundtrip.argtypes = [U, Lg, P, P, P, P, P]
        L.orc_mp_counts.argtypes = [U, P]
        L.orc_d_lut.argtypes = [U, P, P, P]
        L.orc_d_fft.argtypes = [I, U, P, P, P, P]
        _lib = L
    return _lib


def limbs(P):
    return (P + 31) // 32


def _split(buf, n, L):
    buf = np.ascontiguousarray(buf, dtype=np.uint32)
    hdr = buf[:2 * n]
    lim = buf[2 * n:2 * n * (L + 1)]
    return buf, hdr, lim


def mp_fft(n, P, x, L_in=None, Pw=None, nthr=0, code=3):
    """Forward FFT of SoA transform x (L_in limbs, values exact at Pw bits)
    computed at Pw bits (default P); output at ceil(Pw/32) limbs.  code:
    the reference's operation order, 3 (serial_mp_fft3.c, 3m3a), 2
    (serial_mp_fft2.c, 4m2a) or 1 (serial_mp_fft1.c, 4m2a, one butterfly)."""
    L_in = limbs(P) if L_in is None else L_in
    Pw = P if Pw is None else Pw
    Lo = limbs(Pw)
    x, hx, lx = _split(x, n, L_in)
    y = np.zeros(2 * n * (Lo + 1), dtype=np.uint32)
    rc = lib().orc_mp_fft_code(code, n, Pw, L_in, _p(hx), _p(lx), _p(y[:2 * n]),
                               _p(y[2 * n:]), nthr)
    if rc:
        raise RuntimeError(f"orc_mp_fft rc={rc}")
    return y


def mp_fft_time(n, P, x, reps=1, nthr=1):
    """seconds per FFT (conversion and LUT excluded)."""
    L = limbs(P)
    x, hx, lx = _split(x, n, L)
    return lib().orc_mp_fft_time(n, P, L, _p(hx), _p(lx), reps, nthr)


def mp_lut(n, P):
    """LUT r, i, s, d (4 x n/4 reals) as SoA with M = n."""
    L = limbs(P)
    buf = np.zeros(n * (L + 1), dtype=np.uint32)
    rc = lib().orc_mp_lut(n, P, _p(buf[:n]), _p(buf[n:]))
    if rc:
        raise RuntimeError(f"orc_mp_lut rc={rc}")
    return buf


def mp_binop(op, count, P, x, y):
    L = limbs(P)
    x = np.ascontiguousarray(x, dtype=np.uint32)
    y = np.ascontiguousarray(y, dtype=np.uint32)
    z = np.zeros(count * (L + 1), dtype=np.uint32)
    c = count
    rc = lib().orc_mp_binop(op, count, P, _p(x[:c]), _p(x[c:]), _p(y[:c]), _p(y[c:]),
                            _p(z[:c]), _p(z[c:]))
    if rc:
        raise RuntimeError(f"orc_mp_binop rc={rc}")
    return z


def mp_input_in2(n, P):
    L = limbs(P)
    buf = np.zeros(2 * n * (L + 1), dtype=np.uint32)
    lib().orc_mp_input_in2(n, P, _p(buf[:2 * n]), _p(buf[2 * n:]))
    return buf


def err_stats(n, P, x, Lx, star, Ls):
    """(max componentwise ulp_P error, normwise error in units of 2^-P,
    count of components where star is zero and x is not)."""
    x, hx, lx = _split(x, n, Lx)
    s, hs, ls = _split(star, n, Ls)
    out = np.zeros(3, dtype=np.float64)
    rc = lib().orc_err_stats(n, P, Lx, Ls, _p(hx), _p(lx), _p(hs), _p(ls), _p(out))
    if rc:
        raise RuntimeError(f"orc_err_stats rc={rc}")
    return float(out[0]), float(out[1]), int(out[2])


def err_excess(n, P, x, Lx, star, Ls, ea):
    """(max over components of (|x - star| - ulp_P(x)/2)^+ / 2^ea, number of
    components not within half an ulp of star): the absolute error, in units
    of 2^ea, beyond one correct rounding to P bits."""
    x, hx, lx = _split(x, n, Lx)
    s, hs, ls = _split(star, n, Ls)
    out = np.zeros(2, dtype=np.float64)
    rc = lib().orc_err_excess(n, P, Lx, Ls, _p(hx), _p(lx), _p(hs), _p(ls), int(ea), _p(out))
    if rc:
        raise RuntimeError(f"orc_err_excess rc={rc}")
    return float(out[0]), int(out[1])


def max_exponent(x, n):
    """largest exponent of the 2n reals of a SoA transform (None if all are
    zero): the block exponent of the fixed and word forms."""
    h = np.ascontiguousarray(x[:2 * n]).view(np.int32) >> 1
    m = int(h.max())
    return None if m == -(1 << 30) else m


def fixed_bound(n):
    """The fixed form's absolute error bound before its one rounding to P
    bits, in units of 2^(emax - F): 2 (log2 n + 2) n, emax = the largest
    input exponent, F = 32 (L + 1) the fraction bits at L device limbs.
    Derivation, in max-norm over complex values with u = 2^-F: after the
    block exponent is taken off |x| < sqrt(2); the input conversion
    truncates (error < sqrt(2) u); sums are exact; a pass of radix R
    multiplies the errors before it by at most R (every output sums R
    inputs with unit coefficients); before a pass whose inputs are below
    sqrt(2) S (S: product of the radices so far) a twiddle product adds at
    most 3 u of truncation (two real products of <= 1.04 u per part) and
    2.42 S u from the twiddle table's rounding to F bits (per-lane Gauss
    tables; sqrt(2) S u for the panels' c, d tables), and a radix-8 pass's
    internal products by 1/sqrt(2) add 6 u + 8 S u after its radix-2
    stage.  Summed: below n (1.21 log2 n + 1) for radix 4 and
    n (0.81 log2 n + 2) for radix 8, within 2 (log2 n + 2) n."""
    lg = max(int(n).bit_length() - 1, 0)
    return float(2 * (lg + 2) * n)


def fixed_check(n, P, L_dev, x, y, star=None, Ly=None):
    """Check a fixed-form result y (at Ly limbs, default ceil(P/32)) of input
    x (ceil(P/32) limbs) against its contract: every component within half
    an ulp of y plus fixed_bound(n) units of 2^(emax - F) of the exact
    transform (the 2P-bit oracle), and normwise within 2^-P (one rounding)
    plus that bound's share.  Returns (excess in units of 2^(emax - F), max
    ulp, normwise in 2^-P, components beyond half an ulp)."""
    Lo, L2 = limbs(P), limbs(2 * P)
    Ly = Lo if Ly is None else Ly
    if star is None:
        star = mp_fft(n, P, x, Pw=2 * P)
    F = 32 * (L_dev + 1)
    em = max_exponent(x, n)
    em = 0 if em is None else em
    exc, over = err_excess(n, P, y, Ly, star, L2, em - F)
    mu, nw, _ = err_stats(n, P, y, Ly, star, L2)
    lg = max(int(n).bit_length() - 1, 0)
    delta = fixed_bound(n) * 2.0 ** (P - F + 1.5)
    assert exc <= fixed_bound(n), (n, P, exc, fixed_bound(n))
    assert nw <= 1.0 + delta, (n, P, nw, delta)
    return exc, mu, nw, over


def scale_pow2(x, n, s):
    """x 2^s for a SoA transform (exact: the exponent of every nonzero real
    moves by s)."""
    y = np.array(x, dtype=np.uint32, copy=True)
    h = y[:2 * n].view(np.int32)
    nz = h != np.int32(-2**31)
    nz &= h != np.int32(-2**31 + 1)
    h[nz] += np.int32(2 * s)
    return y


def roundtrip(n, P, x, X):
    """the reference's test(): rRMSE (re, im) and max abs error (re, im)."""
    L = limbs(P)
    x, hx, lx = _split(x, n, L)
    X, hX, lX = _split(X, n, L)
    out = np.zeros(4, dtype=np.float64)
    rc = lib().orc_mp_roundtrip(n, P, _p(hx), _p(lx), _p(hX), _p(lX), _p(out))
    if rc:
        raise RuntimeError(f"orc_mp_roundtrip rc={rc}")
    return out


def counts(n):
    c = np.zeros(4, dtype=np.uint64)
    lib().orc_mp_counts(n, _p(c))
    return [int(v) for v in c]


def d_lut(n):
    wr = np.zeros(max(n // 4, 1)); wi = np.zeros(max(n // 4, 1)); cst = np.zeros(5)
    lib().orc_d_lut(n, _p(wr), _p(wi), _p(cst))
    return wr[:n // 4], wi[:n // 4], cst


def d_fft(code, n, x):
    """double CP FFT (code 5 or 6/7) of x = re[n] ++ im[n]."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(2 * n, dtype=np.float64)
    rc = lib().orc_d_fft(code, n, _p(x[:n]), _p(x[n:]), _p(y[:n]), _p(y[n:]))
    if rc:
        raise RuntimeError(f"orc_d_fft rc={rc}")
    return y


# ---- reference builds (oracle/_ref, compiled from the reference sources) ----

REF_MP = {"mp1": "libref_mp1.so", "mp2": "libref_mp2.so", "mp3": "libref_mp3.so",
          "mp4": "libref_mp4.so", "mpp2": "libref_mpp2.so"}
REF_D = {5: "libref_d5.so", 6: "libref_d6.so", 7: "libref_d7.so"}


def have_ref():
    return all(os.path.exists(os.path.join(_REF, f))
               for f in list(REF_MP.values()) + list(REF_D.values()))


def _ref(name):
    if name not in _refs:
        path = os.path.join(_REF, name)
        if not os.path.exists(path):
            raise OSError(f"{path} missing (the reference tree is needed to build it)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        if name.startswith("libref_mp"):
            L.ref_mp_run.argtypes = [ctypes.c_ulong, ctypes.c_long, ctypes.c_int, P, P, P, P, P]
            L.ref_mp_lut.argtypes = [ctypes.c_ulong, ctypes.c_long, P, P]
            L.ref_mp_time.argtypes = [ctypes.c_ulong, ctypes.c_long, P, P, ctypes.c_int]
            L.ref_mp_time.restype = ctypes.c_double
        else:
            L.ref_d_run.argtypes = [ctypes.c_ulong, ctypes.c_int, P, P, P, P, P, P, P]
        _refs[name] = L
    return _refs[name]


def ref_mp(kind, n, P, x=None):
    """run the reference kernel (kind mp3 / mp4 / mpp2).  x: SoA input, or
    None for the reference's own IN=2 input.  Returns (input, output,
    counters)."""
    L = limbs(P)
    M = 2 * n
    xin = np.zeros(M * (L + 1), dtype=np.uint32) if x is None else \
        np.ascontiguousarray(x, dtype=np.uint32).copy()
    y = np.zeros(M * (L + 1), dtype=np.uint32)
    cnt = np.zeros(4, dtype=np.uint64)
    rc = _ref(REF_MP[kind]).ref_mp_run(n, P, 2 if x is None else 0, _p(xin[:M]), _p(xin[M:]),
                                       _p(y[:M]), _p(y[M:]), _p(cnt))
    if rc:
        raise RuntimeError(f"ref_mp_run rc={rc}")
    return xin, y, [int(v) for v in cnt]


def ref_mp_time(kind, n, P, x, reps=1):
    """mean seconds per call of the reference kernel (kind mp3: serial
    CODE 3; mpp2: OpenMP CODE 2, which uses 4 threads)."""
    L = limbs(P)
    x = np.ascontiguousarray(x, dtype=np.uint32)
    return _ref(REF_MP[kind]).ref_mp_time(n, P, _p(x[:2 * n]), _p(x[2 * n:]), reps)


def ref_mp_lut(n, P):
    L = limbs(P)
    buf = np.zeros(n * (L + 1), dtype=np.uint32)
    _ref(REF_MP["mp3"]).ref_mp_lut(n, P, _p(buf[:n]), _p(buf[n:]))
    return buf


def ref_d(code, n, x=None):
    """run the reference double kernel (code 5, 6, 7): returns (input,
    output, (LUTr, LUTi), counters)."""
    xin = np.zeros(2 * n) if x is None else np.ascontiguousarray(x, dtype=np.float64).copy()
    y = np.zeros(2 * n)
    lr = np.zeros(max(n // 4, 1)); li = np.zeros(max(n // 4, 1))
    cnt = np.zeros(4, dtype=np.uint64)
    _ref(REF_D[code]).ref_d_run(n, 2 if x is None else 0, _p(xin[:n]), _p(xin[n:]),
                                _p(y[:n]), _p(y[n:]), _p(lr), _p(li), _p(cnt))
    return xin, y, (lr[:n // 4], li[:n // 4]), [int(v) for v in cnt]
