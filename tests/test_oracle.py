"""The oracle (oracle/liboracle.so, a from-scratch C + MPFR restatement) is
pinned here before anything is checked against it:

* bit-exact against the golden outputs of the reference's own code
  (tests/golden, generated from the compiled reference);
* against the live reference build (oracle/_ref) when it is present;
* against closed forms and invariants that do not depend on the reference
  (impulse, constant, textbook N <= 4, brute-force DFT at 2P bits,
  linearity, the round trip of the reference's test(), Yavne op counts).
"""
import hashlib

import numpy as np
import pytest

from mpfft_amd import inputs as I
from oracle import oracle as O


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


MP_SIZES = [1, 2, 4, 8, 16, 64, 1024]
MP_PRECS = [53, 64, 128, 256, 1024]


@pytest.mark.parametrize("n", MP_SIZES)
@pytest.mark.parametrize("P", MP_PRECS)
def test_oracle_matches_reference_golden(n, P, mp_golden, digests):
    x = I.mp_random(n, P, seed=digests["seed"])
    assert sha(x) == digests["mp"][f"N{n}_P{P}"]["in_sha"], "input generator drifted"
    y = O.mp_fft(n, P, x)
    assert np.array_equal(y, mp_golden[f"out_N{n}_P{P}"])


CODE_CASES = [(c, n, P) for c in (1, 2) for n in MP_SIZES
              for P in (64, 256) + ((1024,) if n <= 64 else ())]


@pytest.mark.parametrize("code,n,P", CODE_CASES)
def test_oracle_codes_match_reference_golden(code, n, P, mp_golden):
    # the alternative operation orders: serial_mp_fft1.c (4m2a, one
    # butterfly) and serial_mp_fft2.c (4m2a, three butterflies)
    x = I.mp_random(n, P, seed=I.DEFAULT_SEED)
    assert np.array_equal(O.mp_fft(n, P, x, code=code), mp_golden[f"c{code}_out_N{n}_P{P}"])


@pytest.mark.parametrize("code", [1, 2])
def test_oracle_codes_large_digest(code, digests):
    n, P = 1 << 16, 256
    x = I.mp_random(n, P, seed=digests["seed"])
    d = digests["mp_codes_large"][f"c{code}_N{n}_P{P}"]
    assert sha(x) == d["in_sha"]
    assert sha(O.mp_fft(n, P, x, code=code)) == d["out_sha"]


def test_codes_agree_where_no_general_butterfly_runs():
    # up to n = 8 every combine butterfly is k = 0 or k = n/8, which CODE 2
    # and CODE 3 treat alike; CODE 1 runs the general formula there too
    P = 256
    for n in (1, 2, 4, 8):
        x = I.mp_random(n, P, seed=30 + n)
        assert np.array_equal(O.mp_fft(n, P, x, code=2), O.mp_fft(n, P, x, code=3))
    x = I.mp_random(16, P, seed=31)
    assert not np.array_equal(O.mp_fft(16, P, x, code=2), O.mp_fft(16, P, x, code=3))
    x = I.mp_random(64, P, seed=32)
    assert not np.array_equal(O.mp_fft(64, P, x, code=1), O.mp_fft(64, P, x, code=2))


@pytest.mark.parametrize("code", [1, 2])
@pytest.mark.parametrize("n", [16, 256])
def test_codes_within_bound_of_2p(code, n):
    P = 128
    x = I.mp_random(n, P, seed=n + code)
    y = O.mp_fft(n, P, x, code=code)
    y2 = O.mp_fft(n, P, x, Pw=2 * P)
    maxulp, norm, zb = O.err_stats(n, P, y, O.limbs(P), y2, O.limbs(2 * P))
    assert zb == 0 and norm < 8.0


@pytest.mark.parametrize("code", [1, 2])
def test_codes_impulse_gives_all_ones(code):
    n, P = 64, 256
    x = _mp_const(n, P, 0.0, 0.0)
    hdr, lim = I.split(x, n, O.limbs(P))
    hdr[0] = 1 << 1
    lim[0, 0] = 0x80000000
    assert np.all(I.to_float(O.mp_fft(n, P, x, code=code), n, O.limbs(P)) == 1.0)


@pytest.mark.parametrize("name", ["impulse", "constant", "zeros"])
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 64])
@pytest.mark.parametrize("P", [128, 1024])
def test_oracle_special_inputs_golden(name, n, P, mp_golden):
    x = mp_golden[f"{name}_in_N{n}_P{P}"]
    assert np.array_equal(O.mp_fft(n, P, x), mp_golden[f"{name}_out_N{n}_P{P}"])


@pytest.mark.parametrize("n,P", [(64, 64), (64, 256), (64, 1024)])
def test_oracle_in2_input_and_output(n, P, mp_golden):
    # the reference's deterministic input (ic() IN=2) and its transform
    x = O.mp_input_in2(n, P)
    assert np.array_equal(x, mp_golden[f"in2_in_N{n}_P{P}"])
    assert np.array_equal(O.mp_fft(n, P, x), mp_golden[f"in2_out_N{n}_P{P}"])


def test_oracle_in2_1024(mp_golden, digests):
    x = O.mp_input_in2(1024, 1024)
    assert sha(x) == digests["in2_in_sha_N1024_P1024"]
    assert np.array_equal(O.mp_fft(1024, 1024, x), mp_golden["in2_out_N1024_P1024"])


def test_oracle_large_digest(digests):
    n, P = 1 << 16, 256
    x = I.mp_random(n, P, seed=digests["seed"])
    assert sha(x) == digests["mp_large"][f"N{n}_P{P}"]["in_sha"]
    assert sha(O.mp_fft(n, P, x)) == digests["mp_large"][f"N{n}_P{P}"]["out_sha"]


WIDE_CASES = [(3, 16, 2048), (2, 16, 2048), (3, 1024, 2048), (2, 1024, 2048),
              (3, 4096, 1500), (2, 4096, 1500), (3, 1 << 16, 2048)]


@pytest.mark.parametrize("code,n,P", WIDE_CASES)
def test_oracle_wide_digest(code, n, P, digests):
    # above 1024 bits (64 limbs): the reference's own outputs as digests
    x = I.mp_random(n, P, seed=digests["seed"])
    d = digests["mp_wide"][f"c{code}_N{n}_P{P}"]
    assert sha(x) == d["in_sha"]
    assert sha(O.mp_fft(n, P, x, code=code)) == d["out_sha"]


@pytest.mark.parametrize("n", [1024, 1 << 16])
def test_op_counts(n, digests):
    # reference counters (serial_mp_fft3.c:123-250) and Yavne's total
    c = O.counts(n)
    assert c == digests["mp_counts"][str(n)]
    lg = n.bit_length() - 1
    assert c[0] == n * lg - 3 * n + 4
    assert c[1] == 3 * n * lg - 3 * n + 4
    assert c[0] + c[1] == 4 * n * lg - 6 * n + 8


def _mp_const(n, P, re, im):
    """SoA transform with every element equal to re + i*im (exact dyadics)."""
    L = O.limbs(P)
    hdr = np.zeros(2 * n, dtype=np.int32)
    lim = np.zeros((L, 2 * n), dtype=np.uint32)
    for part, v in ((0, re), (1, im)):
        sl = slice(part * n, (part + 1) * n)
        if v == 0:
            hdr[sl] = np.int32(-2**31)
            continue
        m, e = np.frexp(abs(v))
        hdr[sl] = (int(e) << 1) | (1 if v < 0 else 0)
        lim[0, sl] = np.uint32(int(m * 2**32))
    return I.join(hdr, lim)


@pytest.mark.parametrize("n", [8, 64, 1024])
def test_impulse_gives_all_ones(n):
    P = 256
    x = _mp_const(n, P, 0.0, 0.0)
    hdr, lim = I.split(x, n, O.limbs(P))
    hdr[0] = 1 << 1
    lim[0, 0] = 0x80000000
    y = I.to_float(O.mp_fft(n, P, x), n, O.limbs(P))
    assert np.all(y == 1.0)


@pytest.mark.parametrize("n", [4, 16, 256])
def test_constant_input(n):
    P = 128
    y = O.mp_fft(n, P, _mp_const(n, P, 0.75, -0.5))
    v = I.to_float(y, n, O.limbs(P))
    assert v[0] == n * (0.75 - 0.5j)
    # the other bins cancel exactly apart from twiddle rounding
    assert np.max(np.abs(v[1:])) < n * 2.0 ** -(P - 8)


def test_textbook_small_sizes():
    P = 64
    for n in (1, 2, 4):
        x = I.mp_random(n, P, seed=3)
        v = I.to_float(x, n, O.limbs(P))
        y = I.to_float(O.mp_fft(n, P, x), n, O.limbs(P))
        ref = np.fft.fft(v)
        assert np.allclose(y, ref, rtol=0, atol=1e-15)


@pytest.mark.parametrize("n", [8, 32, 64])
def test_bruteforce_dft_at_2p(n):
    """direct DFT in exact rational-free form: compare the P-bit FFT with the
    2P-bit FFT and both with numpy's DFT to double accuracy."""
    P = 128
    x = I.mp_random(n, P, seed=n)
    y = O.mp_fft(n, P, x)
    y2 = O.mp_fft(n, P, x, Pw=2 * P)
    maxulp, norm, zb = O.err_stats(n, P, y, O.limbs(P), y2, O.limbs(2 * P))
    assert zb == 0 and norm < 8.0
    v = I.to_float(x, n, O.limbs(P))
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ v
    assert np.allclose(I.to_float(y2, n, O.limbs(2 * P)), dft, rtol=0, atol=1e-12)


def test_linearity():
    n, P = 256, 256
    a = I.mp_random(n, P, seed=1)
    fa = I.to_float(O.mp_fft(n, P, a), n, O.limbs(P))
    # doubling every input doubles every output exactly (exponent shift)
    hdr, lim = I.split(a.copy(), n, O.limbs(P))
    b = a.copy()
    hb, _ = I.split(b, n, O.limbs(P))
    nz = (hb & ~1) != -2**31
    hb[nz] += 2
    fb = I.to_float(O.mp_fft(n, P, b), n, O.limbs(P))
    assert np.array_equal(fb, 2 * fa)


def test_roundtrip_rrmse():
    # the reference's accuracy contract: rRMSE ~ 2^-P (serial/mp/OPTIONS.md:43)
    n, P = 1024, 256
    x = I.mp_random(n, P, seed=2)
    r = O.roundtrip(n, P, x, O.mp_fft(n, P, x))
    assert np.all(r[:2] < 2.0 ** -(P - 6)) and np.all(r[:2] > 0)


def test_err_stats_closed_forms():
    """the error measure itself: zero against an exact copy held at 2P bits,
    exactly one ulp after changing the last bit of one component, and the
    zero-mismatch count"""
    n, P = 64, 128
    L, L2 = O.limbs(P), O.limbs(2 * P)
    y = O.mp_fft(n, P, I.mp_random(n, P, seed=4))
    star = I.relimb(y, n, L, L2)                 # same values, 2P-bit container
    assert O.err_stats(n, P, y, L, star, L2) == (0.0, 0.0, 0)
    bad = y.copy()
    h, lim = I.split(bad, n, L)
    k = int(np.nonzero((h & ~1) != -2**31)[0][5])
    lim[L - 1, k] ^= 1                           # one ulp at P = 128 = 32 L
    mu, nw, zb = O.err_stats(n, P, bad, L, star, L2)
    assert mu == 1.0 and 0 < nw < 1.0 and zb == 0
    z = star.copy()
    hz, lz = I.split(z, n, L2)
    hz[k] = -2**31
    lz[:, k] = 0
    assert O.err_stats(n, P, y, L, z, L2)[2] == 1


def _mant(lim, k):
    """the limbs of component k as one integer, most significant limb first"""
    v = 0
    for w in lim[:, k]:
        v = (v << 32) | int(w)
    return v


def _set_mant(lim, k, v):
    for i in range(lim.shape[0] - 1, -1, -1):
        lim[i, k] = v & 0xFFFFFFFF
        v >>= 32


def _exact_pair(n=64, P=128, seed=4):
    L, L2 = O.limbs(P), O.limbs(2 * P)
    x = I.mp_random(n, P, seed=seed)
    y = O.mp_fft(n, P, x)
    h, _ = I.split(y, n, L)
    k = int(np.argmax(np.where((h & ~1) == -2**31, -2**31, h)))   # largest |y_k|
    return n, P, L, L2, x, y, I.relimb(y, n, L, L2), k, int(h[k]) >> 1


def test_err_excess_closed_forms():
    """the excess beyond half an ulp: 0 against an exact copy; 0 at exactly
    half an ulp (the half-ulp term has the right exponent); a quarter ulp
    past it reads 0.5 in units of half an ulp; a one-ulp change reads 1 such
    unit; a zero result against a nonzero value reads that value's size (the
    unit's sign)"""
    n, P, L, L2, _, y, star, k, e = _exact_pair()
    half = e - P - 1                              # half an ulp of y_k = 2^half
    assert O.err_excess(n, P, y, L, star, L2, half) == (0.0, 0)
    s = star.copy()
    _, ls = I.split(s, n, L2)
    m = _mant(ls, k)
    _set_mant(ls, k, m + (1 << (32 * L2 - P - 1)))          # + ulp/2
    assert O.err_excess(n, P, y, L, s, L2, half) == (0.0, 0)
    _set_mant(ls, k, m + (3 << (32 * L2 - P - 2)))          # + 3/4 ulp
    assert O.err_excess(n, P, y, L, s, L2, half) == (0.5, 1)
    assert O.err_excess(n, P, y, L, s, L2, half - 3) == (4.0, 1)
    bad = y.copy()
    _, lb = I.split(bad, n, L)
    lb[L - 1, k] ^= 1                                        # one ulp
    assert O.err_excess(n, P, bad, L, star, L2, half) == (1.0, 1)
    hz, lz = I.split(bad, n, L)
    hz[k], lz[:, k] = -2**31, 0                              # y_k := 0
    exc, over = O.err_excess(n, P, bad, L, star, L2, e)
    _, lstar = I.split(star, n, L2)
    assert over == 1 and exc == pytest.approx(_mant(lstar, k) / 2.0 ** (32 * L2), rel=1e-15)
    assert 0.5 <= exc < 1.0


def test_max_exponent_and_scale_pow2():
    """the block exponent is the largest header exponent (signs and zeros
    ignored, None for all zeros); scaling by 2^s moves every nonzero
    header exponent by s and keeps the signs, limbs and zeros"""
    n, L = 4, 2
    hdr = np.array([(-5 << 1) | 1, (3 << 1) | 0, -2**31, -2**31 + 1,
                    (-7 << 1) | 1, (2 << 1) | 1, -2**31, (1 << 1)], dtype=np.int32)
    lim = np.full((L, 2 * n), 0x80000000, dtype=np.uint32)
    lim[:, [2, 3, 6]] = 0
    x = I.join(hdr, lim)
    assert O.max_exponent(x, n) == 3
    zeros = I.join(np.full(2 * n, -2**31, dtype=np.int32), np.zeros((L, 2 * n), np.uint32))
    assert O.max_exponent(zeros, n) is None
    for s in (-9, 0, 5):
        y = O.scale_pow2(x, n, s)
        hy, ly = I.split(y, n, L)
        nz = (hdr & ~1) != -2**31
        assert (hy[nz] >> 1 == (hdr[nz] >> 1) + s).all()
        assert (hy & 1 == hdr & 1).all() and (hy[~nz] == hdr[~nz]).all()
        assert np.array_equal(ly, lim) and O.max_exponent(y, n) == 3 + s
    # exact scaling commutes with the correctly rounded transform
    xr = I.mp_random(64, 128, seed=8)
    assert np.array_equal(O.mp_fft(64, 128, O.scale_pow2(xr, 64, -77)),
                          O.scale_pow2(O.mp_fft(64, 128, xr), 64, -77))


def test_fixed_check_accepts_and_rejects():
    """the fixed form's check: a result equal to the exact values passes;
    an extra 900 units of 2^(emax - F) past half an ulp passes at n = 64
    (bound 2 (6 + 2) 64 = 1024) and reads 900; 1100 units fail; a one-ulp
    change fails (a mistake in F, emax or the bound's direction would fail
    one of these)"""
    n, P, L, L2, x, y, star, k, e = _exact_pair()
    Ld = 4
    F = 32 * (Ld + 1)
    emax = O.max_exponent(x, n)
    assert O.fixed_bound(n) == 1024.0
    exc, mu, nw, over = O.fixed_check(n, P, Ld, x, y, star)
    assert (exc, mu, nw, over) == (0.0, 0.0, 0.0, 0)
    j = e - (emax - F)                        # fraction bit of 2^(emax - F) in y_k
    assert P + 12 < j <= 32 * L2
    s = star.copy()
    _, ls = I.split(s, n, L2)
    m = _mant(ls, k) + (1 << (32 * L2 - P - 1))
    _set_mant(ls, k, m + (900 << (32 * L2 - j)))
    exc, _, _, over = O.fixed_check(n, P, Ld, x, y, s)
    assert exc == 900.0 and over == 1
    _set_mant(ls, k, m + (1100 << (32 * L2 - j)))
    with pytest.raises(AssertionError):
        O.fixed_check(n, P, Ld, x, y, s)
    bad = y.copy()
    _, lb = I.split(bad, n, L)
    lb[L - 1, k] ^= 1
    with pytest.raises(AssertionError):
        O.fixed_check(n, P, Ld, x, bad, star)


def test_planted_mistake_is_caught(mp_golden, digests):
    """a one-ulp change anywhere in a 1024-point output must fail the
    bit-exact comparison (the comparison is not vacuous)."""
    n, P = 1024, 1024
    y = O.mp_fft(n, P, I.mp_random(n, P, seed=digests["seed"])).copy()
    L = O.limbs(P)
    y[2 * n + (L - 1) * 2 * n + 777] ^= 1
    assert not np.array_equal(y, mp_golden[f"out_N{n}_P{P}"])


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 1024])
@pytest.mark.parametrize("code", [5, 6, 7])
def test_double_oracle_golden(n, code, d_golden, digests):
    x = I.d_random(n, seed=digests["seed"])
    y = O.d_fft(code, n, x)
    assert np.array_equal(y.view(np.uint64), d_golden[f"out{code}_N{n}"].view(np.uint64))


def test_double_lut_matches_reference(d_golden):
    wr, wi, _ = O.d_lut(1024)
    assert np.array_equal(wr, d_golden["lutr_N1024"])
    assert np.array_equal(wi, d_golden["luti_N1024"])


@pytest.mark.skipif(not O.have_ref(), reason="reference build (oracle/_ref) not present")
def test_oracle_vs_live_reference():
    for n, P in [(512, 200), (128, 96), (2048, 333), (256, 2048), (64, 1800)]:
        x = I.mp_random(n, P, seed=n + P)
        assert np.array_equal(O.mp_fft(n, P, x), O.ref_mp("mp3", n, P, x)[1])
        for code in (1, 2):
            assert np.array_equal(O.mp_fft(n, P, x, code=code),
                                  O.ref_mp(f"mp{code}", n, P, x)[1])
    assert np.array_equal(O.mp_lut(256, 512), O.ref_mp_lut(256, 512))
