"""GPU check of the fixed form (MPFFT_FORM_FIXED: radix-4 Stockham on
fixed-point reals), through the C ABI.

This form does not reproduce the reference's bit pattern; its contract is
accuracy (oracle.fixed_check): every component within half an ulp of the
result plus 2 (log2 n + 2) n units of 2^(emax - F) of the exact transform
(the 2P-bit oracle), F = 32 (L + 1) >= P + 32, emax the largest input
exponent; at P = 1024, n = 2^16 that absolute part is ~2^-20 of an ulp of a
typical output.  The plain form (the reference's order, bit for bit) is
10^5 ulp away at large n, so the bar is the form's own, not the plain
form's.  A wrong index or twiddle anywhere would show as an O(1) error.
"""
import numpy as np
import pytest

import mpfft_amd as M
from mpfft_amd import inputs as I
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def run_fixed(n, P, x, batch=1, form=M.FORM_FIXED):
    Lo = O.limbs(P)
    plan = M.Plan(n, P, form, batch)
    xs = np.concatenate([I.relimb(t, n, Lo, plan.L) for t in np.split(x, batch)])
    y = plan.execute_host(xs)
    plan.close()
    return np.concatenate([I.relimb(t, n, plan.L, Lo) for t in np.split(y, batch)])


def errors(n, P, y, star):
    return O.err_stats(n, P, y, O.limbs(P), star, O.limbs(2 * P))


@pytest.mark.parametrize("n,P", [(1, 128), (2, 128), (4, 256), (8, 256), (16, 1024), (32, 64),
                                 (64, 1024), (128, 100), (1024, 256), (2048, 1024),
                                 (4096, 53), (8192, 512)])
def test_fixed_error_within_bound(n, P):
    x = I.mp_random(n, P, seed=1000 + n + P)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    fixed = run_fixed(n, P, x)
    _, _, zb_f = errors(n, P, fixed, star)
    assert zb_f == 0
    O.fixed_check(n, P, M.limbs_for_precision(P), x, fixed, star)


@pytest.mark.parametrize("n,P,b", [(1, 2048, 1), (4, 1500, 2), (16, 2048, 1), (64, 1056, 3),
                                   (1024, 2048, 1), (4096, 2000, 2)])
def test_fixed_lane_wide(n, P, b):
    # the per-lane kernels at 66-word reals (28-bit digit products): against
    # the 2P-bit oracle no worse than the plain form, and within 2 ulp of the
    # panel kernels (the two differ only next to a rounding boundary)
    xs = [I.mp_random(n, P, seed=1900 + n + P + i) for i in range(b)]
    plan = M.Plan(n, P, M.FORM_FIXED_LANE, b)
    assert plan.L == 64
    plan.close()
    print(f"case n={n} P={P} b={b}", flush=True)
    got = np.split(run_fixed(n, P, np.concatenate(xs), batch=b, form=M.FORM_FIXED_LANE), b)
    for x, y in zip(xs, got):
        star = O.mp_fft(n, P, x, Pw=2 * P)
        _, _, zb_f = errors(n, P, y, star)
        assert zb_f == 0
        O.fixed_check(n, P, 64, x, y, star)
    if n >= 16:
        yp = np.split(run_fixed(n, P, np.concatenate(xs), batch=b, form=M.FORM_FIXED_PANEL), b)
        for y, z in zip(got, yp):
            mu, _, zb = O.err_stats(n, P, y, O.limbs(P), z, O.limbs(P))
            assert zb == 0 and mu <= 2.0, (mu,)


@pytest.mark.parametrize("n", [1, 64, 4096])
def test_fixed_lane_wide_impulse(n):
    # 66-word reals on the per-lane kernels: x = delta_0 gives exactly 1
    P = 2048
    L = O.limbs(P)
    hdr = np.full(2 * n, np.int32(-2**31), dtype=np.int32)
    lim = np.zeros((L, 2 * n), dtype=np.uint32)
    hdr[0], lim[0, 0] = 2, 0x80000000
    y = run_fixed(n, P, I.join(hdr, lim), form=M.FORM_FIXED_LANE)
    h, lm = I.split(y, n, L)
    assert (h[:n] == 2).all() and (lm[0, :n] == 0x80000000).all() and (lm[1:, :n] == 0).all()
    assert ((h[n:] & ~1) == np.int32(-2**31)).all() and (lm[:, n:] == 0).all()


def test_fixed_wide_small_batch_takes_lanes():
    # above 1024 bits MPFFT_FORM_FIXED picks the per-lane kernels where a
    # panel would be mostly padding, and the panels at a full batch
    lane = M.Plan(4096, 2048, M.FORM_FIXED, 1)
    pan = M.Plan(4096, 2048, M.FORM_FIXED_PANEL, 1)
    full = M.Plan(4096, 2048, M.FORM_FIXED, 64)
    assert lane.stages() == 6                    # six radix-4 passes, no conversions
    assert pan.stages() != lane.stages() and full.stages() == pan.stages()
    for p in (lane, pan, full):
        p.close()


def test_fixed_full_size_1024_bits():
    n, P = 1 << 16, 1024
    x = I.mp_random(n, P, seed=77)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    plain = O.mp_fft(n, P, x)
    fixed = run_fixed(n, P, x)
    mu_f, nw_f, _ = errors(n, P, fixed, star)
    mu_p, nw_p, _ = errors(n, P, plain, star)
    assert nw_f <= nw_p and mu_f <= mu_p, (mu_f, nw_f, mu_p, nw_p)
    O.fixed_check(n, P, 32, x, fixed, star)


def test_fixed_batch():
    n, P, b = 512, 256, 3
    xs = [I.mp_random(n, P, seed=300 + i) for i in range(b)]
    got = np.split(run_fixed(n, P, np.concatenate(xs), batch=b), b)
    for i in range(b):
        assert np.array_equal(got[i], run_fixed(n, P, xs[i]))


@pytest.mark.parametrize("n,P", [(16, 256), (32, 64), (64, 256), (256, 128), (512, 200),
                                 (1024, 256), (2048, 64)])
def test_fixed_lane_radix16_many_transforms(n, P):
    # the per-lane radix-16 passes (up to 256 bits) with 100 transforms: small
    # transforms share a block (64 butterflies), passes with Ns < 64 restage
    # their outputs through LDS, the last block is partial; each result is
    # the same transform run alone, bit for bit, and some are within the
    # fixed form's bound against the 2P-bit oracle
    b = 100
    xs = [I.mp_random(n, P, seed=4300 + 3 * i + n + P) for i in range(b)]
    got = np.split(run_fixed(n, P, np.concatenate(xs), batch=b, form=M.FORM_FIXED_LANE), b)
    for i in (0, 1, 37, 63, 64, 99):
        assert np.array_equal(got[i], run_fixed(n, P, xs[i], form=M.FORM_FIXED_LANE)), i
    for i in (0, 99):
        star = O.mp_fft(n, P, xs[i], Pw=2 * P)
        O.fixed_check(n, P, M.limbs_for_precision(P), xs[i], got[i], star)


def test_fixed_input_range_checked():
    # the limb format's exponent range: |exponent| < 2^29 (a real 2^40 times
    # the others is accepted since the block exponent follows the largest)
    n, P = 1024, 256
    x = I.mp_random(n, P, seed=5)
    h, _ = I.split(x, n, O.limbs(P))
    h[17] = np.int32(((1 << 29) << 1) | (h[17] & 1))
    with pytest.raises(M.MpfftError) as e:
        run_fixed(n, P, x)
    assert e.value.code == -3
    # the plan recovers: a good call right after succeeds
    y = run_fixed(n, P, I.mp_random(n, P, seed=6))
    assert y.size == x.size
