"""GPU check of the double-word form (MPFFT_FORM_DD, 53 <= P <= 106) through
the C ABI.

The form has no counterpart in the reference; its contract is an error
bound against the 2P-bit MPFR oracle (the oracle's FFT at 2P bits, whose
own error is far below 2^-P):
  - normwise error (units of 2^-P) at most 1 + (2 + 3 log2 n) 2^(P-104):
    the final rounding to P bits (relative error below 2^-P per
    component), plus the double-word arithmetic (each radix-4 pass at most
    ~12 u^2 relative, u^2 = 2^-106: DWPlusDW below 3u^2, DWTimesDW below
    5u^2, per Joldes-Muller-Popescu 2017, two levels of additions and one
    complex product per element; a radix-16 pass is two such steps), which
    the measured errors stay far below (P = 106: 1.3 at n = 2^10 .. 2^12,
    1.6 at 2^16);
  - no output is nonzero where the exact transform is zero;
  - closed forms: an impulse gives exactly 1 everywhere.
"""
import numpy as np
import pytest
import torch

import mpfft_amd as M
from mpfft_amd import inputs as I
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def run(n, P, xs, form=M.FORM_DD):
    Lo = O.limbs(P)
    plan = M.Plan(n, P, form, len(xs))
    y = plan.execute_host(np.concatenate([I.relimb(t, n, Lo, plan.L) for t in xs]))
    plan.close()
    return [I.relimb(t, n, plan.L, Lo) for t in np.split(y, len(xs))]


def bound(n, P):
    lg = max(1, n.bit_length() - 1)
    return 1.0 + (2 + 3 * lg) * 2.0 ** (P - 104)


def check(n, P, x, y):
    Lo, L2 = O.limbs(P), O.limbs(2 * P)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    mu, nw, zb = O.err_stats(n, P, y, Lo, star, L2)
    assert zb == 0, (n, P)
    assert nw <= bound(n, P), (n, P, nw, bound(n, P))
    return mu, nw


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("P", [53, 64, 100, 106])
def test_dd_error_bound(n, P):
    xs = [I.mp_random(n, P, seed=1300 + n + P + 7 * i) for i in range(3)]
    for x, y in zip(xs, run(n, P, xs)):
        mu, nw = check(n, P, x, y)
    print(f"n={n} P={P}: max {mu:.3g} ulp, normwise {nw:.3g} 2^-P (bound {bound(n, P):.3g})")


def test_dd_beats_plain_at_53_bits():
    # at P = 53 the double-word arithmetic carries ~104 bits: the result is
    # the exact transform rounded once, at most 0.5 ulp per component away
    # from the 2P-bit oracle (where the oracle itself is exact enough),
    # while the plain form accumulates a rounding error per operation
    n, P = 1024, 53
    x = I.mp_random(n, P, seed=17)
    mu, nw = check(n, P, x, run(n, P, [x])[0])
    Lo, L2 = O.limbs(P), O.limbs(2 * P)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    mu_p, nw_p, _ = O.err_stats(n, P, O.mp_fft(n, P, x), Lo, star, L2)
    assert nw <= nw_p and mu <= mu_p, (mu, nw, mu_p, nw_p)


def test_dd_full_size():
    n, P = 1 << 16, 106
    x = I.mp_random(n, P, seed=99)
    mu, nw = check(n, P, x, run(n, P, [x])[0])
    Lo, L2 = O.limbs(P), O.limbs(2 * P)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    mu_p, nw_p, _ = O.err_stats(n, P, O.mp_fft(n, P, x), Lo, star, L2)
    print(f"n={n} P={P}: max {mu:.3g} ulp, normwise {nw:.3g} 2^-P; "
          f"plain form {mu_p:.3g} ulp, {nw_p:.3g} 2^-P")
    # the chosen form at 54 .. 106 bits is this one: no worse than the plain
    # form here, componentwise and normwise
    assert nw <= nw_p and mu <= mu_p, (mu, nw, mu_p, nw_p)


def impulse(n, P):
    L = O.limbs(P)
    hdr = np.full(2 * n, np.int32(-2**31), dtype=np.int32)
    limbs = np.zeros((L, 2 * n), dtype=np.uint32)
    hdr[0] = 2
    limbs[0, 0] = 0x80000000
    return I.join(hdr, limbs)


@pytest.mark.parametrize("n", [1, 2, 16, 1024])
def test_dd_impulse(n):
    P = 106
    for y in run(n, P, [impulse(n, P)] * 2):
        h, lm = I.split(y, n, O.limbs(P))
        assert (h[:n] == 2).all() and (lm[0, :n] == 0x80000000).all() and (lm[1:, :n] == 0).all()
        assert ((h[n:] & ~1) == np.int32(-2**31)).all()


def test_dd_batch_and_stages():
    # transforms of one batch stay apart; the stages one at a time give the
    # whole transform bit for bit
    n, P, b = 2048, 100, 5
    xs = [I.mp_random(n, P, seed=500 + i) for i in range(b)]
    ys = run(n, P, xs)
    for x, y in zip(xs, ys):
        check(n, P, x, y)
    plan = M.Plan(n, P, M.FORM_DD, b)
    S = plan.stages()
    assert S == 4                        # log2 n = 11: radix 2, 16, 4, 16
    x = np.concatenate([I.relimb(t, n, O.limbs(P), plan.L) for t in xs])
    dx = torch.from_numpy(x.view(np.int32)).cuda()
    d1, d2 = torch.empty_like(dx), torch.empty_like(dx)
    plan.execute_device(dx.data_ptr(), d1.data_ptr())
    for s in range(S):
        plan.execute_range(dx.data_ptr(), d2.data_ptr(), s, s + 1)
    plan.check()
    assert torch.equal(d1, d2)
    plan.close()


@pytest.mark.parametrize("n", [16, 32, 64, 128])
def test_dd_many_small_transforms(n):
    # small transforms share the radix-16 pass's blocks (64 butterflies per
    # block, butterfly g of the grid is butterfly g mod n/16 of transform
    # g / (n/16)): 100 transforms fill several blocks and leave the last one
    # partial; each result is within the bound, and equals its transform run
    # alone bit for bit
    P, b = 106, 100
    xs = [I.mp_random(n, P, seed=9100 + 7 * i + n) for i in range(b)]
    ys = run(n, P, xs)
    for i in (0, 1, 37, 63, 64, 99):
        check(n, P, xs[i], ys[i])
        assert np.array_equal(ys[i], run(n, P, [xs[i]])[0]), i


def test_dd_precision_and_range_checked():
    for P in (52, 107, 128):
        with pytest.raises(M.MpfftError) as e:
            M.Plan(64, P, M.FORM_DD, 1)
        assert e.value.code == -2
    n, P = 256, 106
    x = I.mp_random(n, P, seed=8)
    h, _ = I.split(x, n, O.limbs(P))
    h[3] = np.int32(((1 << 29) << 1) | (h[3] & 1))    # beyond the limb format's range
    with pytest.raises(M.MpfftError) as e:
        run(n, P, [x])
    assert e.value.code == -3
    assert len(run(n, P, [I.mp_random(n, P, seed=9)])) == 1


# ---- quad-word form (MPFFT_FORM_QD, 53 <= P <= 212) --------------------------
#
# Contract: normwise error (units of 2^-P) at most 1 + (8 + 32 log2 n)
# 2^(P-212): the final rounding, plus the word arithmetic, whose order-3
# terms are rounded (a product ~10 u^4, a sum ~4 u^4, u^4 = 2^-212; per
# radix-4 pass one complex product and two levels of sums, ~64 u^4 at
# most, log4 n passes).  Measured at P = 212: 0.5-0.7 at n = 16 .. 4096,
# 0.76 at 2^16 (the plain form: 2.9).


def qbound(n, P):
    lg = max(1, n.bit_length() - 1)
    return 1.0 + (8 + 32 * lg) * 2.0 ** (P - 212)


def qcheck(n, P, x, y):
    Lo, L2 = O.limbs(P), O.limbs(2 * P)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    mu, nw, zb = O.err_stats(n, P, y, Lo, star, L2)
    assert zb == 0, (n, P)
    assert nw <= qbound(n, P), (n, P, nw, qbound(n, P))
    return mu, nw


@pytest.mark.parametrize("n", [1, 2, 4, 16, 64, 1024, 2048, 4096])
@pytest.mark.parametrize("P", [53, 106, 160, 200, 212])
def test_qd_error_bound(n, P):
    xs = [I.mp_random(n, P, seed=2300 + n + P + 7 * i) for i in range(3)]
    for x, y in zip(xs, run(n, P, xs, M.FORM_QD)):
        mu, nw = qcheck(n, P, x, y)
    print(f"qd n={n} P={P}: max {mu:.3g} ulp, normwise {nw:.3g} 2^-P (bound {qbound(n, P):.3g})")


def test_qd_full_size():
    n, P = 1 << 16, 212
    x = I.mp_random(n, P, seed=98)
    mu, nw = qcheck(n, P, x, run(n, P, [x], M.FORM_QD)[0])
    Lo, L2 = O.limbs(P), O.limbs(2 * P)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    mu_p, nw_p, _ = O.err_stats(n, P, O.mp_fft(n, P, x), Lo, star, L2)
    print(f"qd n={n} P={P}: max {mu:.3g} ulp, normwise {nw:.3g} 2^-P; "
          f"plain form {mu_p:.3g} ulp, {nw_p:.3g} 2^-P")


@pytest.mark.parametrize("n", [1, 16, 1024])
def test_qd_impulse(n):
    P = 212
    for y in run(n, P, [impulse(n, P)] * 2, M.FORM_QD):
        h, lm = I.split(y, n, O.limbs(P))
        assert (h[:n] == 2).all() and (lm[0, :n] == 0x80000000).all() and (lm[1:, :n] == 0).all()
        assert ((h[n:] & ~1) == np.int32(-2**31)).all()


def test_qd_batch_stages_and_checks():
    n, P, b = 512, 180, 3
    xs = [I.mp_random(n, P, seed=600 + i) for i in range(b)]
    for x, y in zip(xs, run(n, P, xs, M.FORM_QD)):
        qcheck(n, P, x, y)
    plan = M.Plan(n, P, M.FORM_QD, b)
    S = plan.stages()
    assert S == 3                        # log2 n = 9: radix 2, 16, 16
    x = np.concatenate([I.relimb(t, n, O.limbs(P), plan.L) for t in xs])
    dx = torch.from_numpy(x.view(np.int32)).cuda()
    d1, d2 = torch.empty_like(dx), torch.empty_like(dx)
    plan.execute_device(dx.data_ptr(), d1.data_ptr())
    for s in range(S):
        plan.execute_range(dx.data_ptr(), d2.data_ptr(), s, s + 1)
    plan.check()
    assert torch.equal(d1, d2)
    plan.close()
    for Pb in (52, 213):
        with pytest.raises(M.MpfftError) as e:
            M.Plan(64, Pb, M.FORM_QD, 1)
        assert e.value.code == -2
    xb = I.mp_random(256, 212, seed=8)
    h, _ = I.split(xb, 256, O.limbs(212))
    h[3] = np.int32(((1 << 29) << 1) | (h[3] & 1))
    with pytest.raises(M.MpfftError) as e:
        run(256, 212, [xb], M.FORM_QD)
    assert e.value.code == -3
