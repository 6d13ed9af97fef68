"""GPU check of the block exponent of the fixed and word forms: their results
do not depend on the scale of the input.

Each transform's input is divided by 2^emax (emax = its largest exponent)
before the fixed-point or double-/quad-word arithmetic and the output is
multiplied by it afterwards; both are exact exponent shifts.  So:
  - x 2^s gives bit for bit the result of x, times 2^s (s = -300, +200 for
    the fixed form; -1000, +900 for the word forms, whose binary64 words
    could not hold such inputs without the shift);
  - every result meets the form's own bound against the 2P-bit oracle: the
    fixed form's componentwise contract (oracle.fixed_check: half an ulp
    plus 2 (log2 n + 2) n units of 2^(emax - F)), the word forms' normwise
    bounds;
  - the reference's IN=2 input (serial_mp_fft3.c:104-115: 1/(1+k^2),
    1/(1+k^4), values spread over 64 binary orders at n = 2^16) meets the
    same bounds.
"""
import numpy as np
import pytest

import mpfft_amd as M
from mpfft_amd import inputs as I
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_STAR = {}


def star(key, n, P, x):
    if key not in _STAR:
        _STAR[key] = O.mp_fft(n, P, x, Pw=2 * P)
    return _STAR[key]


def run(n, P, xs, form):
    Lo = O.limbs(P)
    plan = M.Plan(n, P, form, len(xs))
    y = plan.execute_host(np.concatenate([I.relimb(t, n, Lo, plan.L) for t in xs]))
    L, got = plan.L, plan.resolved_form()
    plan.close()
    return L, got, [I.relimb(t, n, L, Lo) for t in np.split(y, len(xs))]


@pytest.mark.parametrize("form", [M.FORM_FIXED, M.FORM_FIXED_PANEL, M.FORM_FIXED_LANE,
                                  M.FORM_AUTO])
@pytest.mark.parametrize("n,P", [(1024, 256), (1024, 1024), (65536, 256), (65536, 1024)])
def test_fixed_scale_invariant(form, n, P):
    print(f"case form={form} n={n} P={P}", flush=True)
    x = I.mp_random(n, P, seed=4100 + P)
    x2 = O.mp_input_in2(n, P)
    xs = [x, O.scale_pow2(x, n, -300), O.scale_pow2(x, n, 200), x2]
    L, _, (y, ylo, yhi, y2) = run(n, P, xs, form)
    assert np.array_equal(ylo, O.scale_pow2(y, n, -300))
    assert np.array_equal(yhi, O.scale_pow2(y, n, 200))
    exc, mu, nw, _ = O.fixed_check(n, P, L, x, y, star(("r", n, P), n, P, x))
    exc2, mu2, nw2, _ = O.fixed_check(n, P, L, x2, y2, star(("in2", n, P), n, P, x2))
    # the scaled inputs against the scaled oracle, independently of the above
    O.fixed_check(n, P, L, xs[1], ylo, O.scale_pow2(star(("r", n, P), n, P, x), n, -300))
    print(f"random: excess {exc:.3g} units, {mu:.3g} ulp, {nw:.3g} 2^-P; "
          f"IN=2: excess {exc2:.3g}, {mu2:.3g} ulp, {nw2:.3g} 2^-P")


def test_fixed_wide_range_inside_one_transform():
    # one real 2^40 times the others: the block exponent follows it, the
    # others keep their absolute accuracy 2^(emax - F) (no range error)
    n, P = 1024, 256
    x = I.mp_random(n, P, seed=5)
    h, _ = I.split(x, n, O.limbs(P))
    h[17] = np.int32((40 << 1) | (h[17] & 1))
    for form in (M.FORM_FIXED_LANE, M.FORM_FIXED_PANEL):
        L, _, (y,) = run(n, P, [x], form)
        O.fixed_check(n, P, L, x, y)


def test_exponent_outside_input_range():
    # the limb format's inputs have |exponent| < 2^29; beyond, the transform
    # is refused (MPFFT_ERR_VALUE) and the plan works again afterwards
    n = 256
    for form, P in ((M.FORM_FIXED_LANE, 256), (M.FORM_FIXED_PANEL, 256), (M.FORM_DD, 106),
                    (M.FORM_QD, 212)):
        for e in (1 << 29, -(1 << 29)):
            x = I.mp_random(n, P, seed=7)
            h, _ = I.split(x, n, O.limbs(P))
            h[:] = np.int32(-2**31)                 # all zero but one real
            h[9] = np.int32(e << 1)
            with pytest.raises(M.MpfftError) as err:
                run(n, P, [x], form)
            assert err.value.code == -3, (form, e)
        _, _, (y,) = run(n, P, [I.mp_random(n, P, seed=8)], form)
        assert y.size == 2 * n * (O.limbs(P) + 1)


def dd_bound(n, P, K):
    lg = max(n.bit_length() - 1, 0)
    if K == 2:
        return 1 + (2 + 3 * lg) * 2.0 ** (P - 104)
    return 1 + (8 + 32 * lg) * 2.0 ** (P - 212)


@pytest.mark.parametrize("form,P,K", [(M.FORM_DD, 106, 2), (M.FORM_AUTO, 106, 2),
                                      (M.FORM_QD, 212, 4)])
@pytest.mark.parametrize("n", [1024, 65536])
def test_word_scale_invariant(form, P, K, n):
    print(f"case form={form} n={n} P={P}", flush=True)
    x = I.mp_random(n, P, seed=4200 + P)
    x2 = O.mp_input_in2(n, P)
    xs = [x, O.scale_pow2(x, n, -1000), O.scale_pow2(x, n, 900), x2]
    _, _, (y, ylo, yhi, y2) = run(n, P, xs, form)
    assert np.array_equal(ylo, O.scale_pow2(y, n, -1000))
    assert np.array_equal(yhi, O.scale_pow2(y, n, 900))
    Lo, L2 = O.limbs(P), O.limbs(2 * P)
    for key, xi, yi in (("r", x, y), ("in2", x2, y2)):
        _, nw, zb = O.err_stats(n, P, yi, Lo, star((key, n, P), n, P, xi), L2)
        assert zb == 0 and nw <= dd_bound(n, P, K), (key, nw, dd_bound(n, P, K))


def test_word_tiny_reals_flushed():
    # reals more than 2^900 (double-word) or 2^800 (quad-word) below the
    # transform's largest are taken as zero: the normwise bound still holds
    n = 256
    for form, P, K, gap in ((M.FORM_DD, 106, 2, 901), (M.FORM_QD, 212, 4, 801)):
        x = I.mp_random(n, P, seed=8)
        h, _ = I.split(x, n, O.limbs(P))
        h[3] = np.int32((-gap << 1) | (h[3] & 1))
        _, _, (y,) = run(n, P, [x], form)
        _, nw, zb = O.err_stats(n, P, y, O.limbs(P), O.mp_fft(n, P, x, Pw=2 * P),
                                O.limbs(2 * P))
        assert nw <= dd_bound(n, P, K), (form, nw)


def test_batch_fixed_scale_invariant():
    # the batched form's fixed groups (panels and lanes) carry the block
    # exponent per transform
    ns = np.array([64, 64, 1024, 1024, 256, 256] * 2, dtype=np.int64)
    Ps = np.array([256, 256, 1024, 1024, 512, 512] * 2, dtype=np.int64)
    xs = [I.mp_random(int(n), int(P), seed=4300 + i) for i, (n, P) in enumerate(zip(ns, Ps))]
    shifts = [0, -300, 200, -1000, 77, 0] * 2
    b = M.Batch(ns, Ps, M.FORM_FIXED)

    def go(inp):
        import torch
        host = b.pack([I.relimb(t, int(n), O.limbs(int(P)), int(L))
                       for t, n, P, L in zip(inp, b.ns, b.Ps, b.limbs)])
        dx = torch.from_numpy(host.view(np.int32)).cuda()
        dy = torch.zeros_like(dx)
        b.execute_device(dx.data_ptr(), dy.data_ptr())
        b.check()
        return [I.relimb(t, int(n), int(L), O.limbs(int(P)))
                for t, n, P, L in zip(b.unpack(dy.cpu().numpy().view(np.uint32)), b.ns, b.Ps,
                                      b.limbs)]

    y0 = go(xs)
    y1 = go([O.scale_pow2(t, int(n), s) for t, n, s in zip(xs, ns, shifts)])
    for i, (n, s) in enumerate(zip(ns, shifts)):
        assert np.array_equal(y1[i], O.scale_pow2(y0[i], int(n), s)), i
    for i in (0, 2, 4):
        O.fixed_check(int(ns[i]), int(Ps[i]), int(b.limbs[i]), xs[i], y0[i])
    b.close()
