"""GPU check of the fixed form's panel kernels (MPFFT_FORM_FIXED with batch a
multiple of 64: transforms interleaved in panels of 64, twiddle products on
the int8 matrix cores), through the C ABI.

The contract is the fixed form's (oracle.fixed_check): every component
within half an ulp plus 2 (log2 n + 2) n units of 2^(emax - F) of the
2P-bit oracle.  Two more checks
are specific to the panels:
  - panel and per-lane kernels (MPFFT_FORM_FIXED_LANE) agree within 2 ulp on
    every element of every transform: both stay within a small multiple of
    2^-F (F = P + 32 bits and more) of the exact transform before their one
    rounding to P bits, so they can differ only where the exact value lies
    next to a rounding boundary;
  - every transform of a batch has its own input, so a panel or lane index
    that mixes transforms shows as an O(1) error.
"""
import numpy as np
import pytest

import mpfft_amd as M
from mpfft_amd import inputs as I
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def run(n, P, xs, form):
    Lo = O.limbs(P)
    plan = M.Plan(n, P, form, len(xs))
    y = plan.execute_host(np.concatenate([I.relimb(t, n, Lo, plan.L) for t in xs]))
    plan.close()
    return [I.relimb(t, n, plan.L, Lo) for t in np.split(y, len(xs))]


def inputs(n, P, b, seed):
    return [I.mp_random(n, P, seed=seed + 7919 * i) for i in range(b)]


def check_against_lane(n, P, yp, yl):
    Lo = O.limbs(P)
    for i in range(len(yp)):
        mu, nw, zb = O.err_stats(n, P, yp[i], Lo, yl[i], Lo)
        assert zb == 0 and mu <= 2.0, (i, mu, nw)


def check_against_oracle(n, P, x, y):
    Lo, L2 = O.limbs(P), O.limbs(2 * P)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    _, _, zb_f = O.err_stats(n, P, y, Lo, star, L2)
    assert zb_f == 0
    O.fixed_check(n, P, M.limbs_for_precision(P), x, y, star)


@pytest.mark.parametrize("n,P", [(16, 1024), (32, 256), (64, 53), (128, 32), (256, 64),
                                 (512, 100), (1024, 512), (2048, 1024), (4096, 128),
                                 (8192, 256),
                                 # every other compiled width: W = 5, 8, 14, 26 (W = 8:
                                 # the twiddle tables live outside the slots)
                                 (1024, 96), (2048, 192), (512, 160), (1024, 384),
                                 (256, 768)])
def test_panel_matches_lane_and_oracle(n, P):
    xs = inputs(n, P, 64, seed=9000 + n + P)
    yp = run(n, P, xs, M.FORM_FIXED_PANEL)
    yl = run(n, P, xs, M.FORM_FIXED_LANE)
    check_against_lane(n, P, yp, yl)
    for i in (0, 21, 63):
        check_against_oracle(n, P, xs[i], yp[i])


def test_panel_two_groups():
    # batch 128: two groups of 64 transforms per pass
    n, P = 256, 256
    xs = inputs(n, P, 128, seed=4242)
    yp = run(n, P, xs, M.FORM_FIXED_PANEL)
    yl = run(n, P, xs, M.FORM_FIXED_LANE)
    check_against_lane(n, P, yp, yl)
    check_against_oracle(n, P, xs[100], yp[100])


def test_panel_stage_ranges():
    # the stages run one at a time give the whole transform, bit for bit
    import torch
    n, P, b = 1024, 256, 64
    xs = inputs(n, P, b, seed=77)
    plan = M.Plan(n, P, M.FORM_FIXED_PANEL, b)
    x = np.concatenate([I.relimb(t, n, O.limbs(P), plan.L) for t in xs])
    S = plan.stages()
    # input conversion, a radix-4 pass (log2 n = 10 = 2 + 2 * 4), two radix-16
    # passes, output conversion
    assert S == 1 + 3 + 1
    assert M.Plan(n, 1024, M.FORM_FIXED_PANEL, b).stages() == 1 + 3 + 1   # W = 34 likewise
    assert M.Plan(n, 2048, M.FORM_FIXED_PANEL, b).stages() == 1 + 5 + 1   # W = 66: radix 4
    dx = torch.from_numpy(x.view(np.int32)).cuda()
    d1, d2 = torch.empty_like(dx), torch.empty_like(dx)
    plan.execute_device(dx.data_ptr(), d1.data_ptr())
    for s in range(S):
        plan.execute_range(dx.data_ptr(), d2.data_ptr(), s, s + 1)
    plan.check()
    assert torch.equal(d1, d2)
    plan.close()


def test_panel_input_range_checked():
    n, P = 1024, 256
    xs = inputs(n, P, 64, seed=5)
    h, _ = I.split(xs[40], n, O.limbs(P))
    h[17] = np.int32(((1 << 29) << 1) | (h[17] & 1))   # beyond the limb format's range
    with pytest.raises(M.MpfftError) as e:
        run(n, P, xs, M.FORM_FIXED_PANEL)
    assert e.value.code == -3
    y = run(n, P, inputs(n, P, 64, seed=6), M.FORM_FIXED_PANEL)
    assert len(y) == 64


def test_panel_full_size_1024_bits():
    n, P = 1 << 16, 1024
    xs = inputs(n, P, 64, seed=31)
    yp = run(n, P, xs, M.FORM_FIXED)
    yl = run(n, P, xs, M.FORM_FIXED_LANE)
    check_against_lane(n, P, yp, yl)
    for i in (0, 63):
        check_against_oracle(n, P, xs[i], yp[i])


def impulse(n, P, pos):
    """x[pos] = 1 (0.1b * 2^1), every other real +0"""
    L = O.limbs(P)
    hdr = np.full(2 * n, np.int32(-2**31), dtype=np.int32)
    limbs = np.zeros((L, 2 * n), dtype=np.uint32)
    hdr[pos] = 2
    limbs[0, pos] = 0x80000000
    return I.join(hdr, limbs)


@pytest.mark.parametrize("form,batch", [(M.FORM_FIXED, 64), (M.FORM_FIXED_LANE, 2)])
@pytest.mark.parametrize("n,P", [(16, 1024), (32, 256), (1024, 64), (2048, 1024)])
def test_fixed_impulse(form, batch, n, P):
    # x = delta_0: every output exactly 1 + 0i (each pass adds zeros and
    # multiplies only zeros by twiddles); x = delta_1: X[k] = w^k, within the
    # plain form's error against the 2P-bit oracle
    Lo = O.limbs(P)
    for y in run(n, P, [impulse(n, P, 0)] * batch, form):
        h, lm = I.split(y, n, Lo)
        assert (h[:n] == 2).all() and (lm[0, :n] == 0x80000000).all() and (lm[1:, :n] == 0).all()
        assert ((h[n:] & ~1) == np.int32(-2**31)).all() and (lm[:, n:] == 0).all()
    x1 = impulse(n, P, 1)
    ys = run(n, P, [x1] * batch, form)
    check_against_oracle(n, P, x1, ys[0])
    check_against_oracle(n, P, x1, ys[-1])


def test_panel_large_n():
    # N = 2^20 (plan: five radix-16 passes; buffers of
    # 64 x 2^21 reals), P = 128: every transform against the per-lane kernels,
    # two against the 2P-bit oracle
    n, P = 1 << 20, 128
    # 16 random transforms, each also scaled by 2^-1, 2^-2, 2^-3 (an exponent
    # shift: exact, and a transform that took another's data would differ
    # by a power of two)
    base = inputs(n, P, 16, seed=2020)
    xs = []
    for sh in range(4):
        for x in base:
            h, lm = I.split(x.copy(), n, O.limbs(P))
            nz = (h & ~1) != np.int32(-2**31)
            h[nz] -= np.int32(2 * sh)
            xs.append(I.join(h, lm))
    plan = M.Plan(n, P, M.FORM_FIXED_PANEL, 64)
    assert plan.panel_plan() == 1 and plan.stages() == 1 + 5 + 1
    plan.close()
    yp = run(n, P, xs, M.FORM_FIXED_PANEL)
    yl = run(n, P, xs, M.FORM_FIXED_LANE)
    check_against_lane(n, P, yp, yl)
    for i in (0, 63):
        check_against_oracle(n, P, xs[i], yp[i])


@pytest.mark.parametrize("n,P,b", [(16, 2048, 64), (256, 1500, 64), (4096, 2048, 3)])
def test_panel_2048_bits(n, P, b):
    # reals of 66 words (P up to 2048, 64 limbs): the panel kernels only;
    # against the 2P-bit oracle, no worse than the plain order (the oracle's
    # restatement of the reference at P bits)
    xs = inputs(n, P, b, seed=2048 + n)
    plan = M.Plan(n, P, M.FORM_FIXED, b)
    assert plan.L == 64
    plan.close()
    yp = run(n, P, xs, M.FORM_FIXED)
    for i in sorted({0, b - 1}):
        check_against_oracle(n, P, xs[i], yp[i])


def test_panel_2048_bits_forms_checked():
    # beyond 1024 bits the word forms refuse (and the panels below n = 16);
    # the plain op orders and both fixed-form kernel families run up to 2048
    # bits
    for form in (M.FORM_DD, M.FORM_QD):
        with pytest.raises(M.MpfftError) as e:
            M.Plan(64, 2048, form, 1)
        assert e.value.code == -2
    for form in (M.FORM_MP, M.FORM_MP_C1, M.FORM_MP_C2, M.FORM_FIXED, M.FORM_FIXED_LANE):
        M.Plan(64, 2048, form, 1).close()
        with pytest.raises(M.MpfftError) as e:
            M.Plan(64, 2049, form, 1)
        assert e.value.code == -2
    with pytest.raises(M.MpfftError) as e:
        M.Plan(8, 2048, M.FORM_FIXED_PANEL, 64)
    assert e.value.code == -1
    M.Plan(8, 2048, M.FORM_FIXED, 64).close()           # per-lane kernels below n = 16


def test_panel_2048_bits_large_n_impulse():
    # N = 2^20 at 2048 bits, one transform (a padded panel): x = delta_0
    # gives exactly 1 everywhere (ten passes of zeros added and multiplied;
    # the random-input error contract is checked against the oracle above)
    n, P = 1 << 20, 2048
    y = run(n, P, [impulse(n, P, 0)], M.FORM_FIXED)[0]
    h, lm = I.split(y, n, O.limbs(P))
    assert (h[:n] == 2).all() and (lm[0, :n] == 0x80000000).all() and (lm[1:, :n] == 0).all()
    assert ((h[n:] & ~1) == np.int32(-2**31)).all() and (lm[:, n:] == 0).all()


# ---- radix-16 passes (MPFFT_PANEL_RADIX=16) ----------------------------------
#
# The radix-16 panel pass (eight waves, 16 positions of 32 transforms, the
# DFT16 as 4 x 4 through LDS, two twiddles per wave on the matrix cores) is
# the default plan up to 34 words per real; MPFFT_PANEL_RADIX=16 makes it
# explicit here, and the radix-4 and radix-8 passes it replaced stay
# reachable (MPFFT_PANEL_RADIX=4 / 8) and tested below.  Same bar as above.

@pytest.fixture
def radix16(monkeypatch):
    monkeypatch.setenv("MPFFT_PANEL_RADIX", "16")


@pytest.mark.parametrize("n,P", [(16, 1024), (32, 256), (64, 53), (128, 32), (256, 1024),
                                 (512, 100), (2048, 1024), (4096, 768), (8192, 512),
                                 (1 << 16, 1024)])
def test_panel_radix16(radix16, n, P):
    print(f"case radix16 n={n} P={P}", flush=True)
    b = 64
    plan = M.Plan(n, P, M.FORM_FIXED_PANEL, b)
    lg = n.bit_length() - 1
    r = lg % 4
    W = plan.L + 2
    first = 0 if r == 0 else (2 if (r == 3 and W > 18) else 1)
    assert plan.stages() == 2 + first + lg // 4, (plan.stages(), lg)
    plan.close()
    xs = inputs(n, P, b, seed=9700 + n + P)
    yp = run(n, P, xs, M.FORM_FIXED_PANEL)
    yl = run(n, P, xs, M.FORM_FIXED_LANE)
    check_against_lane(n, P, yp, yl)
    for i in (0, 33, 63):
        check_against_oracle(n, P, xs[i], yp[i])


def test_panel_radix16_stage_ranges(radix16):
    import torch
    n, P, b = 4096, 1024, 70                   # two panels, the second partial
    xs = inputs(n, P, b, seed=78)
    plan = M.Plan(n, P, M.FORM_FIXED_PANEL, b)
    x = np.concatenate([I.relimb(t, n, O.limbs(P), plan.L) for t in xs])
    S = plan.stages()
    assert S == 1 + 3 + 1
    dx = torch.from_numpy(x.view(np.int32)).cuda()
    d1, d2 = torch.empty_like(dx), torch.empty_like(dx)
    plan.execute_device(dx.data_ptr(), d1.data_ptr())
    for s in range(S):
        plan.execute_range(dx.data_ptr(), d2.data_ptr(), s, s + 1)
    plan.check()
    assert torch.equal(d1, d2)
    plan.close()
    y = np.split(d1.cpu().numpy().view(np.uint32), b)
    for i in (0, 64, 69):
        check_against_oracle(n, P, xs[i], I.relimb(y[i], n, M.limbs_for_precision(P), O.limbs(P)))


@pytest.mark.parametrize("rad,n,P", [("4", 256, 1024), ("4", 8192, 256), ("8", 1024, 256),
                                     ("8", 128, 32), ("4", 1 << 16, 1024)])
def test_panel_other_radixes(monkeypatch, rad, n, P):
    # the radix-4 and radix-8 plans, on request
    monkeypatch.setenv("MPFFT_PANEL_RADIX", rad)
    print(f"case radix{rad} n={n} P={P}", flush=True)
    xs = inputs(n, P, 64, seed=9900 + n + P)
    yp = run(n, P, xs, M.FORM_FIXED_PANEL)
    yl = run(n, P, xs, M.FORM_FIXED_LANE)
    check_against_lane(n, P, yp, yl)
    check_against_oracle(n, P, xs[5], yp[5])


# ---- four-step plan (MPFFT_PANEL_PLAN=fs; n = 16^m >= 4096, up to 34 words) --
#
# n = n1 256: radix-16 passes over n1 (columns 32 consecutive (transform, n2);
# the first converts the limb input), the twiddles w^(k1 n2) in the first
# pass over 256 points (columns 32 transforms), its last pass rounding into
# the limb output; tiled buffers between the passes.  The same contract and
# the same cross-check against the per-lane kernels as above; batches that
# leave a block of 32 columns or transforms partial; the default plan (with
# conversion passes) against it.

@pytest.fixture
def four_step(monkeypatch):
    monkeypatch.setenv("MPFFT_PANEL_PLAN", "fs")


@pytest.mark.parametrize("n,P,b", [(4096, 1024, 64), (4096, 256, 70), (4096, 53, 3),
                                   (4096, 128, 33), (1 << 16, 256, 40), (4096, 160, 64),
                                   (4096, 768, 32), (1 << 16, 1024, 64)])
def test_four_step_plan(four_step, n, P, b):
    print(f"case four-step n={n} P={P} batch={b}", flush=True)
    plan = M.Plan(n, P, M.FORM_FIXED_PANEL, b)
    lg = n.bit_length() - 1
    assert plan.panel_plan() == 2 and plan.stages() == lg // 4
    plan.close()
    xs = inputs(n, P, b, seed=12000 + n + P + b)
    yp = run(n, P, xs, M.FORM_FIXED_PANEL)
    yl = run(n, P, xs, M.FORM_FIXED_LANE)
    check_against_lane(n, P, yp, yl)
    for i in sorted({0, b // 2, b - 1}):
        check_against_oracle(n, P, xs[i], yp[i])


def test_four_step_large_n(four_step):
    # n = 2^20: three passes over n1 = 4096, then the twist and the last pass
    n, P, b = 1 << 20, 128, 32
    print(f"case four-step n={n} P={P} batch={b}", flush=True)
    plan = M.Plan(n, P, M.FORM_FIXED_PANEL, b)
    assert plan.panel_plan() == 2 and plan.stages() == 5
    plan.close()
    xs = inputs(n, P, b, seed=12077)
    yp = run(n, P, xs, M.FORM_FIXED_PANEL)
    yl = run(n, P, xs, M.FORM_FIXED_LANE)
    check_against_lane(n, P, yp, yl)
    check_against_oracle(n, P, xs[31], yp[31])


def test_four_step_stage_ranges(four_step):
    # its stages run one at a time give the whole transform, bit for bit
    import torch
    n, P, b = 1 << 16, 128, 35
    xs = inputs(n, P, b, seed=79)
    plan = M.Plan(n, P, M.FORM_FIXED_PANEL, b)
    assert plan.panel_plan() == 2
    x = np.concatenate([I.relimb(t, n, O.limbs(P), plan.L) for t in xs])
    S = plan.stages()
    assert S == 4
    dx = torch.from_numpy(x.view(np.int32)).cuda()
    d1, d2 = torch.empty_like(dx), torch.empty_like(dx)
    plan.execute_device(dx.data_ptr(), d1.data_ptr())
    for s in range(S):
        plan.execute_range(dx.data_ptr(), d2.data_ptr(), s, s + 1)
    plan.check()
    assert torch.equal(d1, d2)
    plan.close()
    y = np.split(d1.cpu().numpy().view(np.uint32), b)
    for i in (0, 34):
        check_against_oracle(n, P, xs[i], I.relimb(y[i], n, M.limbs_for_precision(P), O.limbs(P)))


def test_four_step_against_default_plan(monkeypatch):
    # the default plan (conversion passes) and the four-step plan on the same
    # inputs: both within the contract, and within 2 ulp of each other
    n, P, b = 4096, 1024, 64
    xs = inputs(n, P, b, seed=81)
    plan = M.Plan(n, P, M.FORM_FIXED_PANEL, b)
    assert plan.panel_plan() == 1 and plan.stages() == 1 + 3 + 1
    plan.close()
    yd = run(n, P, xs, M.FORM_FIXED_PANEL)
    monkeypatch.setenv("MPFFT_PANEL_PLAN", "fs")
    y4 = run(n, P, xs, M.FORM_FIXED_PANEL)
    check_against_lane(n, P, y4, yd)
    for i in (0, 63):
        check_against_oracle(n, P, xs[i], y4[i])
        check_against_oracle(n, P, xs[i], yd[i])


def test_four_step_input_range_checked(four_step):
    # the first pass converts the input: an exponent beyond the limb
    # format's range is flagged there
    n, P = 4096, 256
    xs = inputs(n, P, 32, seed=15)
    h, _ = I.split(xs[20], n, O.limbs(P))
    h[4000] = np.int32(((1 << 29) << 1) | (h[4000] & 1))
    with pytest.raises(M.MpfftError) as e:
        run(n, P, xs, M.FORM_FIXED_PANEL)
    assert e.value.code == -3
    y = run(n, P, inputs(n, P, 32, seed=16), M.FORM_FIXED_PANEL)
    assert len(y) == 32
