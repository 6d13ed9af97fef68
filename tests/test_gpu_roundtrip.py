"""GPU inverse transform and round-trip error measures, the reference's
test() (lab/serial/mp/serial_mp_fft3.c:346-427), through the C ABI.

The inverse is the reference's swap trick: forward transform of the
swapped input (re <-> im), division by N, swap back.  Bar: bit-exact
against the oracle's forward transform composed with that trick.  The
round-trip statistics (rRMSE per part, max |error|) are accumulated in
double on the GPU and at P bits by the oracle: relative tolerance 1e-9.
"""
import numpy as np
import pytest

import mpfft_amd as M
from mpfft_amd import inputs as I
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ZERO = np.int32(-2**31)


def swap_parts(buf, n, L):
    """exchange the real and imaginary halves of every row of a transform"""
    h, lim = I.split(np.array(buf, dtype=np.uint32), n, L)
    h2 = np.concatenate([h[n:], h[:n]])
    l2 = np.concatenate([lim[:, n:], lim[:, :n]], axis=1)
    return I.join(h2, l2)


def scale_pow2(buf, n, L, lg):
    """divide every nonzero real by 2^lg (exponent shift)"""
    h, lim = I.split(np.array(buf, dtype=np.uint32), n, L)
    h = h.copy()
    nz = (h & ~1) != ZERO
    h[nz] = (((h[nz] >> 1) - lg) << 1) | (h[nz] & 1)
    return I.join(h, lim)


def oracle_inverse(n, P, X):
    """swap(FFT(swap(X))) / n with the oracle's forward transform"""
    L = O.limbs(P)
    lg = n.bit_length() - 1
    return swap_parts(scale_pow2(O.mp_fft(n, P, swap_parts(X, n, L)), n, L, lg), n, L)


def device_inverse(n, P, X, batch=1):
    torch = pytest.importorskip("torch")
    plan = M.Plan(n, P, M.FORM_MP, batch)
    Lo = O.limbs(P)
    xs = np.concatenate([I.relimb(t, n, Lo, plan.L) for t in np.split(X, batch)])
    dx = torch.from_numpy(xs.view(np.int32)).cuda()
    dy = torch.empty_like(dx)
    plan.execute_inverse(dx.data_ptr(), dy.data_ptr())
    plan.sync()
    y = dy.cpu().numpy().view(np.uint32)
    plan.close()
    return np.concatenate([I.relimb(t, n, plan.L, Lo) for t in np.split(y, batch)])


@pytest.mark.parametrize("n,P", [(1, 128), (2, 128), (4, 256), (8, 64), (64, 1024),
                                 (1024, 256), (2048, 1024), (512, 100), (256, 2048)])
def test_inverse_matches_oracle(n, P):
    x = I.mp_random(n, P, seed=700 + n + P)
    X = O.mp_fft(n, P, x)
    got = device_inverse(n, P, X)
    assert np.array_equal(got, oracle_inverse(n, P, X))


def test_inverse_batch():
    n, P, b = 256, 512, 3
    Xs = [O.mp_fft(n, P, I.mp_random(n, P, seed=800 + i)) for i in range(b)]
    got = np.split(device_inverse(n, P, np.concatenate(Xs), batch=b), b)
    for i in range(b):
        assert np.array_equal(got[i], oracle_inverse(n, P, Xs[i]))


@pytest.mark.parametrize("n,P", [(1024, 2048), (4096, 1500), (16, 1056)])
def test_roundtrip_wide_error_scale(n, P):
    # 64 limbs: the statistics themselves underflow the doubles (2^-P), so
    # the check is on their log2 fields: the reference's round trip of a
    # correctly rounded P-bit transform stays within a few units of 2^-P
    x = I.mp_random(n, P, seed=950 + n + P)
    plan = M.Plan(n, P, M.FORM_MP, 1)
    st = plan.roundtrip_host(I.relimb(x, n, O.limbs(P), plan.L))[0]
    plan.close()
    for j in (4, 5):                       # log2 rRMSE, real and imaginary parts
        assert -P - 4 < st[j] < -P + 6, (j, st[j])
    for j in (6, 7):                       # log2 max |error|
        assert -P - 4 < st[j] < -P + 8, (j, st[j])


@pytest.mark.parametrize("n,P", [(64, 128), (1024, 256), (1024, 1024), (256, 53)])
def test_roundtrip_stats_match_oracle(n, P):
    x = I.mp_random(n, P, seed=900 + n + P)
    plan = M.Plan(n, P, M.FORM_MP, 1)
    Lo = O.limbs(P)
    st = plan.roundtrip_host(I.relimb(x, n, Lo, plan.L))[0]
    plan.close()
    X = O.mp_fft(n, P, x)
    ref = O.roundtrip(n, P, x, X)          # rRMSE re, im; max|err| re, im (P-bit sums)
    for j in range(4):
        assert ref[j] > 0
        # compare through log2: the doubles are subnormal near P = 1024
        assert abs(st[4 + j] - np.log2(ref[j])) < 1e-6, (j, st, ref)
        assert st[j] == pytest.approx(ref[j], rel=1e-6)


def test_roundtrip_full_size_accuracy():
    # N = 2^16 at 1024 bits: the round trip stays within a few ulps (rRMSE
    # of the reference's test() is ~2^-P times a small factor)
    n, P = 1 << 16, 1024
    plan = M.Plan(n, P, M.FORM_MP, 1)
    x = I.mp_random(n, P, seed=17, L=plan.L)
    st = plan.roundtrip_host(x)[0]
    plan.close()
    assert -P - 8 < st[4] < -P + 6 and -P - 8 < st[5] < -P + 6, st
    assert st[6] < -P + 8 and st[7] < -P + 8, st


def test_inverse_rejects_double_plan():
    plan = M.Plan(64, 0, M.FORM_D6, 1)
    with pytest.raises(M.MpfftError) as e:
        plan.execute_inverse(1, 1)
    assert e.value.code == -8
    plan.close()
