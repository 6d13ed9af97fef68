"""GPU checks at N = 2^20, one transform each, against the 2P-bit oracle
(computed on the host; ~10-40 s per transform with OpenMP):

  - the wide precisions, 512 and 2048 bits: the fixed form on the per-lane
    kernels and on the panels (the matrix-core twiddle products; a single
    transform is one panel with 63 zero transforms) within the fixed form's
    own bound (oracle.fixed_check), and no worse than the plain form, whose
    result is also checked: bit for bit against the oracle's reference-order
    transform at 512 bits, against the 2P-bit oracle at 2048 bits (its bits
    at 2048 are pinned by the reference digests up to 2^16);
  - the precision forms, double-word at 106 bits and quad-word at 212 bits,
    within their normwise bounds and no worse than the plain form at the
    same precision.
"""
import numpy as np
import pytest

import mpfft_amd as M
from mpfft_amd import inputs as I
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N20 = 1 << 20


def run(n, P, x, form):
    Lo = O.limbs(P)
    plan = M.Plan(n, P, form, 1)
    y = plan.execute_host(I.relimb(x, n, Lo, plan.L))
    L = plan.L
    plan.close()
    return L, I.relimb(y, n, L, Lo)


@pytest.mark.parametrize("P", [512, 2048])
def test_wide_fixed_at_2p20(P):
    n = N20
    print(f"case n=2^20 P={P}: oracle", flush=True)
    x = I.mp_random(n, P, seed=5100 + P)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    Lo, L2 = O.limbs(P), O.limbs(2 * P)
    _, yp = run(n, P, x, M.FORM_MP)
    if P <= 512:
        assert np.array_equal(yp, O.mp_fft(n, P, x)), "plain form differs from the oracle"
    mu_p, nw_p, _ = O.err_stats(n, P, yp, Lo, star, L2)
    for form in (M.FORM_FIXED_LANE, M.FORM_FIXED_PANEL):
        print(f"case n=2^20 P={P} form={form}", flush=True)
        L, y = run(n, P, x, form)
        exc, mu, nw, over = O.fixed_check(n, P, L, x, y, star)
        assert mu <= mu_p and nw <= nw_p, (form, mu, nw, mu_p, nw_p)
        print(f"form {form}: excess {exc:.3g} units, {mu:.3g} ulp, {nw:.3g} 2^-P "
              f"({over} beyond half an ulp); plain {mu_p:.3g} ulp, {nw_p:.3g} 2^-P")


@pytest.mark.parametrize("form,P", [(M.FORM_DD, 106), (M.FORM_QD, 212)])
def test_word_forms_at_2p20(form, P):
    n = N20
    print(f"case n=2^20 P={P} form={form}", flush=True)
    x = I.mp_random(n, P, seed=5200 + P)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    Lo, L2 = O.limbs(P), O.limbs(2 * P)
    _, y = run(n, P, x, form)
    mu, nw, zb = O.err_stats(n, P, y, Lo, star, L2)
    lg = 20
    bound = (1 + (2 + 3 * lg) * 2.0 ** (P - 104) if form == M.FORM_DD
             else 1 + (8 + 32 * lg) * 2.0 ** (P - 212))
    assert zb == 0 and nw <= bound, (nw, bound)
    _, yp = run(n, P, x, M.FORM_MP)
    mu_p, nw_p, _ = O.err_stats(n, P, yp, Lo, star, L2)
    assert nw <= nw_p, (nw, nw_p)
    print(f"{mu:.3g} ulp, {nw:.3g} 2^-P (bound {bound:.3g}); plain {mu_p:.3g} ulp, "
          f"{nw_p:.3g} 2^-P")
