"""GPU parity of the alternative operation orders, through the C ABI:
CODE 1 (lab/serial/mp/serial_mp_fft1.c: 4m2a, one butterfly for every k)
and CODE 2 (serial_mp_fft2.c: 4m2a, special butterflies k = 0 and n/8).

Bar: bit-exact against the compiled reference's outputs (tests/golden) and
against the oracle, which tests/test_oracle.py pins to them.
"""
import hashlib

import numpy as np
import pytest

import mpfft_amd as M
from mpfft_amd import inputs as I
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FORMS = {1: M.FORM_MP_C1, 2: M.FORM_MP_C2}
CASES = [(c, n, P) for c in (1, 2) for n in [1, 2, 4, 8, 16, 64, 1024]
         for P in (64, 256) + ((1024,) if n <= 64 else ())]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(code, n, P, x, batch=1):
    Lo = O.limbs(P)
    plan = M.Plan(n, P, FORMS[code], batch)
    xs = np.concatenate([I.relimb(t, n, Lo, plan.L) for t in np.split(x, batch)])
    y = plan.execute_host(xs)
    plan.close()
    return np.concatenate([I.relimb(t, n, plan.L, Lo) for t in np.split(y, batch)])


@pytest.mark.parametrize("code,n,P", CASES)
def test_codes_match_reference_golden(code, n, P, mp_golden):
    x = I.mp_random(n, P, seed=I.DEFAULT_SEED)
    assert np.array_equal(run(code, n, P, x), mp_golden[f"c{code}_out_N{n}_P{P}"])


@pytest.mark.parametrize("code", [1, 2])
@pytest.mark.parametrize("n,P,seed", [(4096, 1024, 5), (2048, 100, 6), (512, 333, 7),
                                      (1024, 2048, 8), (256, 1111, 9)])
def test_codes_match_oracle(code, n, P, seed):
    x = I.mp_random(n, P, seed=seed)
    assert np.array_equal(run(code, n, P, x), O.mp_fft(n, P, x, code=code))


@pytest.mark.parametrize("code", [1, 2])
def test_codes_large_digest(code, digests):
    n, P = 1 << 16, 256
    x = I.mp_random(n, P, seed=digests["seed"])
    d = digests["mp_codes_large"][f"c{code}_N{n}_P{P}"]
    assert sha(x) == d["in_sha"]
    assert sha(run(code, n, P, x)) == d["out_sha"]


@pytest.mark.parametrize("code", [1, 2])
def test_codes_batch(code):
    n, P, b = 256, 1024, 3
    xs = [I.mp_random(n, P, seed=90 + i) for i in range(b)]
    got = np.split(run(code, n, P, np.concatenate(xs), batch=b), b)
    for i in range(b):
        assert np.array_equal(got[i], O.mp_fft(n, P, xs[i], code=code))
