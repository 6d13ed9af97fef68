"""GPU parity of the multiprecision plain form, through the C ABI.

Bar: bit-exact.  Every check compares the HIP result with the reference's
own outputs (tests/golden, made by the compiled reference) or with the
oracle, which tests/test_oracle.py pins to them.
"""
import ctypes
import hashlib

import numpy as np
import pytest

import mpfft_amd as M
from mpfft_amd import inputs as I
from mpfft_amd import mpfr_arrays as MA
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MP_SIZES = [1, 2, 4, 8, 16, 64, 1024]
MP_PRECS = [53, 64, 128, 256, 1024]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_fft(n, P, x, batch=1, form=M.FORM_MP):
    """x: SoA-limb transform(s) at ceil(P/32) limbs; result likewise."""
    Lo = O.limbs(P)
    plan = M.Plan(n, P, form, batch)
    xs = np.concatenate([I.relimb(t, n, Lo, plan.L)
                         for t in np.split(np.asarray(x, dtype=np.uint32), batch)])
    y = plan.execute_host(xs)
    plan.close()
    return np.concatenate([I.relimb(t, n, plan.L, Lo) for t in np.split(y, batch)])


# ---- device arithmetic ------------------------------------------------------

def operands(count, P, seed):
    """count reals x, y with structured exponent gaps, cancellations, ties
    and signed zeros (SoA at ceil(P/32) limbs)."""
    L = O.limbs(P)
    rng = np.random.default_rng(seed)
    x = I.mp_random(count // 2, P, seed=seed)
    y = I.mp_random(count // 2, P, seed=seed + 1)
    hx, lx = I.split(x, count // 2, L)
    hy, ly = I.split(y, count // 2, L)
    ex = hx >> 1
    gaps = rng.choice([0, 0, 1, 2, 31, 32, 33, 64, P - 1, P, P + 1, 32 * L + 31, 32 * L + 33,
                       3 * P], size=count)
    gaps = gaps * rng.choice([1, -1], size=count)
    hy[:] = (((ex - gaps) << 1) | (hy & 1)).astype(np.int32)
    # cancellation: y = x with low bits changed
    c = rng.random(count) < 0.15
    ly[:, c] = lx[:, c]
    hy[c] = hx[c] ^ rng.integers(0, 2, size=int(c.sum()), dtype=np.int32)
    ly[L - 1, c] ^= np.uint32(1 << (32 * L - P))
    # signed zeros
    for h, lim in ((hx, lx), (hy, ly)):
        z = rng.random(count) < 0.03
        h[z] = np.int32(-2**31) | rng.integers(0, 2, size=int(z.sum()), dtype=np.int32)
        lim[:, z] = 0
    return x, y


@pytest.mark.parametrize("P", [53, 64, 100, 128, 160, 256, 333, 512, 700, 1024, 1500, 2048])
@pytest.mark.parametrize("op", [0, 1, 2])
def test_device_ops_match_mpfr(P, op):
    count = 8192
    x, y = operands(count, P, seed=P * 3 + op)
    ref = O.mp_binop(op, count, P, x, y)
    Ld, Lo = M.limbs_for_precision(P), O.limbs(P)
    got = M.mp_binop(op, I.relimb(x, count // 2, Lo, Ld), I.relimb(y, count // 2, Lo, Ld),
                     count, P)
    got = I.relimb(got, count // 2, Ld, Lo)
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, f"first mismatching word {bad[:5]}"


# ---- FFT parity with the reference's outputs --------------------------------

@pytest.mark.parametrize("n", MP_SIZES)
@pytest.mark.parametrize("P", MP_PRECS)
def test_fft_matches_reference_golden(n, P, mp_golden, digests):
    x = I.mp_random(n, P, seed=digests["seed"])
    print(f"case n={n} P={P}", flush=True)
    y = run_fft(n, P, x)
    assert np.array_equal(y, mp_golden[f"out_N{n}_P{P}"])


@pytest.mark.parametrize("name", ["impulse", "constant", "zeros"])
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 64])
@pytest.mark.parametrize("P", [128, 1024])
def test_fft_special_inputs(name, n, P, mp_golden):
    y = run_fft(n, P, mp_golden[f"{name}_in_N{n}_P{P}"])
    assert np.array_equal(y, mp_golden[f"{name}_out_N{n}_P{P}"])


@pytest.mark.parametrize("n,P", [(64, 64), (64, 256), (64, 1024)])
def test_fft_reference_in2_input(n, P, mp_golden):
    y = run_fft(n, P, mp_golden[f"in2_in_N{n}_P{P}"])
    assert np.array_equal(y, mp_golden[f"in2_out_N{n}_P{P}"])


@pytest.mark.parametrize("n,P,seed", [(4096, 256, 1), (2048, 1024, 2), (512, 100, 3),
                                      (256, 333, 4), (128, 700, 5), (8192, 64, 6),
                                      (32, 32 * 24, 7), (1 << 13, 128, 8)])
def test_fft_matches_oracle(n, P, seed):
    x = I.mp_random(n, P, seed=seed)
    print(f"case n={n} P={P} seed={seed}", flush=True)
    assert np.array_equal(run_fft(n, P, x), O.mp_fft(n, P, x))


# subtree size S finished in LDS (MPFFT_SUB_SIZE; 0: every level on HBM,
# None: the default), sizes below, at and above the transform size, both
# operation orders (CODE 3, CODE 1)
@pytest.mark.parametrize("n,P,S", [(4096, 256, "0"), (4096, 256, None), (4096, 256, "8"),
                                   (4096, 256, "16"), (4096, 256, "64"), (4096, 256, "8192"),
                                   (2048, 512, "256"), (2048, 512, "32"), (1 << 14, 64, "2048"),
                                   (1024, 200, "128"), (8, 128, "512"), (16, 128, "8"),
                                   (32, 300, "512"), (1 << 15, 128, "512")])
def test_subtrees_in_lds_any_size(monkeypatch, n, P, S):
    if S is None:
        monkeypatch.delenv("MPFFT_SUB_SIZE", raising=False)
    else:
        monkeypatch.setenv("MPFFT_SUB_SIZE", S)
    x = I.mp_random(n, P, seed=9000 + n + P)
    print(f"case n={n} P={P} S={S}", flush=True)
    for code, form in ((3, M.FORM_MP), (1, M.FORM_MP_C1)):
        assert np.array_equal(run_fft(n, P, x, form=form), O.mp_fft(n, P, x, code=code)), code


def test_subtree_stages(monkeypatch):
    # stage 0 holds the leaves and the levels finished in LDS; the rest are
    # one combine level each, and running them one by one gives the whole
    # transform
    monkeypatch.setenv("MPFFT_SUB_SIZE", "512")
    n, P = 1 << 14, 256
    plan = M.Plan(n, P, M.FORM_MP, 2)
    assert plan.stages() == 1 + (14 - 9)          # levels 2^10 .. 2^14 on HBM
    x = np.concatenate([I.relimb(I.mp_random(n, P, seed=77 + i), n, O.limbs(P), plan.L)
                        for i in range(2)])
    whole = plan.execute_host(x)
    import torch
    dx = torch.from_numpy(x.view(np.int32)).cuda()
    dy = torch.empty_like(dx)
    st = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    for s in range(plan.stages()):
        plan.execute_range(dx.data_ptr(), dy.data_ptr(), s, s + 1, st)
    torch.cuda.synchronize()
    assert np.array_equal(dy.cpu().numpy().view(np.uint32), whole)
    plan.close()


@pytest.mark.parametrize("key", ["N65536_P256", "N65536_P1024", "N1048576_P128"])
def test_fft_large_matches_reference_digest(key, digests):
    n = int(key.split("_")[0][1:]); P = int(key.split("_")[1][1:])
    x = I.mp_random(n, P, seed=digests["seed"])
    assert sha(x) == digests["mp_large"][key]["in_sha"]
    print(f"case {key}", flush=True)
    assert sha(run_fft(n, P, x)) == digests["mp_large"][key]["out_sha"]


WIDE_KEYS = ["c3_N16_P2048", "c2_N16_P2048", "c3_N1024_P2048", "c2_N1024_P2048",
             "c3_N4096_P1500", "c2_N4096_P1500", "c3_N65536_P2048"]


@pytest.mark.parametrize("key", WIDE_KEYS)
def test_fft_wide_matches_reference_digest(key, digests):
    # above 1024 bits (64 limbs, 28-bit digit products): CODE 3 and CODE 2
    # against the reference's own outputs
    code, nk, pk = key.split("_")
    n, P = int(nk[1:]), int(pk[1:])
    form = M.FORM_MP if code == "c3" else M.FORM_MP_C2
    x = I.mp_random(n, P, seed=digests["seed"])
    assert sha(x) == digests["mp_wide"][key]["in_sha"]
    print(f"case {key}", flush=True)
    assert sha(run_fft(n, P, x, form=form)) == digests["mp_wide"][key]["out_sha"]


@pytest.mark.parametrize("n,P,seed", [(1, 2048, 1), (2, 1025, 2), (4, 2000, 3), (8, 2048, 4),
                                      (64, 1056, 5), (512, 2047, 6)])
def test_fft_wide_matches_oracle(n, P, seed):
    x = I.mp_random(n, P, seed=seed)
    print(f"case n={n} P={P} seed={seed}", flush=True)
    assert np.array_equal(run_fft(n, P, x), O.mp_fft(n, P, x))


def wide_special_inputs(n, P, seed):
    """impulse, constant 0.75 - 0.75i, and a mix of signed zeros with a few
    random entries, as SoA transforms at ceil(P/32) limbs"""
    L, M2 = O.limbs(P), 2 * n
    hdr = np.full(M2, -2**31, dtype=np.int32)
    lim = np.zeros((L, M2), dtype=np.uint32)
    hdr[0], lim[0, 0] = 2, 0x80000000
    out = [I.join(hdr, lim)]
    hdr = np.zeros(M2, dtype=np.int32)
    hdr[n:] = 1
    lim = np.zeros((L, M2), dtype=np.uint32)
    lim[0, :] = 0xC0000000
    out.append(I.join(hdr, lim))
    rng = np.random.default_rng(seed)
    hx, lx = I.split(I.mp_random(n, P, seed=seed), n, L)
    hx, lx = hx.copy(), lx.copy()
    z = rng.random(M2) < 0.8
    hx[z] = np.int32(-2**31) | rng.integers(0, 2, size=int(z.sum()), dtype=np.int32)
    lx[:, z] = 0
    out.append(I.join(hx, lx))
    return out


@pytest.mark.parametrize("n", [4, 16, 256])
@pytest.mark.parametrize("P", [1500, 2048])
def test_fft_wide_special_inputs(n, P):
    # 64 limbs: impulse, constant and signed-zero inputs bit-exact to the
    # oracle (its special-input behaviour is pinned to the reference's
    # fixtures up to 1024 bits and follows MPFR's signed-zero rules beyond)
    for k, x in enumerate(wide_special_inputs(n, P, seed=n + P)):
        print(f"case n={n} P={P} input {k}", flush=True)
        assert np.array_equal(run_fft(n, P, x), O.mp_fft(n, P, x)), k


def test_batch_equals_single_transforms():
    n, P, b = 256, 256, 3
    xs = [I.mp_random(n, P, seed=40 + i) for i in range(b)]
    yb = run_fft(n, P, np.concatenate(xs), batch=b)
    for i in range(b):
        assert np.array_equal(np.split(yb, b)[i], O.mp_fft(n, P, xs[i]))


def test_device_pointer_execution():
    torch = pytest.importorskip("torch")
    n, P, b = 1024, 256, 2
    plan = M.Plan(n, P, M.FORM_MP, b)
    xs = np.concatenate([I.mp_random(n, P, seed=60 + i) for i in range(b)])
    dx = torch.from_numpy(xs.view(np.int32)).cuda()
    dy = torch.empty_like(dx)
    plan.execute_device(dx.data_ptr(), dy.data_ptr())
    plan.sync()
    y = dy.cpu().numpy().view(np.uint32)
    for i in range(b):
        assert np.array_equal(np.split(y, b)[i], O.mp_fft(n, P, np.split(xs, b)[i]))


# ---- the reference entry point on mpfr_t arrays -----------------------------

def mpfr_pair_from_soa(n, P, x):
    return MA.pair_from_limbs(n, P, x, O.limbs(P))


def soa_from_mpfr_pair(n, P, re, im):
    return MA.limbs_from_pair(n, O.limbs(P), re, im)


@pytest.mark.parametrize("entry", ["v1", "v2"])
@pytest.mark.parametrize("n,P", [(1024, 1024), (64, 256), (16, 53)])
def test_mpfr_entry_point_matches_reference(entry, n, P, mp_golden, digests):
    x = I.mp_random(n, P, seed=digests["seed"])
    inr, ini = mpfr_pair_from_soa(n, P, x)
    outr, outi = MA.MpfrArray(n, P), MA.MpfrArray(n, P)
    lib = M.lib()
    if entry == "v1":
        lib.mp_sr24_cp_fft(n, 1, 0, inr.ptr(), ini.ptr(), outr.ptr(), outi.ptr())
    else:
        lib.mp_sr24_cp_fft_v2(n, 0, 1, 0, inr.ptr(), ini.ptr(), outr.ptr(), outi.ptr())
    assert lib.mpfft_last_error() == 0
    y = soa_from_mpfr_pair(n, P, outr, outi)
    assert np.array_equal(y, mp_golden[f"out_N{n}_P{P}"])
    # the input arrays are left untouched (out-of-place)
    assert np.array_equal(soa_from_mpfr_pair(n, P, inr, ini), x)


def test_mpfr_entry_point_large_chunked(digests):
    # N = 2^16 at 1024 bits: the marshalling runs in chunks on several host
    # threads, overlapped with the transfers; the result is the reference's
    n, P = 1 << 16, 1024
    x = I.mp_random(n, P, seed=digests["seed"])
    assert sha(x) == digests["mp_large"][f"N{n}_P{P}"]["in_sha"]
    inr, ini = mpfr_pair_from_soa(n, P, x)
    outr, outi = MA.MpfrArray(n, P), MA.MpfrArray(n, P)
    lib = M.lib()
    lib.mp_sr24_cp_fft(n, 1, 0, inr.ptr(), ini.ptr(), outr.ptr(), outi.ptr())
    assert lib.mpfft_last_error() == 0
    y = soa_from_mpfr_pair(n, P, outr, outi)
    assert sha(y) == digests["mp_large"][f"N{n}_P{P}"]["out_sha"]


@pytest.mark.parametrize("key", ["c3_N1024_P2048", "c3_N4096_P1500"])
def test_mpfr_entry_point_wide(key, digests):
    # the drop-in call above 1024 bits: 64-limb significands marshalled from
    # mpfr_t, the reference's own result
    _, nk, pk = key.split("_")
    n, P = int(nk[1:]), int(pk[1:])
    x = I.mp_random(n, P, seed=digests["seed"])
    inr, ini = mpfr_pair_from_soa(n, P, x)
    outr, outi = MA.MpfrArray(n, P), MA.MpfrArray(n, P)
    lib = M.lib()
    lib.mp_sr24_cp_fft(n, 1, 0, inr.ptr(), ini.ptr(), outr.ptr(), outi.ptr())
    assert lib.mpfft_last_error() == 0
    assert sha(soa_from_mpfr_pair(n, P, outr, outi)) == digests["mp_wide"][key]["out_sha"]


def test_mpfr_plan_timing():
    n, P = 8192, 256
    plan = M.Plan(n, P, M.FORM_MP, 1)
    x = I.mp_random(n, P, seed=5)
    inr, ini = mpfr_pair_from_soa(n, P, x)
    outr, outi = MA.MpfrArray(n, P), MA.MpfrArray(n, P)
    M.check(M.lib().mpfft_execute_mpfr(plan.h, inr.ptr(), ini.ptr(), outr.ptr(), outi.ptr()),
            "execute_mpfr")
    t = np.zeros(4)
    M.check(M.lib().mpfft_plan_last_timing(plan.h, ctypes.c_void_p(t.ctypes.data)), "timing")
    plan.close()
    assert np.all(t > 0) and t[3] >= t[0] + t[1]
    assert np.array_equal(soa_from_mpfr_pair(n, P, outr, outi), O.mp_fft(n, P, x))


@pytest.mark.parametrize("form,P", [(M.FORM_FIXED, 256), (M.FORM_DD, 106), (M.FORM_QD, 212),
                                    (M.FORM_FIXED, 2048)])
def test_mpfr_entry_point_other_forms(form, P):
    # the drop-in mpfr_t path runs every limb-layout form: each result equals
    # the same plan's device execution bit for bit (same kernels), and
    # stays within the plain form's error against the 2P-bit oracle
    n = 1024
    plan = M.Plan(n, P, form, 1)
    x = I.mp_random(n, P, seed=21)
    inr, ini = mpfr_pair_from_soa(n, P, x)
    outr, outi = MA.MpfrArray(n, P), MA.MpfrArray(n, P)
    M.check(M.lib().mpfft_execute_mpfr(plan.h, inr.ptr(), ini.ptr(), outr.ptr(), outi.ptr()),
            "execute_mpfr")
    y = soa_from_mpfr_pair(n, P, outr, outi)
    yh = I.relimb(plan.execute_host(I.relimb(x, n, O.limbs(P), plan.L)), n, plan.L, O.limbs(P))
    plan.close()
    assert np.array_equal(y, yh)
    Lo, L2 = O.limbs(P), O.limbs(2 * P)
    star = O.mp_fft(n, P, x, Pw=2 * P)
    _, nw, zb = O.err_stats(n, P, y, Lo, star, L2)
    _, nw_p, _ = O.err_stats(n, P, O.mp_fft(n, P, x), Lo, star, L2)
    assert zb == 0 and nw <= max(nw_p, 1.0), (nw, nw_p)


def test_mpfr_entry_point_errors():
    lib = M.lib()
    n, P = 16, 128
    x = I.mp_random(n, P, seed=5)
    inr, ini = mpfr_pair_from_soa(n, P, x)
    outr, outi = MA.MpfrArray(n, P), MA.MpfrArray(n, P)
    MA.MpfrArray.mpfr().mpfr_set_nan(inr.ptr(3))
    lib.mp_sr24_cp_fft(n, 1, 0, inr.ptr(), ini.ptr(), outr.ptr(), outi.ptr())
    assert lib.mpfft_last_error() == -3
    # arrays of different precision
    inr2, ini2 = mpfr_pair_from_soa(n, P, x)
    out_other = MA.MpfrArray(n, P + 64)
    lib.mp_sr24_cp_fft(n, 1, 0, inr2.ptr(), ini2.ptr(), outr.ptr(), out_other.ptr())
    assert lib.mpfft_last_error() == -2
    # and a good call right after still works
    lib.mp_sr24_cp_fft(n, 1, 0, inr2.ptr(), ini2.ptr(), outr.ptr(), outi.ptr())
    assert lib.mpfft_last_error() == 0
    assert np.array_equal(soa_from_mpfr_pair(n, P, outr, outi), O.mp_fft(n, P, x))
