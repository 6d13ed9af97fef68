"""GPU check of mpfft_roundtrip_stats_noalloc (the error statistics of the
reference's test() with the kernel's partial sums in the library's fixed
workspace instead of an allocated buffer), through the C ABI: on two calls
in a row, the second on a smaller input, its results equal those of
mpfft_roundtrip_stats bit for bit (the same kernel and the same host
combination of the same partial sums); beyond its workspace it returns
MPFFT_ERR_SCRATCH without running."""
import ctypes

import numpy as np
import pytest

import mpfft_amd as M
from mpfft_amd import inputs as I

pytestmark = pytest.mark.gpu


def _stats(plan, dx, dy, noalloc):
    out = np.full(8 * plan.batch, -1.0)
    f = M.lib().mpfft_roundtrip_stats_noalloc if noalloc else M.lib().mpfft_roundtrip_stats
    rc = f(plan.h, ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()), None,
           ctypes.c_void_p(out.ctypes.data))
    return rc, out


def test_noalloc_stats_equal_allocating_entry_point():
    import torch
    for n, P, b, seed in ((4096, 256, 8, 61), (256, 1024, 3, 62)):   # the second smaller
        print(f"case noalloc n={n} P={P} batch={b}", flush=True)
        plan = M.Plan(n, P, M.FORM_MP, b)
        x = np.concatenate([I.relimb(I.mp_random(n, P, seed=seed + i), n, I.limbs_for(P), plan.L)
                            for i in range(b)])
        dx = torch.from_numpy(x.view(np.int32)).cuda()
        dX, dy = torch.empty_like(dx), torch.empty_like(dx)
        plan.execute_device(dx.data_ptr(), dX.data_ptr())
        plan.execute_inverse(dX.data_ptr(), dy.data_ptr())
        plan.sync()
        rc0, got = _stats(plan, dx, dy, True)
        rc1, ref = _stats(plan, dx, dy, False)
        assert rc0 == 0 and rc1 == 0, (rc0, rc1)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (got, ref)
        # the round trip error sits at the precision: rRMSE ~ 2^-P, log2 given
        assert np.all(got.reshape(b, 8)[:, 4:6] < -(P - 8)), got
        plan.close()


def test_noalloc_stats_limit():
    import torch
    n, P, b = 1 << 16, 64, 103          # 512 blocks x 103 x 10 doubles > 2^19
    plan = M.Plan(n, P, M.FORM_MP, b)
    d = torch.zeros(plan.words * b, dtype=torch.int32, device="cuda")
    rc, _ = _stats(plan, d, d, True)
    assert rc == -9
    plan.close()
