"""GPU check of the chosen form (MPFFT_FORM_AUTO) through the C ABI.

Contract: a plan of MPFFT_FORM_AUTO resolves at creation to the double-word
form for 53 <= P <= 106 and to the fixed form otherwise (panel or per-lane
kernels by the fixed form's own rule), reports that through
mpfft_plan_form, and runs exactly the resolved form's kernels: its output
equals, bit for bit, that of a plan created with the resolved form.  The
batched form applies the same rule per group.  The resolved forms' own
accuracy contracts are checked in test_gpu_dd.py, test_gpu_fixed.py and
test_gpu_panel.py.
"""
import numpy as np
import pytest
import torch

import mpfft_amd as M
from mpfft_amd import inputs as I
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def run(n, P, xs, form):
    Lo = O.limbs(P)
    plan = M.Plan(n, P, form, len(xs))
    got = plan.resolved_form()
    y = plan.execute_host(np.concatenate([I.relimb(t, n, Lo, plan.L) for t in xs]))
    plan.close()
    return got, [I.relimb(t, n, plan.L, Lo) for t in np.split(y, len(xs))]


@pytest.mark.parametrize("n,P,b,expect", [
    (1024, 106, 2, M.FORM_DD), (16, 53, 1, M.FORM_DD), (4096, 100, 3, M.FORM_DD),
    (4096, 1024, 64, M.FORM_FIXED_PANEL), (4096, 1024, 1, M.FORM_FIXED_LANE),
    (256, 40, 1, M.FORM_FIXED_LANE), (1024, 107, 2, M.FORM_FIXED_LANE),
    (64, 2048, 1, M.FORM_FIXED_PANEL), (4096, 2048, 1, M.FORM_FIXED_LANE),
])
def test_auto_resolves_and_runs_the_resolved_form(n, P, b, expect):
    xs = [I.mp_random(n, P, seed=2600 + n + P + i) for i in range(b)]
    print(f"case n={n} P={P} b={b}", flush=True)
    got, ya = run(n, P, xs, M.FORM_AUTO)
    assert got == expect, (got, expect)
    got2, yr = run(n, P, xs, expect)
    assert got2 == expect
    for u, v in zip(ya, yr):
        assert np.array_equal(u, v)


def test_plan_form_of_every_form():
    cases = [(M.FORM_MP, 256), (M.FORM_MP_C1, 256), (M.FORM_MP_C2, 256), (M.FORM_QD, 200),
             (M.FORM_DD, 100), (M.FORM_FIXED_LANE, 256), (M.FORM_FIXED_PANEL, 256)]
    for form, P in cases:
        plan = M.Plan(64, P, form, 1)
        assert plan.resolved_form() == form
        plan.close()
    with pytest.raises(M.MpfftError) as e:
        M.Plan(64, 4096, M.FORM_AUTO, 1)
    assert e.value.code == -2


def test_batch_auto_groups():
    # groups at 100 bits take the double-word form, the others the fixed
    # form's kernels; each transform equals a plan of its group's form
    ns = np.array([64, 64, 1024, 16, 256, 256, 4096], dtype=np.int64)
    Ps = np.array([100, 100, 106, 256, 1024, 1024, 2048], dtype=np.int64)
    xs = [I.mp_random(int(n), int(P), seed=2700 + i) for i, (n, P) in enumerate(zip(ns, Ps))]
    b = M.Batch(ns, Ps, M.FORM_AUTO)
    forms = b.forms()
    host = b.pack([I.relimb(x, int(n), O.limbs(int(P)), int(L))
                   for x, n, P, L in zip(xs, b.ns, b.Ps, b.limbs)])
    dx = torch.from_numpy(host.view(np.int32)).cuda()
    dy = torch.zeros_like(dx)
    b.execute_device(dx.data_ptr(), dy.data_ptr())
    b.check()
    ys = b.unpack(dy.cpu().numpy().view(np.uint32))
    for i in range(len(xs)):
        n, P = int(b.ns[i]), int(b.Ps[i])
        expect = M.FORM_DD if 53 <= P <= 106 else None
        if expect is not None:
            assert forms[i] == expect
        else:
            assert forms[i] in (M.FORM_FIXED_LANE, M.FORM_FIXED_PANEL)
        cnt = int(((b.ns == n) & (b.Ps == P)).sum())
        plan = M.Plan(n, P, int(forms[i]), cnt)
        same = [j for j in range(len(xs)) if int(b.ns[j]) == n and int(b.Ps[j]) == P]
        yp = np.split(plan.execute_host(np.concatenate(
            [I.relimb(xs[j], n, O.limbs(P), plan.L) for j in same])), cnt)
        plan.close()
        assert np.array_equal(ys[i], yp[same.index(i)]), (n, P)
    b.close()


def test_batch_quad_word_groups():
    # the quad-word form as a batch form: every transform equals a plan of
    # that form over its group; a precision beyond the form is refused
    ns = np.array([256, 16, 256, 1024], dtype=np.int64)
    Ps = np.array([150, 212, 150, 60], dtype=np.int64)
    xs = [I.mp_random(int(n), int(P), seed=2800 + i) for i, (n, P) in enumerate(zip(ns, Ps))]
    b = M.Batch(ns, Ps, M.FORM_QD)
    assert (b.forms() == M.FORM_QD).all()
    host = b.pack([I.relimb(x, int(n), O.limbs(int(P)), int(L))
                   for x, n, P, L in zip(xs, b.ns, b.Ps, b.limbs)])
    dx = torch.from_numpy(host.view(np.int32)).cuda()
    dy = torch.zeros_like(dx)
    b.execute_device(dx.data_ptr(), dy.data_ptr())
    b.check()
    ys = b.unpack(dy.cpu().numpy().view(np.uint32))
    b.close()
    for i in range(len(xs)):
        n, P = int(ns[i]), int(Ps[i])
        same = [j for j in range(len(xs)) if int(ns[j]) == n and int(Ps[j]) == P]
        plan = M.Plan(n, P, M.FORM_QD, len(same))
        yp = np.split(plan.execute_host(np.concatenate(
            [I.relimb(xs[j], n, O.limbs(P), plan.L) for j in same])), len(same))
        plan.close()
        assert np.array_equal(ys[i], yp[same.index(i)]), (n, P)
    with pytest.raises(M.MpfftError) as e:
        M.Batch(np.array([64]), np.array([300]), M.FORM_QD)
    assert e.value.code == -2
