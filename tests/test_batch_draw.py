"""The batched workload's seeded draw (no GPU): n log-uniform over
2^lg_lo..2^lg_hi, P uniform over the precision set, reproducible."""
import numpy as np

from mpfft_amd import inputs as I


def test_draw_reproducible_and_in_range():
    ns, Ps = I.batch_draw(4096)
    ns2, Ps2 = I.batch_draw(4096)
    assert np.array_equal(ns, ns2) and np.array_equal(Ps, Ps2)
    lg = np.log2(ns).astype(int)
    assert (1 << lg == ns).all() and lg.min() == 6 and lg.max() == 14
    assert set(Ps.tolist()) == set(I.BATCH_PRECISIONS)
    # roughly uniform: every log2 n and every P within 20 % of its share
    assert (np.abs(np.bincount(lg - 6) - 4096 / 9) < 0.2 * 4096 / 9).all()
    for P in I.BATCH_PRECISIONS:
        assert abs((Ps == P).sum() - 1024) < 0.2 * 1024
    ns3, _ = I.batch_draw(4096, seed=1)
    assert not np.array_equal(ns, ns3)


def test_draw_inputs_distinct_per_transform():
    ns = np.array([64, 64], dtype=np.int64)
    Ps = np.array([128, 128], dtype=np.int64)
    a, b = I.batch_inputs(ns, Ps)
    assert a.size == b.size == 2 * 64 * 5 and not np.array_equal(a, b)
    a2, _ = I.batch_inputs(ns, Ps)
    assert np.array_equal(a, a2)
