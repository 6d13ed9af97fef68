"""Seeded synthetic inputs (the workload generator used by tests and bench).

Multiprecision inputs: real and imaginary parts i.i.d. with a random sign,
geometric exponent (e = 0 with probability 1/2, -1 with 1/4, ...) and a
P-bit significand 1bbb..b with all P-1 trailing bits random, which is
approximately uniform on (-1, 1).  Double inputs: uniform[-1, 1) binary64.

The generator is splitmix64 in counter form: draw t of stream `key` is
mix(key + (t + 1) * 0x9E3779B97F4A7C15), which is exactly the t-th output of
a splitmix64 sequence seeded with `key`; being counter-based it vectorises
in numpy.  Stream keys: mix(seed ^ (array_id * 0xD1B54A32D192ED03)).
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
DEFAULT_SEED = 0x6D7046465420
HDR_ZERO = np.int32(-2**31)

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def stream_key(seed, array_id):
    with np.errstate(over="ignore"):
        return _mix(np.uint64(seed) ^ (np.uint64(array_id) * np.uint64(0xD1B54A32D192ED03)))


def draws(key, start, count):
    """count consecutive 64-bit draws of stream key, from draw index start."""
    t = np.arange(start, start + count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(np.uint64(key) + (t + np.uint64(1)) * GAMMA)


def limbs_for(P):
    return (P + 31) // 32


def mp_random(n, P, seed=DEFAULT_SEED, L=None, array_ids=(0, 1)):
    """n complex points at P bits as one SoA-limb transform:
    returns uint32 array of 2n*(L+1) words (hdr[2n] then limbs[L][2n])."""
    L = limbs_for(P) if L is None else L
    Lp = limbs_for(P)
    per = 1 + (Lp + 1) // 2                     # draws per element
    hdr = np.empty(2 * n, dtype=np.int32)
    limb = np.zeros((L, 2 * n), dtype=np.uint32)
    for part, aid in enumerate(array_ids):
        key = stream_key(seed, aid)
        d = draws(key, 0, n * per).reshape(n, per)
        w0 = d[:, 0]
        neg = (w0 >> np.uint64(63)).astype(np.int32)
        g = (w0 & np.uint64((1 << 62) - 1)) | np.uint64(1 << 62)
        # trailing zeros of g -> geometric exponent 0, -1, -2, ...
        tz = np.zeros(n, dtype=np.int64)
        x = g.copy()
        for b in (32, 16, 8, 4, 2, 1):
            m = (x & np.uint64((1 << b) - 1)) == 0
            tz[m] += b
            x[m] >>= np.uint64(b)
        e = -tz
        words = np.empty((n, 2 * (per - 1)), dtype=np.uint32)
        words[:, 0::2] = (d[:, 1:] >> np.uint64(32)).astype(np.uint32)
        words[:, 1::2] = (d[:, 1:] & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        mant = words[:, :Lp].copy()
        mant[:, 0] |= np.uint32(0x80000000)
        W = 32 * Lp - P
        if W:
            mant[:, Lp - 1] &= np.uint32((0xFFFFFFFF << W) & 0xFFFFFFFF)
        sl = slice(part * n, (part + 1) * n)
        hdr[sl] = ((e.astype(np.int64) << 1) | neg).astype(np.int32)
        limb[:Lp, sl] = mant.T
    return np.concatenate([hdr.view(np.uint32), limb.reshape(-1)])


BATCH_PRECISIONS = (128, 256, 512, 1024)


def batch_draw(count=4096, seed=DEFAULT_SEED, lg_lo=6, lg_hi=14, precs=BATCH_PRECISIONS):
    """The batched workload: `count` transforms with log2 n uniform over
    lg_lo..lg_hi and P uniform over precs (stream 1000 of the seed).
    Returns (n, P) as int64 arrays."""
    d = draws(stream_key(seed, 1000), 0, 2 * count)
    lg = lg_lo + (d[0::2] % np.uint64(lg_hi - lg_lo + 1)).astype(np.int64)
    P = np.asarray(precs, dtype=np.int64)[(d[1::2] % np.uint64(len(precs))).astype(np.int64)]
    return np.left_shift(1, lg).astype(np.int64), P


def batch_inputs(ns, Ps, seed=DEFAULT_SEED):
    """inputs of the batched workload: transform i from arrays 2 + 2i, 3 + 2i"""
    return [mp_random(int(n), int(P), seed=seed, array_ids=(2 + 2 * i, 3 + 2 * i))
            for i, (n, P) in enumerate(zip(ns, Ps))]


def d_random(n, seed=DEFAULT_SEED, array_ids=(0, 1)):
    """n complex points uniform[-1,1) binary64: array of 2n (re then im)."""
    out = np.empty(2 * n, dtype=np.float64)
    for part, aid in enumerate(array_ids):
        d = draws(stream_key(seed, aid), 0, n)
        u = (d >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        out[part * n:(part + 1) * n] = 2.0 * u - 1.0
    return out


def split(buf, n, L):
    """view a SoA-limb transform as (hdr int32[2n], limbs uint32[L, 2n])."""
    buf = np.asarray(buf, dtype=np.uint32)
    return buf[:2 * n].view(np.int32), buf[2 * n:2 * n * (L + 1)].reshape(L, 2 * n)


def join(hdr, limbs):
    return np.concatenate([np.asarray(hdr, dtype=np.int32).view(np.uint32),
                           np.asarray(limbs, dtype=np.uint32).reshape(-1)])


def relimb(buf, n, L_from, L_to):
    """change the limb count of a SoA-limb transform (pad with zero limbs, or
    drop limbs that must be zero)."""
    hdr, limbs = split(buf, n, L_from)
    out = np.zeros((L_to, 2 * n), dtype=np.uint32)
    k = min(L_from, L_to)
    out[:k] = limbs[:k]
    if L_to < L_from and np.any(limbs[L_to:]):
        raise ValueError("dropping nonzero limbs")
    return join(hdr, out)


def to_float(buf, n, L):
    """approximate values as complex128 (for orientation checks only)."""
    hdr, limbs = split(buf, n, L)
    h = hdr.astype(np.int64)
    zero = (h & ~1) == -2**31
    e = h >> 1
    top = limbs[0].astype(np.float64) * 2.0 ** -32
    if L > 1:
        top = top + limbs[1].astype(np.float64) * 2.0 ** -64
    v = np.where(zero, 0.0, np.ldexp(top, np.where(zero, 0, e)))
    v = np.where(h & 1, -v, v)
    return v[:n] + 1j * v[n:]
