"""Generate the golden fixtures from the reference itself.

Runs the reference kernels compiled from the reference sources (oracle/_ref,
built by `make -C oracle` where the reference tree exists) on the seeded
inputs of mpfft_amd.inputs and on the reference's own IN=2 input, and
writes:

  mp_golden.npz      multiprecision outputs (SoA limbs, include/mpfft.h
                     layout) of serial CODE 3 (serial_mp_fft3.c:121-253),
                     and of CODE 1 / CODE 2 (serial_mp_fft1.c, serial_mp_fft2.c)
  d_golden.npz       double outputs of CODE 5 / 6 / 7 (serial_d_fft5/6/7.c)
  digests.json       SHA-256 of large outputs, input digests, op counters

Usage: python tests/golden/gen_golden.py
       python tests/golden/gen_golden.py --wide-only   (only the digests
       above 1024 bits, added to the existing digests.json)
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mpfft_amd import inputs as I  # noqa: E402
from oracle import oracle as O      # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
SEED = I.DEFAULT_SEED

MP_SIZES = [1, 2, 4, 8, 16, 64, 1024]
MP_PRECS = [53, 64, 128, 256, 1024]
MP_LARGE = [(1 << 16, 256), (1 << 16, 1024), (1 << 20, 128)]
D_SIZES = [1, 2, 4, 8, 16, 32, 64, 1024]
D_LARGE = [1 << 16]
# above 1024 bits (64 limbs): CODE 3 and CODE 2 outputs as digests
MP_WIDE = [(16, 2048), (1024, 2048), (4096, 1500), (1 << 16, 2048)]


def wide(dig):
    dig["mp_wide"] = {}
    for n, P in MP_WIDE:
        x = I.mp_random(n, P, seed=SEED)
        for code in ((3, 2) if n <= 4096 else (3,)):
            _, y, _ = O.ref_mp(f"mp{code}", n, P, x)
            dig["mp_wide"][f"c{code}_N{n}_P{P}"] = {"in_sha": sha(x), "out_sha": sha(y)}
        print("wide", n, P, flush=True)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def special_inputs(n, P):
    """impulse, constant and a signed-zero mix as SoA transforms."""
    L = O.limbs(P)
    M = 2 * n
    out = {}
    # impulse: x[0] = 1
    hdr = np.full(M, -2**31, dtype=np.int32)
    lim = np.zeros((L, M), dtype=np.uint32)
    hdr[0] = 1 << 1                      # 0.1b * 2^1 = 1
    lim[0, 0] = 0x80000000
    out["impulse"] = I.join(hdr, lim)
    # constant x[k] = 0.75 - 0.75i
    hdr = np.zeros(M, dtype=np.int32)
    lim = np.zeros((L, M), dtype=np.uint32)
    hdr[:n] = 0 << 1
    hdr[n:] = (0 << 1) | 1
    lim[0, :] = 0xC0000000
    out["constant"] = I.join(hdr, lim)
    # signed zeros with a few nonzero entries
    rng = np.random.default_rng(11)
    x = I.mp_random(n, P, seed=SEED + 99)
    hx, lx = I.split(x, n, L)
    hx = hx.copy(); lx = lx.copy()
    z = rng.random(M) < 0.8
    hx[z] = np.int32(-2**31) | rng.integers(0, 2, size=int(z.sum()), dtype=np.int32)
    lx[:, z] = 0
    out["zeros"] = I.join(hx, lx)
    return out


def main():
    if not O.have_ref():
        sys.exit("oracle/_ref libraries missing: make -C oracle with the reference tree present")
    if "--wide-only" in sys.argv:
        path = os.path.join(OUT, "digests.json")
        with open(path) as f:
            dig = json.load(f)
        wide(dig)
        with open(path, "w") as f:
            json.dump(dig, f, indent=1, sort_keys=True)
        print("added wide digests to", path)
        return
    mp = {}
    dig = {"seed": SEED, "mp": {}, "mp_large": {}, "mp_counts": {}, "d": {}, "d_large": {}}
    for n in MP_SIZES:
        for P in MP_PRECS:
            x = I.mp_random(n, P, seed=SEED)
            _, y, cnt = O.ref_mp("mp3", n, P, x)
            mp[f"out_N{n}_P{P}"] = y
            dig["mp"][f"N{n}_P{P}"] = {"in_sha": sha(x), "out_sha": sha(y)}
            dig["mp_counts"][str(n)] = cnt
        for P in ((128, 1024) if n <= 64 else ()):
            for name, x in special_inputs(n, P).items():
                _, y, _ = O.ref_mp("mp3", n, P, x)
                mp[f"{name}_in_N{n}_P{P}"] = x
                mp[f"{name}_out_N{n}_P{P}"] = y
    # the reference's own deterministic input IN=2 (serial_mp_fft3.c:104-115)
    for n, P in [(64, 64), (64, 256), (64, 1024)]:
        x, y, _ = O.ref_mp("mp3", n, P)
        mp[f"in2_in_N{n}_P{P}"] = x
        mp[f"in2_out_N{n}_P{P}"] = y
    x, y, _ = O.ref_mp("mp3", 1024, 1024)
    mp["in2_out_N1024_P1024"] = y
    dig["in2_in_sha_N1024_P1024"] = sha(x)
    # the alternative operation orders CODE 1 and CODE 2 (serial_mp_fft1.c,
    # serial_mp_fft2.c) on the same seeded inputs as the CODE 3 fixtures
    dig["mp_codes_large"] = {}
    for code in (1, 2):
        for n in MP_SIZES:
            for P in (64, 256) + ((1024,) if n <= 64 else ()):
                x = I.mp_random(n, P, seed=SEED)
                _, y, _ = O.ref_mp(f"mp{code}", n, P, x)
                mp[f"c{code}_out_N{n}_P{P}"] = y
        n, P = 1 << 16, 256
        x = I.mp_random(n, P, seed=SEED)
        _, y, _ = O.ref_mp(f"mp{code}", n, P, x)
        dig["mp_codes_large"][f"c{code}_N{n}_P{P}"] = {"in_sha": sha(x), "out_sha": sha(y)}
    for n, P in MP_LARGE:
        x = I.mp_random(n, P, seed=SEED)
        _, y, cnt = O.ref_mp("mp3", n, P, x)
        dig["mp_large"][f"N{n}_P{P}"] = {"in_sha": sha(x), "out_sha": sha(y)}
        dig["mp_counts"][str(n)] = cnt
        print("large", n, P, flush=True)
    wide(dig)
    np.savez_compressed(os.path.join(OUT, "mp_golden.npz"), **mp)

    d = {}
    for n in D_SIZES:
        x = I.d_random(n, seed=SEED)
        for code in (5, 6, 7):
            _, y, (lr, li), cnt = O.ref_d(code, n, x)
            d[f"out{code}_N{n}"] = y
            dig["d"][f"code{code}_N{n}"] = {"in_sha": sha(x), "out_sha": sha(y), "counts": cnt}
        if n == 1024:
            d["lutr_N1024"] = lr
            d["luti_N1024"] = li
            xi, yi, _, _ = O.ref_d(6, n)
            d["in2_in_N1024"] = xi
            for code in (5, 6, 7):
                d[f"in2_out{code}_N1024"] = O.ref_d(code, n, xi)[1]
    for n in D_LARGE:
        x = I.d_random(n, seed=SEED)
        for code in (5, 6, 7):
            _, y, _, _ = O.ref_d(code, n, x)
            dig["d_large"][f"code{code}_N{n}"] = {"in_sha": sha(x), "out_sha": sha(y)}
    np.savez_compressed(os.path.join(OUT, "d_golden.npz"), **d)
    with open(os.path.join(OUT, "digests.json"), "w") as f:
        json.dump(dig, f, indent=1, sort_keys=True)
    print("wrote fixtures to", OUT)


if __name__ == "__main__":
    main()
