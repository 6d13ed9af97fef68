"""Round-trip accuracy tool: the GPU counterpart of the reference's test()
(lab/serial/mp/serial_mp_fft3.c:346-427).

    python -m mpfft_amd.roundtrip [--n N] [--prec P] [--batch B] [--seed S]

Forward transform of seeded uniform[-1,1) input, inverse by the swap trick
(swap re/im, forward transform, divide by N, swap back), then the
reference's accuracy measures against the input: the relative root-mean-
square error per part and the largest absolute error.  Prints the
reference's summary line per transform, then one JSON line.
"""
import argparse
import json
import math

import mpfft_amd as M
from mpfft_amd import inputs as I


def fmt(log2v):
    """log2 value -> '%.4e' text, valid far below the double range"""
    if log2v == -math.inf:
        return "0.0000e+00"
    if math.isnan(log2v):
        return "nan"
    d = log2v * math.log10(2.0)
    e = math.floor(d)
    return f"{10 ** (d - e):.4f}e{e:+03d}"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--prec", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=1)
    ap.add_argument("--seed", type=int, default=I.DEFAULT_SEED)
    a = ap.parse_args(argv)
    plan = M.Plan(a.n, a.prec, M.FORM_MP, a.batch)
    import numpy as np
    x = np.concatenate([I.mp_random(a.n, a.prec, seed=a.seed + b, L=plan.L)
                        for b in range(a.batch)])
    st = plan.roundtrip_host(x)
    plan.close()
    rows = []
    for b in range(a.batch):
        s = st[b]
        print(f"Relative root-mean-square error (rRMSE): {fmt(s[4])} (real part), "
              f"{fmt(s[5])} (imaginary part)")
        rows.append({"log2_rrmse": [s[4], s[5]], "log2_max_abs_err": [s[6], s[7]]})
    print(json.dumps({"n": a.n, "prec": a.prec, "seed": a.seed, "transforms": rows}))


if __name__ == "__main__":
    main()
