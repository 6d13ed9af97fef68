#!/usr/bin/env python3
"""Benchmark of the multiprecision split-radix FFT on one MI355X:
mp-FFTs/s at N = 2^16, 1024 and 256 bits, for the plain form (bit-identical
to the reference's MPFR path) and the fixed form (radix-4 Stockham on
fixed-point reals), on panels (twiddle products on the int8 matrix cores,
"fixed") and on lanes ("fixed_lane"); the "chosen" form is the fastest one whose error
against a 2P-bit oracle is no worse than the plain form's, componentwise
(max ulp) and normwise.  Also: the dominant kernel's roofline, the CPU
baseline and the drop-in call from mpfr_t arrays.

    python bench.py [--steps K] [--warmup W] [--n N] [--prec P] [--batch B]
                    [--batched-count C]

Beside the headline line: the double-word FP64 form at 106 bits against
the plain and fixed forms at 128 bits ("precision_form"), N = 2^20 at 512
and 2048 bits, fixed against plain ("wide"), and the batched
workload of C transforms of uneven size and precision ("batched"), one
graph launch of the batched forms against the plain form run one transform
at a time.

A step is one pass of a form over one batch of B synthetic transforms
whose inputs are already resident in HBM.  Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X published FP32 vector peak 157.3 TFLOP/s = CUs * SIMDs * lanes * clock
# * 2 (FMA) * 2 (packed): the full-rate VALU lane-instruction rate is a
# quarter of it.  v_mad_u64_u32 (one 32x32->64 product, plus a 64-bit
# addend) issues at that full rate on gfx950 (mpfft_amd/valu_probe).
VALU_LANE_OPS = 157.3e12 / 4
MFMA_I8_PEAK_T = 5000.0          # dense int8 matrix-core peak, TOP/s

PROFILES = {
    "plain": ("mp_level_kernel_N65536_P1024.txt", "mpfft::mp_level_kernel"),
    "fixed": ("fxp_r16_kernel_N65536_P1024.txt", "mpfft::fxp_r16_kernel"),
    "fixed_lane": ("fx_r4_kernel_N65536_P1024.txt", "mpfft::fx_r4_kernel"),
}


def panel_plan(n, W):
    """the panel passes (radix, Ns) of the default plan, as fxp_plan in
    mpfft_amd/csrc/kernels.h: radix 16 up to 34 words per real (a first pass
    of radix 2, 4 or 8 when log2 n is not a multiple of 4; radix 2 then 4
    for a remainder of 3 above 18 words), radix 4 beyond"""
    lg = n.bit_length() - 1
    rb = 4 if W <= 34 else 2
    out, ns = [], 1
    rem = lg % rb
    first = ([2, 4] if rem == 3 and W > 18 else [1 << rem]) if rem else []
    for R in first + [1 << rb] * ((lg - rem) // rb):
        out.append((R, ns))
        ns *= R
    return out


def panel_mfma_ops(n, W, B):
    """int8 MFMA operations (2 per multiply-add) the panel passes issue per
    step of B transforms: every complex twiddle product is 8 tiles() MFMAs of
    32x32x32 per wave (two parts, four per live tile); radix 16: per item
    (16 positions, 32 transforms) eight waves' external products (none in a
    first pass) and four waves' internal ones; radix 4 (4 positions, 64
    transforms): the three waves with r != 0 when k != 0; radix 2: none"""
    per = 8 * mfma_tiles(W) * 2 * 32 ** 3
    groups = (B + 63) // 64
    tot = 0
    for R, ns in panel_plan(n, W):
        if R == 16:
            items = n // 16 * groups * 2
            tot += items * ((8 if ns > 1 else 0) + 4) * per
        elif R == 4:
            tot += 3 * (n // 4) * (ns - 1) // ns * groups * per
        elif R == 8:
            tot += (7 * (n // 8) * (ns - 1) // ns + 2 * (n // 8)) * groups * per
    return tot


def cmul(n):
    lg = n.bit_length() - 1
    return n * lg - 3 * n + 4 if n >= 8 else 0


def fixed_products(n):
    """real twiddle products of the fixed form per transform: a radix-4
    pass with sub-transform size Ns multiplies lanes r = 1..3 of every j
    with j mod Ns != 0 by a twiddle, three real products each"""
    lg = n.bit_length() - 1
    ns, tot = (2 if lg & 1 else 1), 0
    while ns < n:
        tot += 9 * (n // 4) * (ns - 1) // ns
        ns *= 4
    return tot


def mfma_tiles(W):
    """32x32x32 int8 tiles of one real twiddle product at W words: tile
    (p, q) of the Toeplitz product is issued when it holds a twiddle digit
    (mf_live in mpfft_amd/csrc/kernels.h)"""
    nxb, ntb, c0, ncol = 4 * W, 4 * W - 3, 4 * W - 8, 4 * W + 4
    t = 0
    for p in range((ncol + 31) // 32):
        for q in range((nxb + 31) // 32):
            mmax = min(ncol - 32 * p, 32) - 1
            kmax = min(nxb - 32 * q, 32) - 1
            lo, hi = c0 + 32 * (p - q) - kmax, c0 + 32 * (p - q) + mmax
            t += mmax >= 0 and kmax >= 0 and hi >= 0 and lo <= ntb - 1
    return t


def profiled_traffic(form, n, P, batch):
    """HBM bytes per launch of the form's dominant kernel (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration) from the committed
    rocprofv3 summary, when it covers this workload; None otherwise."""
    try:
        for line in open(os.path.join(ROOT, "profiles", PROFILES[form][0])):
            f = line.split()
            if f and f[0] == "dominant_kernel_traffic_bytes_per_launch":
                if f[2:] == [f"N={n}", f"P={P}", f"batch={batch}"]:
                    return int(f[1])
    except OSError:
        pass
    return None


def measure(M, I, torch, n, P, batch, steps, warmup, seed, form):
    """Time `steps` transforms of `batch` on device-resident data: the
    throughput from HIP events (on the launch stream) around `steps` calls of
    the whole transform, then the parts of a step from events between stage
    ranges in as many steps again: plain form: leaf pass (with the levels
    finished in LDS) | combine levels on HBM; fixed form on panels: input
    conversion | radix passes | output conversion; fixed form on lanes: the
    radix passes (the first and last also convert from and to floats); word
    forms: first pass | the others."""
    plan = M.Plan(n, P, form, batch)
    Lc = I.limbs_for(P)
    host = np.concatenate([I.relimb(I.mp_random(n, P, seed=seed + b), n, Lc, plan.L)
                           for b in range(batch)])
    x = torch.from_numpy(host.view(np.int32)).cuda()
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    S = plan.stages()
    panel = form in M.FIXED_FORMS and plan.resolved_form() == M.FORM_FIXED_PANEL
    cuts = [0, S] if form in M.FIXED_FORMS and not panel else ([0, 1, S - 1, S] if panel else [0, 1, S])
    for _ in range(warmup):
        plan.execute_range(x.data_ptr(), y.data_ptr(), 0, S, sp)
    torch.cuda.synchronize()
    # throughput: one call of the library per step, as a caller runs it
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    t_start.record(stream)
    for i in range(steps):
        plan.execute_range(x.data_ptr(), y.data_ptr(), 0, S, sp)
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    # the parts of a step, timed in as many steps again with events between
    # the stage ranges (the extra calls and events cost host time that can
    # leave a fast form's GPU waiting, so they stay out of the throughput)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in cuts] for _ in range(steps)]
    for i in range(steps):
        ev[i][0].record(stream)
        for c in range(len(cuts) - 1):
            plan.execute_range(x.data_ptr(), y.data_ptr(), cuts[c], cuts[c + 1], sp)
            ev[i][c + 1].record(stream)
    torch.cuda.synchronize()
    plan.check()
    parts = [sum(e[c].elapsed_time(e[c + 1]) for e in ev) / steps for c in range(len(cuts) - 1)]
    out0 = I.relimb(y[: plan.words].cpu().numpy().view(np.uint32), n, plan.L, Lc)
    x0 = I.relimb(host[: plan.words], n, plan.L, Lc)
    plan.close()
    return dict(total_ms=total_ms, fft_s=batch * steps / (total_ms * 1e-3), parts_ms=parts,
                S=S, out0=out0, x0=x0, panel=panel, L=plan.L)


FP64_PEAK_T = 78.6              # MI355X FP64 peak, TFLOP/s


def word_fp64_ops(n, K):
    """FP64 operations per transform of the word forms (an FMA counts 2),
    counted from dd_fft.hip: TwoSum 6, Fast2Sum 3, TwoProd 3; double-word
    sum 20, product 10; quad-word sum 71, product 99 (41 + 30 and 69 + 30
    with the renormalisation); complex product 4 products + 2 sums, complex
    sum 2 sums, DFT4 8 complex sums.  Radix-r butterflies as the kernels run
    them (twiddle products skipped at k = 0; the radix-16 DFT's 8 internal
    complex products), over the pass plan of kernels.h (mw_plan); the limb
    conversions of the first and last pass are integer work, not counted."""
    add, mul = (20, 10) if K == 2 else (71, 99)
    cadd, cmul = 2 * add, 4 * mul + 2 * add
    dft4 = 8 * cadd
    lg = n.bit_length() - 1
    plan = []
    if lg >= 4:
        plan += [2] if lg & 1 else []
        plan.append(16)
        plan += [4] if lg & 2 else []
        plan += [16] * (lg // 4 - 1)
    else:
        if lg & 1:
            plan.append(2)
        while (1 << sum(r.bit_length() - 1 for r in plan)) < n:
            plan.append(4)
    ops, Ns = 0, 1
    for R in plan:
        base = {2: 2 * cadd, 4: dft4, 16: 8 * dft4 + 8 * cmul}[R]
        tw = (R - 1) * cmul
        ops += (n // R) * (base + tw * (Ns - 1) / Ns)
        Ns *= R
    return ops


def measure_precision_form(M, I, torch, O, n, P, Pcmp, batch, steps, seed):
    """The double-word form (P <= 106 bits) beside the plain and the fixed
    form at Pcmp bits, all at N = n on `batch` device-resident transforms;
    each form's error against its own 2P-bit oracle on transform 0."""
    out = {"workload": f"N={n}, batch {batch}: double-word form at {P} bits, plain and fixed "
                       f"forms at {Pcmp} bits; quad-word form at {2 * P} bits, plain and "
                       f"fixed forms at {2 * Pcmp} bits"}
    for name, form, prec in (("double_word", M.FORM_DD, P), ("plain", M.FORM_MP, Pcmp),
                             ("fixed", M.FORM_FIXED, Pcmp), ("quad_word", M.FORM_QD, 2 * P),
                             ("plain_2x", M.FORM_MP, 2 * Pcmp),
                             ("fixed_2x", M.FORM_FIXED, 2 * Pcmp)):
        r = measure(M, I, torch, n, prec, batch, steps, 1, seed, form)
        star = O.mp_fft(n, prec, r["x0"], Pw=2 * prec)
        mu, nw, zb = O.err_stats(n, prec, r["out0"], I.limbs_for(prec), star,
                                 I.limbs_for(2 * prec))
        out[name] = {"precision_bits": prec, "fft_per_s": round(r["fft_s"], 1),
                     "error_vs_2P_oracle": {"max_ulp": mu, "normwise_2^-P": nw,
                                            "zero_mismatch": zb}}
        if form in (M.FORM_DD, M.FORM_QD):
            ops = word_fp64_ops(n, 2 if form == M.FORM_DD else 4)
            rate = ops * r["fft_s"] / 1e12
            out[name]["fp64"] = {"ops_per_fft": int(ops), "achieved": round(rate, 2),
                                 "peak": FP64_PEAK_T, "unit": "TFLOP/s",
                                 "frac": round(rate / FP64_PEAK_T, 4)}
    out["double_word_vs_plain"] = round(out["double_word"]["fft_per_s"] /
                                        out["plain"]["fft_per_s"], 2)
    out["double_word_vs_fixed"] = round(out["double_word"]["fft_per_s"] /
                                        out["fixed"]["fft_per_s"], 2)
    out["quad_word_vs_plain_2x"] = round(out["quad_word"]["fft_per_s"] /
                                         out["plain_2x"]["fft_per_s"], 2)
    out["quad_word_vs_fixed_2x"] = round(out["quad_word"]["fft_per_s"] /
                                         out["fixed_2x"]["fft_per_s"], 2)
    return out


def measure_wide(M, I, torch, O, lgn, steps, seed, err=True):
    """N = 2^lgn at 512 bits (batch 64) and at 2048 bits (batch 64 and
    batch 1): the fixed form (on panels at batch 64, on lanes at batch 1)
    against the plain form (the reference's op order, 64-limb kernels at
    2048 bits).  One generated transform is copied into every slot of the
    batch on the device (the kernels' time does not depend on the values);
    transform 0 of every run is checked against the 2P-bit oracle (max ulp,
    normwise, and for the fixed form its own bound: half an ulp plus
    2 (log2 n + 2) n units of 2^(emax - F))."""
    n = 1 << lgn
    out = {"workload": f"N=2^{lgn}: 512 bits batch 64, 2048 bits batch 64 and batch 1; "
                       "fixed (panels at batch 64, lanes at batch 1) and plain"}
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    both = (("fixed", M.FORM_FIXED), ("plain", M.FORM_MP))
    stars = {}
    t_err = 0.0
    for P, B in ((512, 64), (2048, 64), (2048, 1)):
        x1 = I.mp_random(n, P, seed=seed)
        if err and P not in stars:
            t0 = time.time()
            stars = {P: O.mp_fft(n, P, x1, Pw=2 * P)}
            t_err += time.time() - t0
        for name, form in both:
            plan = M.Plan(n, P, form, B)
            one = torch.from_numpy(I.relimb(x1, n, I.limbs_for(P), plan.L).view(np.int32)).cuda()
            x = one.repeat(B)
            del one
            y = torch.empty_like(x)
            plan.execute_device(x.data_ptr(), y.data_ptr(), sp)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(steps):
                plan.execute_device(x.data_ptr(), y.data_ptr(), sp)
            e1.record(stream)
            torch.cuda.synchronize()
            plan.check()
            ms = e0.elapsed_time(e1) / steps
            key = f"{name}_P{P}" + ("" if B == 64 else f"_batch{B}")
            out[key] = {"batch": B, "ms_per_step": round(ms, 3),
                        "fft_per_s": round(B / ms * 1e3, 2)}
            if err:
                t0 = time.time()
                y0 = I.relimb(y[: plan.words].cpu().numpy().view(np.uint32), n, plan.L,
                              I.limbs_for(P))
                mu, nw, zb = O.err_stats(n, P, y0, I.limbs_for(P), stars[P],
                                         I.limbs_for(2 * P))
                e = {"max_ulp": mu, "normwise_2^-P": nw, "zero_mismatch": zb}
                if form == M.FORM_FIXED:
                    try:
                        exc = O.fixed_check(n, P, plan.L, x1, y0, stars[P])[0]
                        e["within_fixed_bound"], e["excess_units"] = True, exc
                    except AssertionError:
                        e["within_fixed_bound"] = False
                out[key]["error_vs_2P_oracle"] = e
                t_err += time.time() - t0
            plan.close()
            del x, y
            torch.cuda.empty_cache()
    for suf in ("P512", "P2048", "P2048_batch1"):
        out[f"fixed_vs_plain_{suf}"] = round(out[f"fixed_{suf}"]["fft_per_s"] /
                                             out[f"plain_{suf}"]["fft_per_s"], 3)
    if err:
        out["error_seconds"] = round(t_err, 1)
    return out


def valu_probe():
    """measured VALU issue rates on this GPU (mpfft_amd/valu_probe, built by
    python -m mpfft_amd.build from mpfft_amd/csrc/valu_probe.hip): lane-
    operations per second of v_mad_u64_u32 and the other forms it times"""
    import subprocess
    exe = os.path.join(ROOT, "mpfft_amd", "valu_probe")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        raise RuntimeError(f"valu_probe rc={r.returncode}: {r.stderr[-300:]}")
    rates = {}
    for line in r.stdout.splitlines():
        if "Tops/s" in line:
            name, val = line.split("Tops/s")[0].rsplit(None, 1)
            rates[name.strip()] = float(val)
    return {"unit": "T lane-ops/s", "rates": rates,
            "full_rate_lane_ops_published": round(VALU_LANE_OPS / 1e12, 2)}


def device_copy_rate(torch, mib=2048, reps=5):
    """HBM rate a plain device-to-device copy reaches (read + write bytes
    over its time, HIP events): the practical streaming ceiling the pass
    kernels are compared with beside the published peak"""
    src = torch.empty(mib << 18, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    rate = 2 * (mib << 20) * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return rate


def auto_form_name(M, n, P, B):
    """what a plan of the library's chosen form (MPFFT_FORM_AUTO) runs at
    this configuration"""
    names = {M.FORM_FIXED_PANEL: "fixed (panels)", M.FORM_FIXED_LANE: "fixed_lane",
             M.FORM_DD: "double_word"}
    plan = M.Plan(n, P, M.FORM_AUTO, B)
    f = plan.resolved_form()
    plan.close()
    return names.get(f, str(f))


def measure_mpfr(M, I, n, P, seed, reps=3):
    """the drop-in path: one transform from and to mpfr_t arrays
    (marshalling, PCIe transfers and the GPU transform), per call"""
    import ctypes
    from mpfft_amd import mpfr_arrays as MA
    Lc = I.limbs_for(P)
    x = I.mp_random(n, P, seed=seed)
    inr, ini = MA.pair_from_limbs(n, P, x, Lc)
    outr, outi = MA.MpfrArray(n, P), MA.MpfrArray(n, P)
    plan = M.Plan(n, P, M.FORM_MP, 1)
    call = lambda: M.check(M.lib().mpfft_execute_mpfr(plan.h, inr.ptr(), ini.ptr(),
                                                       outr.ptr(), outi.ptr()), "execute_mpfr")
    call()
    t = np.zeros(4)
    w0 = time.perf_counter()
    acc = np.zeros(4)
    for _ in range(reps):
        call()
        M.check(M.lib().mpfft_plan_last_timing(plan.h, ctypes.c_void_p(t.ctypes.data)), "timing")
        acc += t
    wall = (time.perf_counter() - w0) / reps * 1e3
    plan.close()
    acc /= reps
    return {"N": n, "precision_bits": P, "call_ms": round(wall, 3),
            "marshal_in_ms": round(acc[0], 3), "gpu_transform_ms": round(acc[1], 3),
            "marshal_out_ms": round(acc[2], 3),
            "note": "one transform per call from mpfr_t arrays, PCIe transfers included"}


def measure_lds_subtrees(M, I, torch, n, precs, sizes, batch, steps, seed):
    """The plain form with its levels up to node size S finished in LDS
    (MPFFT_SUB_SIZE = S; S = 0: every level an HBM sweep, the default): FFT/s
    and the two stage times (leaves with the LDS levels | the HBM levels) per
    precision and S.  Results are bit-identical for every S."""
    out = {"workload": f"plain form, N={n}, batch {batch}", "unit": "FFT/s"}
    old = os.environ.get("MPFFT_SUB_SIZE")
    try:
        for P in precs:
            row = {}
            for S in sizes:
                os.environ["MPFFT_SUB_SIZE"] = str(S)
                r = measure(M, I, torch, n, P, batch, steps, 1, seed, M.FORM_MP)
                row[f"S{S}"] = {"fft_per_s": round(r["fft_s"], 1),
                                "stage_ms": [round(x, 4) for x in r["parts_ms"]],
                                "hbm_level_launches": r["S"] - 1}
            out[f"P{P}"] = row
    finally:
        if old is None:
            os.environ.pop("MPFFT_SUB_SIZE", None)
        else:
            os.environ["MPFFT_SUB_SIZE"] = old
    return out


def measure_batched(M, I, torch, seed, count, steps):
    """The batched workload: `count` transforms, log2 n uniform over 6..14, P
    over {128, 256, 512, 1024}, inputs resident in HBM.  Times one graph
    launch of the batched fixed form and of the batched plain form, and the
    plain form run one transform at a time (a batch-1 plan per (n, P), one
    launch sequence per transform, as a caller of the reference's one-call-
    per-transform interface would); checks the two plain runs agree bit for
    bit."""
    ns, Ps = I.batch_draw(count, seed=seed)
    xs = I.batch_inputs(ns, Ps, seed=seed)
    t0 = time.time()
    bf = M.Batch(ns, Ps, M.FORM_FIXED)
    create_s = time.time() - t0
    bm = M.Batch(ns, Ps, M.FORM_MP)
    assert np.array_equal(bf.offsets, bm.offsets) and bf.words == bm.words
    host = bf.pack([I.relimb(x, int(n), I.limbs_for(int(P)), int(L))
                    for x, n, P, L in zip(xs, ns, Ps, bf.limbs)])
    del xs
    dx = torch.from_numpy(host.view(np.int32)).cuda()
    del host
    dy, dy2 = torch.empty_like(dx), torch.empty_like(dx)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def timed(fn, k):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(k):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    out = {"workload": f"{count} transforms, n = 2^6..2^14 log-uniform, P in {{128, 256, 512, "
                       f"1024}} bits ({len(set(zip(ns.tolist(), Ps.tolist())))} (n, P) groups), "
                       "inputs resident in HBM",
           "words": int(bf.words)}
    t0 = time.time()
    bf.tune(dx.data_ptr(), dy.data_ptr())
    tune_s = time.time() - t0
    ms_f = timed(lambda: bf.execute_device(dx.data_ptr(), dy.data_ptr(), sp), steps)
    bf.check()
    forms = bf.forms()
    out["fixed_graph"] = {"fft_per_s": round(count / ms_f * 1e3, 1), "ms": round(ms_f, 3),
                          "streams": bf.schedule()[0],
                          "groups_on_panels": int(len(set(
                              (int(n), int(P)) for n, P, f in zip(ns, Ps, forms)
                              if f == M.FORM_FIXED_PANEL))),
                          "create_s": round(create_s, 2), "tune_s": round(tune_s, 2)}
    bf.close()
    bm.tune(dx.data_ptr(), dy.data_ptr())
    ms_m = timed(lambda: bm.execute_device(dx.data_ptr(), dy.data_ptr(), sp), steps)
    out["plain_graph"] = {"fft_per_s": round(count / ms_m * 1e3, 1), "ms": round(ms_m, 3),
                          "streams": bm.schedule()[0]}
    plans = {}
    for n, P in set(zip(ns.tolist(), Ps.tolist())):
        plans[(n, P)] = M.Plan(n, P, M.FORM_MP, 1)
    keys = [(int(n), int(P)) for n, P in zip(ns, Ps)]
    offs = [int(o) * 4 for o in bm.offsets]
    xb, yb = dx.data_ptr(), dy2.data_ptr()

    def one_by_one():
        for k, o in zip(keys, offs):
            plans[k].execute_device(xb + o, yb + o, sp)

    ms_1 = timed(one_by_one, 1)
    out["plain_per_transform"] = {"fft_per_s": round(count / ms_1 * 1e3, 1), "ms": round(ms_1, 3)}
    out["plain_graph_bit_identical_to_per_transform"] = bool(torch.equal(dy, dy2))
    out["fixed_graph_vs_plain_per_transform"] = round(ms_1 / ms_f, 2)
    out["plain_graph_vs_plain_per_transform"] = round(ms_1 / ms_m, 2)
    for p in plans.values():
        p.close()
    bm.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 16)
    ap.add_argument("--prec", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--prec2", type=int, default=256, help="second precision (0 = skip)")
    ap.add_argument("--seed", type=int, default=0x6D7046465420)
    ap.add_argument("--forms", default="plain,fixed,fixed_lane", help="forms to measure")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-err", action="store_true", help="skip the 2P-bit oracle error")
    ap.add_argument("--no-mpfr", action="store_true", help="skip the mpfr_t-array call timing")
    ap.add_argument("--dd-prec", type=int, default=106,
                    help="precision of the double-word form leg (0 = skip; the quad-word "
                         "form runs at twice it); compared with the plain and fixed forms "
                         "at 128 (256) bits")
    ap.add_argument("--dd-n2", type=int, default=1 << 20,
                    help="size of the second precision-form leg (0 = skip)")
    ap.add_argument("--dd-n2-batch", type=int, default=16,
                    help="batch of the second precision-form leg")
    ap.add_argument("--wide-lgn", type=int, default=20,
                    help="log2 N of the wide-precision leg (512 and 2048 bits; 0 = skip)")
    ap.add_argument("--batched-count", type=int, default=4096,
                    help="transforms of the batched uneven-(n, P) workload (0 = skip)")
    ap.add_argument("--sub-precs", default="128,256",
                    help="precisions of the plain form's LDS-subtree leg ('' = skip)")
    ap.add_argument("--sub-sizes", default="0,128,512",
                    help="subtree sizes finished in LDS for that leg (0 = none)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch
    import mpfft_amd as M
    from mpfft_amd import inputs as I

    n, P, B = a.n, a.prec, a.batch
    Lc = I.limbs_for(P)
    form_ids = {"plain": M.FORM_MP, "fixed": M.FORM_FIXED, "fixed_lane": M.FORM_FIXED_LANE,
                "fixed_panel": M.FORM_FIXED_PANEL, "double_word": M.FORM_DD,
                "quad_word": M.FORM_QD}
    forms = [f for f in a.forms.split(",") if f]
    if "plain" not in forms:
        forms.insert(0, "plain")
    if a.calib_copy_mib:
        src = torch.empty(a.calib_copy_mib << 18, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    res = {f: measure(M, I, torch, n, P, B, a.steps, a.warmup, a.seed, form_ids[f])
           for f in forms}
    out = {}
    for f in forms:
        r = res[f]
        out[f] = {"fft_per_s": round(r["fft_s"], 3)}
        if f == "plain":
            out[f]["kernel_ms_per_step"] = {"leaf": round(r["parts_ms"][0], 4),
                                            "levels": round(r["parts_ms"][1], 4),
                                            "level_launches": r["S"] - 1}
        elif form_ids[f] in (M.FORM_DD, M.FORM_QD):
            out[f]["kernel_ms_per_step"] = {"first_pass": round(r["parts_ms"][0], 4),
                                            "other_passes": round(r["parts_ms"][1], 4),
                                            "pass_launches": r["S"]}
        elif r["panel"]:
            out[f]["kernel_ms_per_step"] = {"in": round(r["parts_ms"][0], 4),
                                            "passes": round(r["parts_ms"][1], 4),
                                            "out": round(r["parts_ms"][2], 4),
                                            "pass_launches": r["S"] - 2}
        else:
            out[f]["kernel_ms_per_step"] = {"passes": round(r["parts_ms"][0], 4),
                                            "pass_launches": r["S"]}
    if a.prec2:
        for f in forms:
            r2 = measure(M, I, torch, n, a.prec2, B, max(2, a.steps // 2), 1, a.seed, form_ids[f])
            out[f][f"fft_per_s_P{a.prec2}"] = round(r2["fft_s"], 3)

    # accuracy of every form against the 2P-bit oracle on transform 0; the
    # chosen form is the fastest one no worse than the plain form
    from oracle import oracle as O
    chosen = "plain"
    if not a.no_err:
        t0 = time.time()
        x0 = res["plain"]["x0"]
        star = O.mp_fft(n, P, x0, Pw=2 * P)
        for f in forms:
            mu, nw, zb = O.err_stats(n, P, res[f]["out0"], Lc, star, I.limbs_for(2 * P))
            out[f]["error_vs_2P_oracle"] = {"max_ulp": mu, "normwise_2^-P": nw,
                                            "zero_mismatch": zb}
        out["plain"]["bit_identical_to_reference_order"] = bool(
            np.array_equal(res["plain"]["out0"], O.mp_fft(n, P, x0)))
        ep = out["plain"]["error_vs_2P_oracle"]
        for f in forms:
            e = out[f]["error_vs_2P_oracle"]
            if (e["max_ulp"] <= ep["max_ulp"] and e["normwise_2^-P"] <= ep["normwise_2^-P"]
                    and e["zero_mismatch"] == 0 and res[f]["fft_s"] > res[chosen]["fft_s"]):
                chosen = f
        err_s = round(time.time() - t0, 2)

    rc = res[chosen]
    fft_s = rc["fft_s"]
    ms_step = rc["total_ms"] / a.steps
    # roofline of the chosen form's dominant kernel.
    #   plain, fixed on lanes (VALU-bound): algorithmic schoolbook 32x32->64
    #     limb products at the precision's limb count L, i.e. the form's real
    #     products per transform times L^2 (plain: cmul(N), the reference's
    #     own count, in the combine levels; fixed: the twiddle products of
    #     the radix passes), over that kernel's event time;
    #   fixed on panels (fxp_r4_kernel): HBM bytes a radix pass must move,
    #     one read and one write of 2N fixed-point reals of W = L + 2 words
    #     per transform, over the average pass time; the matrix-core side is
    #     reported beside it (int8 MFMA operations the pass issues).
    if rc["panel"]:
        W = rc["L"] + 2
        S = rc["S"]
        kern_s = rc["parts_ms"][1] * 1e-3
        launches = S - 2
        plan = panel_plan(n, W)
        bytes_launch = 2 * (2 * n * W * 4) * B
        achieved = bytes_launch * launches / kern_s / 1e9
        mfma_ops = panel_mfma_ops(n, W, B)
        copy_gbs = device_copy_rate(torch)
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(chosen, n, P, B),
                "kernel": PROFILES[chosen][1] if plan[-1][0] == 16 else "mpfft::fxp_r4_kernel",
                "kernel_avg_us": round(kern_s / launches * 1e6, 2),
                "algorithmic_bytes_per_launch": bytes_launch,
                "passes": [R for R, _ in plan],
                "levels_per_sweep": round((n.bit_length() - 1) / launches, 2),
                "mfma": {"achieved": round(mfma_ops / kern_s / 1e12, 2), "peak": MFMA_I8_PEAK_T,
                         "unit": "TOP/s (int8)", "frac": round(mfma_ops / kern_s / 1e12
                                                               / MFMA_I8_PEAK_T, 4)},
                "device_copy": {"achieved": round(copy_gbs, 1), "unit": "GB/s",
                                "what": "2 GiB device-to-device copy (torch), read + write bytes "
                                        "over its time",
                                "pass_frac_of_copy": round(achieved / copy_gbs, 4)}}
    else:
        prod_per_fft = (cmul(n) if chosen == "plain" else fixed_products(n)) * Lc * Lc
        kern_s = rc["parts_ms"][1 if chosen == "plain" else 0] * 1e-3
        launches = rc["S"] - (1 if chosen == "plain" else 0)
        achieved = prod_per_fft * B / kern_s / 1e12
        peak = VALU_LANE_OPS / 1e12
        roof = {"bound": "valu", "achieved": round(achieved, 4), "peak": round(peak, 4),
                "unit": "T(32x32->64 limb products)/s", "frac": round(achieved / peak, 4),
                "traffic": profiled_traffic(chosen, n, P, B),
                "kernel": PROFILES[chosen][1],
                "kernel_avg_us": round(kern_s / launches * 1e6, 2),
                "algorithmic_products_per_fft": prod_per_fft}
    bytes_fft = (4 * n + 3 * n // 4) * (4 * Lc + 4)
    line = {
        "metric": "mp-FFTs/s at N=2^16 @256/1024 bit, chosen vs plain GPU form at matched ulp error",
        "value": round(fft_s, 3),
        "unit": "FFT/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"mp{P} (u32 limbs)",
        "data": "synthetic",
        "config": {"workload": f"multiprecision FFT N={n}, P={P} bits, batch {B} transforms "
                               f"per step, {chosen} form",
                   "N": n, "precision_bits": P, "batch": B, "form": chosen,
                   "parallelism": "single GPU"},
        "forms": out,
        "chosen": chosen,
        "chosen_vs_plain": round(fft_s / res["plain"]["fft_s"], 4),
        # the same form at the second precision of the metric (256 bits)
        **({f"chosen_vs_plain_P{a.prec2}": round(out[chosen][f"fft_per_s_P{a.prec2}"] /
                                                 out["plain"][f"fft_per_s_P{a.prec2}"], 4)}
           if a.prec2 and chosen != "plain" else {}),
        "library_auto_form": auto_form_name(M, n, P, B),
        "roofline": roof,
        "roofline_hbm": {"achieved": round(bytes_fft * fft_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(bytes_fft * fft_s / 1e9 / HBM_PEAK_GBS, 5),
                         "compulsory_bytes_per_fft": bytes_fft},
    }
    if not a.no_err:
        line["error_seconds"] = err_s
    if not a.no_mpfr:
        side_leg(line, "mpfr_path", lambda: measure_mpfr(M, I, n, P, a.seed))
    if a.dd_prec:
        side_leg(line, "precision_form",
                 lambda: measure_precision_form(M, I, torch, O, n, a.dd_prec, 128, B,
                                                max(2, a.steps // 2), a.seed))
    if a.dd_prec and a.dd_n2:
        side_leg(line, "precision_form_N2",
                 lambda: measure_precision_form(M, I, torch, O, a.dd_n2, a.dd_prec, 128,
                                                a.dd_n2_batch, 2, a.seed))
    if a.wide_lgn:
        side_leg(line, "wide", lambda: measure_wide(M, I, torch, O, a.wide_lgn, 2, a.seed,
                                                    err=not a.no_err))
    if a.sub_precs:
        side_leg(line, "plain_lds_subtrees",
                 lambda: measure_lds_subtrees(M, I, torch, n,
                                              [int(v) for v in a.sub_precs.split(",") if v],
                                              [int(v) for v in a.sub_sizes.split(",") if v],
                                              B, max(2, a.steps // 2), a.seed))
    side_leg(line, "valu_probe", valu_probe)
    if line["roofline"]["bound"] == "valu" and "rates" in line.get("valu_probe", {}):
        mad = line["valu_probe"]["rates"].get("v_mad_u64_u32 (C)")
        if mad:
            line["roofline"]["measured_mad_u64_rate"] = mad
            line["roofline"]["frac_of_measured_mad_u64"] = round(
                line["roofline"]["achieved"] / mad, 4)
    if a.batched_count:
        side_leg(line, "batched", lambda: measure_batched(M, I, torch, a.seed, a.batched_count,
                                                          max(2, a.steps // 2)))
    if not a.no_cpu:
        side_leg(line, None, lambda: cpu_legs(O, I, n, P, a.seed))
    print(json.dumps(line))


def side_leg(line, key, fn):
    """Run one side leg of the bench.  A Python-level failure (an allocation
    the device cannot hold, a library status) is recorded in the line under
    the leg's key instead of losing the headline measurement; key None: fn
    returns several keys."""
    try:
        r = fn()
        if key is None:
            line.update(r)
        else:
            line[key] = r
    except Exception as e:                      # noqa: BLE001
        msg = f"{type(e).__name__}: {e}"
        line[key or "cpu_baseline"] = {"error": msg}
        print(f"bench: {key or 'cpu'} leg failed: {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """the host CPU: model name, logical CPUs, CPUs usable by this process,
    and the OpenMP thread count of the port leg (OMP_NUM_THREADS if set,
    else the usable CPUs, at most 16: one GPU's share of the host's cores on
    a machine of several GPUs)"""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else \
        (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    omp = int(env) if env.isdigit() and int(env) > 0 else min(usable, 16)
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable,
            "omp_threads": omp}


def cpu_legs(O, I, n, P, seed):
    """bounded CPU baselines; returns the line's cpu_* keys"""
    line = {}
    # bounded sample of 3 repeats (the reference's M >= 3): the reference's
    # OpenMP CODE 2 (4 threads, hard-coded at parallel_mp_fft2.c:262) when
    # its build is present, and the oracle restatement on the host's CPU
    # share and on 4 threads.  The host: CPU model, logical CPUs, and the
    # CPUs this process may run on (a machine of several GPUs may give one
    # GPU's work a share of its cores; OMP_NUM_THREADS states it)
    host = host_cpu()
    nthr = host["omp_threads"]
    xs = I.mp_random(n, P, seed=seed)
    cpu = {}
    if O.have_ref():
        t = O.ref_mp_time("mpp2", n, P, xs, 3)
        cpu = {"value": round(1.0 / t, 4), "unit": "FFT/s", "cores": 4, "kind": "reference",
               "sample": f"3 FFTs N={n} P={P}, reference OpenMP CODE 2 (4 threads, as the "
                         "reference hard-codes), LUT and conversion excluded",
               "host": host}
    t = O.mp_fft_time(n, P, xs, 3, nthr)
    port = {"value": round(1.0 / t, 4), "unit": "FFT/s", "cores": nthr, "kind": "port",
            "sample": f"3 FFTs N={n} P={P}, oracle restatement, {nthr} OpenMP threads, "
                      "LUT and conversion excluded", "host": host}
    t4 = O.mp_fft_time(n, P, xs, 3, 4)
    port4 = dict(port, value=round(1.0 / t4, 4), cores=4,
                 sample=f"3 FFTs N={n} P={P}, oracle restatement, 4 OpenMP threads, "
                        "LUT and conversion excluded")
    if cpu:
        line["cpu_baseline"] = cpu
        line["cpu_port"] = port
    else:
        line["cpu_baseline"] = port
    line["cpu_port_4_threads"] = port4
    return line


if __name__ == "__main__":
    main()
