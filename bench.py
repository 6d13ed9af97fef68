#!/usr/bin/env python3
"""Benchmark of the multiprecision split-radix conjugate-pair FFT on one
MI355X: mp-FFTs/s at N = 2^16 and 1024-bit precision (plain GPU form,
bit-identical to the reference's MPFR path), with the 256-bit point, the
kernel roofline, the error against a 2P-bit oracle and the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--n N] [--prec P] [--batch B]

A step is one pass of the FFT over one batch of B synthetic transforms
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


def cmul(n):
    lg = n.bit_length() - 1
    return n * lg - 3 * n + 4 if n >= 8 else 0


def cadd(n):
    lg = n.bit_length() - 1
    return 3 * n * lg - 3 * n + 4 if n >= 8 else {1: 0, 2: 4, 4: 16}[n]


PROFILE = os.path.join(ROOT, "profiles", "mp_level_kernel_N65536_P1024.txt")


def profiled_traffic(n, P, batch):
    """HBM bytes per level-kernel launch (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration) from the committed rocprofv3 summary, when
    it covers this workload; None otherwise."""
    try:
        for line in open(PROFILE):
            f = line.split()
            if f and f[0] == "level_kernel_traffic_bytes_per_launch":
                if f[2:] == [f"N={n}", f"P={P}", f"batch={batch}"]:
                    return int(f[1])
    except OSError:
        pass
    return None


def measure(M, I, torch, n, P, batch, steps, warmup, seed):
    plan = M.Plan(n, P, M.FORM_MP, batch)
    Lc = I.limbs_for(P)
    host = np.concatenate([I.relimb(I.mp_random(n, P, seed=seed + b), n, Lc, plan.L)
                           for b in range(batch)])
    x = torch.from_numpy(host.view(np.int32)).cuda()
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    S = plan.stages()
    for _ in range(warmup):
        plan.execute_range(x.data_ptr(), y.data_ptr(), 0, S, sp)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    t_start.record(stream)
    for i in range(steps):
        ev[i][0].record(stream)
        plan.execute_range(x.data_ptr(), y.data_ptr(), 0, 1, sp)     # leaf pass
        ev[i][1].record(stream)
        plan.execute_range(x.data_ptr(), y.data_ptr(), 1, S, sp)     # combine levels
        ev[i][2].record(stream)
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    total_ms = t_start.elapsed_time(t_end)
    leaf_ms = sum(e[0].elapsed_time(e[1]) for e in ev)
    lvl_ms = sum(e[1].elapsed_time(e[2]) for e in ev)
    out0 = y[: plan.words].cpu().numpy().view(np.uint32)
    x0 = host[: plan.words]
    plan_L = plan.L
    plan.close()
    return dict(total_ms=total_ms, leaf_ms=leaf_ms, lvl_ms=lvl_ms, wall_s=wall, S=S,
                out0=I.relimb(out0, n, plan_L, Lc), x0=I.relimb(x0, n, plan_L, Lc), L=plan_L)


def measure_form(M, I, torch, n, P, batch, steps, warmup, seed, form):
    """FFT/s of a limb-layout form timed as whole transforms (device
    buffers resident), plus its first output for the error check"""
    plan = M.Plan(n, P, form, batch)
    Lc = I.limbs_for(P)
    host = np.concatenate([I.relimb(I.mp_random(n, P, seed=seed + b), n, Lc, plan.L)
                           for b in range(batch)])
    x = torch.from_numpy(host.view(np.int32)).cuda()
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    for _ in range(warmup):
        plan.execute_device(x.data_ptr(), y.data_ptr(), sp)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        plan.execute_device(x.data_ptr(), y.data_ptr(), sp)
    t1.record(stream)
    torch.cuda.synchronize()
    plan.check()
    ms = t0.elapsed_time(t1)
    out0 = I.relimb(y[: plan.words].cpu().numpy().view(np.uint32), n, plan.L, Lc)
    plan.close()
    return dict(total_ms=ms, fft_s=batch * steps / (ms * 1e-3), out0=out0)


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
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-err", action="store_true", help="skip the 2P-bit oracle error")
    ap.add_argument("--no-mpfr", action="store_true", help="skip the mpfr_t-array call timing")
    ap.add_argument("--fixed", action="store_true", help="also measure the fixed form")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch
    import mpfft_amd as M
    from mpfft_amd import inputs as I

    n, P, B = a.n, a.prec, a.batch
    if a.calib_copy_mib:
        src = torch.empty(a.calib_copy_mib << 18, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    r = measure(M, I, torch, n, P, B, a.steps, a.warmup, a.seed)
    ms_step = r["total_ms"] / a.steps
    fft_s = B * a.steps / (r["total_ms"] * 1e-3)

    # roofline of the dominant kernel (the combine-level kernel): algorithmic
    # schoolbook 32x32->64 limb products = cmul(N) * L^2 per transform
    Lc = I.limbs_for(P)
    nlev = r["S"] - 1
    prod_per_fft = cmul(n) * Lc * Lc
    lvl_s_per_step = r["lvl_ms"] * 1e-3 / a.steps
    achieved = prod_per_fft * B / lvl_s_per_step / 1e12          # T products / s
    # peak: one v_mad_u64_u32 per 32x32->64 product at the full VALU rate
    peak = VALU_LANE_OPS / 1e12
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
        "config": {"workload": f"plain multiprecision CP split-radix FFT, N={n}, P={P} bits, "
                               f"batch {B} transforms per step",
                   "N": n, "precision_bits": P, "batch": B, "form": "plain",
                   "parallelism": "single GPU"},
        "forms": {"plain": {"fft_per_s": round(fft_s, 3)}, "chosen": "plain",
                  "chosen_vs_plain": 1.0},
        "kernel_ms": {"leaf_per_step": round(r["leaf_ms"] / a.steps, 4),
                      "levels_per_step": round(r["lvl_ms"] / a.steps, 4),
                      "level_launches_per_step": nlev,
                      "mp_level_kernel_avg_us": round(r["lvl_ms"] / a.steps / nlev * 1e3, 2)},
        "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": round(peak, 4),
                     "unit": "T(32x32->64 limb products)/s", "frac": round(achieved / peak, 4),
                     "traffic": profiled_traffic(n, P, B),
                     "kernel": "mpfft::mp_level_kernel",
                     "algorithmic_per_fft": prod_per_fft},
        "roofline_hbm": {"achieved": round(bytes_fft * fft_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(bytes_fft * fft_s / 1e9 / HBM_PEAK_GBS, 5),
                         "compulsory_bytes_per_fft": bytes_fft},
    }
    if a.prec2:
        r2 = measure(M, I, torch, n, a.prec2, B, max(2, a.steps // 2), 1, a.seed)
        line["forms"][f"plain_P{a.prec2}"] = {
            "fft_per_s": round(B * max(2, a.steps // 2) / (r2["total_ms"] * 1e-3), 3)}

    if a.fixed:
        fx = measure_form(M, I, torch, n, P, B, a.steps, a.warmup, a.seed, M.FORM_FIXED)
        line["forms"]["fixed"] = {"fft_per_s": round(fx["fft_s"], 3)}
        fx_out0 = fx["out0"]
        if a.prec2:
            fx2 = measure_form(M, I, torch, n, a.prec2, B, max(2, a.steps // 2), 1, a.seed,
                               M.FORM_FIXED)
            line["forms"][f"fixed_P{a.prec2}"] = {"fft_per_s": round(fx2["fft_s"], 3)}
    if not a.no_mpfr:
        line["mpfr_path"] = measure_mpfr(M, I, n, P, a.seed)

    from oracle import oracle as O
    if not a.no_err:
        t0 = time.time()
        x0 = r["x0"]
        star = O.mp_fft(n, P, x0, Pw=2 * P)
        mu, nw, zb = O.err_stats(n, P, r["out0"], Lc, star, I.limbs_for(2 * P))
        same = bool(np.array_equal(r["out0"], O.mp_fft(n, P, x0)))
        line["error_vs_2P_oracle"] = {"max_ulp": mu, "normwise_2^-P": nw, "zero_mismatch": zb,
                                      "bit_identical_to_P_oracle": same,
                                      "seconds": round(time.time() - t0, 2)}
        if a.fixed:
            mu2, nw2, zb2 = O.err_stats(n, P, fx_out0, Lc, star, I.limbs_for(2 * P))
            line["forms"]["fixed"]["error_vs_2P_oracle"] = {"max_ulp": mu2, "normwise_2^-P": nw2,
                                                           "zero_mismatch": zb2}
    if not a.no_cpu:
        # bounded sample: the reference's OpenMP CODE 2 (4 threads, hard-coded
        # at parallel_mp_fft2.c:262) when its build is present, and the oracle
        # restatement on every host core
        nthr = min(os.cpu_count() or 1, 16)
        xs = I.mp_random(n, P, seed=a.seed)
        cpu = {}
        if O.have_ref():
            t = O.ref_mp_time("mpp2", n, P, xs, 2)
            cpu = {"value": round(1.0 / t, 4), "unit": "FFT/s", "cores": 4, "kind": "reference",
                   "sample": f"2 FFTs N={n} P={P}, reference OpenMP CODE 2 (4 threads), "
                             "LUT and conversion excluded"}
        t = O.mp_fft_time(n, P, xs, 2, nthr)
        port = {"value": round(1.0 / t, 4), "unit": "FFT/s", "cores": nthr, "kind": "port",
                "sample": f"2 FFTs N={n} P={P}, oracle restatement, {nthr} OpenMP threads, "
                          "LUT and conversion excluded"}
        if cpu:
            line["cpu_baseline"] = cpu
            line["cpu_port"] = port
        else:
            line["cpu_baseline"] = port
    print(json.dumps(line))


if __name__ == "__main__":
    main()
