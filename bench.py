"""EMC iteration throughput on one MI355X: pattern x rotation pairs per second.

Workload (default): the 128^3 model / 128x128 patterns / 50,000 patterns /
50,100 rotations configuration (n = 10 rotation table, antipodal half), with
synthetic data.  One step = one EMC iteration exactly as the reference driver
sequences it (emc.c:989-1084), through the emc.h C-ABI: model max, expand,
likelihood, responsibilities, sum P log P, model reset, scaling update, M-step
+ insertion, normalisation.  Inputs are resident in HBM before timing starts.

Printed: one JSON line.  `value` is the fast form's (FP16 matrix cores with
hi/lo splitting) pairs/s; the plain FP32 form (the reference's method on FP32
matrix cores) is timed in the same run and reported beside it, with the
correlation of the two final models.  `roofline` is the likelihood GEMM,
timed with HIP events on the library's stream during the timed iterations;
`cpu_baseline` is the oracle's FP32 restatement of the reference on the host
cores, on a pattern sample, extrapolated to the full pattern count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP16 = 2.5e15     # MI355X dense FP16 MFMA, FLOP/s
PEAK_FP32 = 157.3e12   # MI355X FP32 matrix, FLOP/s
PEAK_HBM = 8.0e12      # bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def phase_times(L):
    import ctypes as C
    ms = (C.c_double * 10)()
    n = (C.c_longlong * 10)()
    k = L.emc_hip_phase_times(ms, n, 10)
    return {L.emc_hip_phase_name(j).decode(): (ms[j], n[j]) for j in range(k)}


def run_form(E, W, precision, warmup, steps, sigma=0.4):
    """Returns (seconds per iteration, phase times over the timed steps,
    final model)."""
    L = E.lib()
    E.set_option("precision", precision)
    E.set_option("insert", "atomic")
    E.set_option("timing", "0")
    prob = E.EmcProblem(W.cfg.side, W.images, W.mask, W.q, W.w, W.model0)
    for it in range(warmup):
        prob.iteration(sigma)
        log(f"  {precision} warmup {it} done")
    E.set_option("timing", "1")
    L.emc_hip_reset_phase_times()
    L.emc_hip_synchronize()
    t0 = time.perf_counter()
    for it in range(steps):
        prob.iteration(sigma)
    L.emc_hip_synchronize()
    dt = (time.perf_counter() - t0) / steps
    E.check()
    ph = phase_times(L)
    E.set_option("timing", "0")
    model = prob.model()
    prob.free()
    L.emc_hip_invalidate()
    log(f"  {precision}: {dt * 1e3:.1f} ms/iteration")
    return dt, ph, model


def copy_calibration(mib):
    """A plain device-to-device copy of known size: the reference point that
    converts the FETCH_SIZE / WRITE_SIZE counters into HBM bytes."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}), flush=True)


def read_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_hbm_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cc(a, b):
    a = a.astype(np.float64) - a.mean()
    b = b.astype(np.float64) - b.mean()
    return float((a * b).sum() / np.sqrt((a * a).sum() * (b * b).sum()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--patterns", type=int, default=0, help="override the pattern count")
    ap.add_argument("--plain-steps", type=int, default=-1, help="timed FP32 iterations (default = steps)")
    ap.add_argument("--cpu-patterns", type=int, default=8, help="pattern sample for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device three times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              "(the pattern-sharded multi-GPU form is a design note only)", file=sys.stderr)
        sys.exit(2)

    import emc_amd as E
    from emc_amd import configs

    t0 = time.perf_counter()
    W = configs.build(args.config, args.patterns or None)
    N, R, n_m = W.N, W.R, W.n_m
    log(f"workload {W.cfg.name}: N={N} R={R} side={W.cfg.side} n_m={n_m} "
        f"(data {time.perf_counter() - t0:.1f}s)")

    plain_steps = args.steps if args.plain_steps < 0 else args.plain_steps
    dt_fast, ph_fast, model_fast = run_form(E, W, "fp16", args.warmup, args.steps)
    res_plain = None
    if not args.no_plain and plain_steps > 0:
        res_plain = run_form(E, W, "fp32", args.warmup, plain_steps)

    pairs = float(N) * R
    value = pairs / dt_fast
    flop_gemm = 2.0 * N * R * n_m  # algorithmic FLOP of one E-step contraction
    like_ms, like_n = ph_fast["likelihood"]
    like_avg = like_ms / 1e3 / max(like_n, 1)
    achieved = flop_gemm / like_avg

    out = {
        "metric": "pattern x rotation pairs/s per EMC iteration (FP16-MFMA hi/lo vs plain FP32)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_fast * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16 (hi/lo split) x fp32 accumulate; fp64 normalisation",
        "data": "synthetic (Gaussian-blob object, Poisson patterns), inputs resident in HBM",
        "config": {"workload": W.cfg.name, "side": W.cfg.side, "patterns": N, "rotations": R,
                   "masked_pixels": n_m, "sampling_n": W.cfg.n_rot, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": "likelihood GEMM (k_estep16)", "achieved": achieved / 1e12,
                     "peak": PEAK_FP16 / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_FP16,
                     "traffic": read_traffic("k_estep16", W.cfg.name), "launch_ms": like_avg * 1e3,
                     "algorithmic_flop_per_launch": flop_gemm},
        "phases_ms_per_step": {k: v[0] / args.steps for k, v in ph_fast.items()},
    }
    if res_plain is not None:
        dt_plain, ph_plain, model_plain = res_plain
        out["plain_fp32"] = {"value": pairs / dt_plain, "ms_per_step": dt_plain * 1e3,
                             "speedup_fast_vs_plain": dt_plain / dt_fast,
                             "phases_ms_per_step": {k: v[0] / plain_steps for k, v in ph_plain.items()},
                             "likelihood_tflops": flop_gemm / (ph_plain["likelihood"][0] / 1e3 /
                                                               max(ph_plain["likelihood"][1], 1)) / 1e12}
        if plain_steps == args.steps:
            out["model_cc_fast_vs_plain"] = cc(model_fast, model_plain)

    if not args.no_cpu:
        from oracle import oracle as O
        Ns = min(args.cpu_patterns, N)
        log(f"cpu baseline: FP32 restatement on {Ns} patterns x {R} rotations ...")
        tr = O.time_reference_iteration(W.model0, W.images[:Ns], W.mask, W.cfg.side, W.x, W.y, W.z, W.q, W.w, N)
        out["cpu_baseline"] = {
            "value": pairs / tr["full_s"], "unit": "pairs/s", "cores": tr["threads"], "kind": "port",
            "sample": (f"one iteration of the oracle's FP32 restatement on {Ns} of {N} patterns x all {R} "
                       f"rotations ({tr['sample_s']:.1f} s); per-pattern phases scaled by {N}/{Ns}, "
                       f"expand/insert/normalise counted once: {tr['full_s']:.1f} s per full iteration"),
            "phases_s": tr["phases"],
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
