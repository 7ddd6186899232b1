"""EMC iteration throughput on one MI355X: pattern x rotation pairs per second.

Workload (default): the 128^3 model / 128x128 patterns / 50,000 patterns /
50,100 rotations configuration (n = 10 rotation table, antipodal half) with
the learned rotation proposal, on synthetic data, 20 iterations from the
random initial model, with the workload's likelihood (Poisson: photon counts
contracted with the log of the slices).  One step = one EMC iteration exactly
as the reference driver sequences it (emc.c:989-1084), through the emc.h
C-ABI: model max, expand, likelihood, responsibilities, sum P log P, model
reset, scaling update, M-step + insertion, normalisation.  Inputs are
resident in HBM before timing starts.

Schedule of the learned form: iterations 1..T dense (FP16 hi/lo matrix
cores), the proposal network trained after each on that iteration's
responsibilities; prediction of the top-k coarse cells of every pattern;
then proposal iterations to the end of the run.  In each, a random 2% of the
patterns (and those in full search) are scored on the whole table first;
when more than `--max-fail` of the audited patterns hold less than 0.999 of
their posterior in their candidates, the iteration runs on the whole table,
else every other pattern is scored on the fine rotations of its candidate
cells (re-centred on its last posterior, then the network's cells).  The
first proposal iteration (untimed) audits every pattern; the rest are timed.
The plain FP32 form (the same equations on FP32 matrix cores) runs the same
number of iterations from the same start, its last `steps` timed; the final
models are correlated, and one more iteration of each from the same state
gives the accuracy of a single iteration.

Printed: one JSON line.  `value` = N * R / (seconds per timed iteration), the
effective pattern x rotation pairs per second; the pairs actually scored are
reported beside it.  `roofline` is the matrix-core kernel that took most of
the likelihood time (the candidate E-step, or the dense likelihood GEMM),
timed with HIP events on the library's stream during the timed iterations.
`cpu_baseline` is the oracle's FP32 restatement of the reference's CPU path
on the host cores, on a pattern sample, extrapolated to the full pattern
count.
"""
from __future__ import annotations

import argparse
import ctypes as C
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
NPHASE = 15
INSERT = ["atomic"]  # insertion mode of every run (--insert)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def phase_times(L):
    ms = (C.c_double * NPHASE)()
    n = (C.c_longlong * NPHASE)()
    k = L.emc_hip_phase_times(ms, n, NPHASE)
    return {L.emc_hip_phase_name(j).decode(): (ms[j], n[j]) for j in range(k)}


def proposal_stats(L):
    out = (C.c_double * 22)()
    k = L.emc_hip_proposal_stats(out, 22)
    keys = ["iteration", "full_search", "audited", "audit_below_min_mass", "audit_min_mass", "audit_mean_mass",
            "pairs_scored", "gated_at_prediction", "train_loss", "train_ms", "predict_ms", "whole_table",
            "candidate_pairs", "recentred", "whole_table_iteration", "shared_cells", "recentre_ms", "group_pairs",
            "resid_pairs", "shared_rotations", "respons_cleared", "core_cells_truncated"]
    return {keys[j]: out[j] for j in range(k)}


def timed(E, prob, sigma, steps, learned=False):
    """`steps` iterations; total and per-iteration wall time (every iteration
    ends with a host read of sum P log P, so the stream is drained), phase
    times, and for the learned form each iteration's proposal statistics with
    its own likelihood-phase time."""
    L = E.lib()
    E.set_option("timing", "1")
    L.emc_hip_reset_phase_times()
    L.emc_hip_synchronize()
    t0 = time.perf_counter()
    per, hist, its, refined = [], [], [], []
    prev = {}
    for _ in range(steps):
        t1 = time.perf_counter()
        hist.append(prob.iteration(sigma))
        its.append(time.perf_counter() - t1)
        refined.append(L.emc_hip_refine_count())
        if learned:
            st = proposal_stats(L)
            ph = phase_times(L)
            for k in ("likelihood", "groups", "shared", "audit_gemm"):
                st[k + "_ms"] = ph[k][0] - prev.get(k, 0.0)
                prev[k] = ph[k][0]
            per.append(st)
    L.emc_hip_synchronize()
    dt = (time.perf_counter() - t0) / steps
    E.check()
    ph = phase_times(L)
    E.set_option("timing", "0")
    return dt, ph, per, hist, its, float(np.mean(refined))


def run_dense(E, W, precision, warmup, steps, sigma, likelihood, pre=0):
    """`pre` + `warmup` untimed iterations, then `steps` timed ones."""
    L = E.lib()
    E.set_option("likelihood", likelihood)
    E.set_option("sparse", "on" if W.cfg.csr and precision == "fp16" else "off")
    E.set_option("precision", precision)
    E.set_option("insert", INSERT[0])
    E.set_option("proposal", "off")
    E.set_option("timing", "0")
    prob = E.EmcProblem(W.cfg.side, W.images, W.mask, W.q, W.w, W.model0)
    history = []
    t0 = time.perf_counter()
    for it in range(pre + warmup):
        history.append(prob.iteration(sigma))
        log(f"  {precision} untimed iteration {it + 1} done")
    dt, ph, _, hist, its, refined = timed(E, prob, sigma, steps)
    total_s = time.perf_counter() - t0
    history += hist
    model = prob.model()
    prob.free()
    L.emc_hip_invalidate()
    log(f"  {precision}: {dt * 1e3:.1f} ms/iteration")
    return {"dt": dt, "phases": ph, "model": model, "history": history, "iteration_s": its, "total_s": total_s,
            "refined": refined}


def run_learned(E, W, warmup, steps, sigma, likelihood, train_iters, epochs, lr, topk, n_coarse, max_fail,
                seed=0, batch=256, track=True, shared=512, one_step=True):
    L = E.lib()
    E.set_option("likelihood", likelihood)
    E.set_option("sparse", "on" if W.cfg.csr else "off")
    E.set_option("precision", "fp16")
    E.set_option("insert", INSERT[0])
    E.set_option("proposal", "off")
    E.set_option("audit_fraction", "0.02")
    E.set_option("min_mass", "0.999")
    E.set_option("audit_max_fail", str(max_fail))
    E.set_option("proposal_gate", "audit")
    E.set_option("proposal_track", "on" if track else "off")
    E.set_option("proposal_shared", str(shared))
    E.set_option("timing", "0")
    N, R, side = W.N, W.R, W.cfg.side
    prob = E.EmcProblem(side, W.images, W.mask, W.q, W.w, W.model0)
    t_start = time.perf_counter()
    if L.emc_hip_proposal_init(E.fptr(W.q), R, n_coarse, side, topk, batch, seed) != 0:
        E.check()
    dense_s = train_s = 0.0
    losses, history = [], []
    for it in range(train_iters):
        t0 = time.perf_counter()
        history.append(prob.iteration(sigma))
        L.emc_hip_synchronize()
        t1 = time.perf_counter()
        loss = C.c_double()
        ep = epochs[min(it, len(epochs) - 1)]
        if L.emc_hip_proposal_train(prob.d_images.vp, prob.d_mask.vp, side * side, prob.d_respons.vp, N, R, ep,
                                    lr, C.byref(loss)) != 0:
            E.check()
        t2 = time.perf_counter()
        dense_s += t1 - t0
        train_s += t2 - t1
        losses.append(loss.value)
        log(f"  dense iteration {it + 1}: {(t1 - t0) * 1e3:.0f} ms, training {(t2 - t1) * 1e3:.0f} ms, "
            f"loss {loss.value:.4f}")
    model_at_prediction = prob.model()  # the state the network was trained on
    t0 = time.perf_counter()
    if L.emc_hip_proposal_predict(prob.d_images.vp, prob.d_mask.vp, side * side, N) != 0:
        E.check()
    predict_s = time.perf_counter() - t0
    gated = proposal_stats(L)["gated_at_prediction"]
    pmass = np.zeros(N, np.float32)
    cand = np.zeros((N, topk), np.int32)
    L.emc_hip_proposal_get(E.iptr(cand), None, E.fptr(pmass), N)
    pq = np.percentile(pmass, [1, 10, 50]).tolist()
    per_cell = np.bincount(cand.reshape(-1), minlength=L.emc_hip_proposal_cells(None, None))
    log(f"  prediction {predict_s * 1e3:.0f} ms; {gated:.0f} patterns start in whole-table search; "
        f"network top-k mass q1/q10/q50 {pq}; entries per cell mean {per_cell.mean():.0f} "
        f"p99 {np.percentile(per_cell, 99):.0f} max {per_cell.max()}")
    E.set_option("proposal", "on")
    first_audit = None
    for it in range(warmup):
        # the first proposal iteration audits every pattern: the search mode
        # of each then follows its measured captured mass
        E.set_option("audit_fraction", "1" if it == 0 else "0.02")
        history.append(prob.iteration(sigma))
        st = proposal_stats(L)
        if it == 0:
            first_audit = st
        log(f"  proposal warmup {it + 1}: whole table {st['whole_table']:.0f}, audit below min mass "
            f"{st['audit_below_min_mass']:.0f}/{st['audited']:.0f}, min mass {st['audit_min_mass']:.6f}, "
            f"mean {st['audit_mean_mass']:.6f}, whole-table iteration {st['whole_table_iteration']:.0f}")
    E.set_option("audit_fraction", "0.02")
    dt, ph, per, hist, its, refined = timed(E, prob, sigma, steps, learned=True)
    for k, st in enumerate(per):
        log(f"  proposal iteration {train_iters + warmup + k + 1}: {its[k] * 1e3:.1f} ms, "
            f"{'whole table' if st['whole_table_iteration'] else 'candidates'}, audit below min mass "
            f"{st['audit_below_min_mass']:.0f}/{st['audited']:.0f} (mean mass {st['audit_mean_mass']:.5f}), "
            f"full search {st['full_search']:.0f}, re-centred {st['recentred']:.0f}, "
            f"shared cells {st['shared_cells']:.0f}, pairs {st['pairs_scored']:.3g}, "
            f"re-centring {st['recentre_ms']:.1f} ms host, group-block pairs {st['group_pairs']:.3g}, "
            f"per-cell pairs {st['resid_pairs']:.3g}")
    history += hist
    total_s = time.perf_counter() - t_start
    model = prob.model()
    state_model, state_scaling = model.copy(), prob.scaling().copy()
    res = {"dt": dt, "phases": ph, "model": model, "scaling": state_scaling, "per_step": per, "iteration_s": its,
           "refined": refined, "model_at_prediction": model_at_prediction,
           "dense_s": dense_s, "train_s": train_s, "predict_s": predict_s, "losses": losses, "gated": gated,
           "total_s": total_s, "predicted_mass_q1_q10_q50": pq, "history": history, "first_audit": first_audit}
    if not one_step:
        prob.free()
        E.set_option("proposal", "off")
        E.set_option("proposal_track", "on")
        E.set_option("proposal_shared", "512")
        L.emc_hip_invalidate()
        log(f"  learned: {dt * 1e3:.1f} ms/iteration")
        return res
    # one more iteration from this state, learned and plain FP32, for the
    # accuracy of a single iteration (free of the run-to-run divergence of
    # the iterated map)
    prob.iteration(sigma)
    one_st = proposal_stats(L)
    one_learned = prob.model()
    prob.free()
    E.set_option("proposal", "off")
    L.emc_hip_invalidate()
    E.set_option("precision", "fp32")
    ref = E.EmcProblem(side, W.images, W.mask, W.q, W.w, state_model, scaling=state_scaling)
    ref.iteration(sigma)
    one_plain = ref.model()
    ref.free()
    L.emc_hip_invalidate()
    log(f"  learned: {dt * 1e3:.1f} ms/iteration")
    res.update({"one_step_cc": cc(one_learned, one_plain), "one_step_whole_table": one_st["whole_table_iteration"]})
    return res


def fresh_capture(E, W, model, fresh, topk, min_mass=0.999):
    """Patterns the run never saw (same object and generator): the network
    loaded in the library predicts their top-k coarse cells; their posterior
    comes from FP16 whole-table E-steps on `model` (the model of the last
    training iteration, the frame the network learned), the scaling updated
    once between two E-steps; the captured mass is the posterior each
    pattern's predicted cells hold.  What a trained network gains where no
    previous posterior exists to track."""
    L = E.lib()
    M = fresh.shape[0]
    for k, v in (("precision", "fp16"), ("proposal", "off")):
        E.set_option(k, v)
    prob = E.EmcProblem(W.cfg.side, fresh, W.mask, W.q, W.w, model)
    if L.emc_hip_proposal_predict(prob.d_images.vp, prob.d_mask.vp, W.cfg.side ** 2, M) != 0:
        E.check()
    cand = np.zeros((M, topk), np.int32)
    L.emc_hip_proposal_get(E.iptr(cand), None, None, M)
    prob.get_slices()
    prob.likelihood(0.4)
    prob.normalize_respons()
    prob.update_scaling()
    prob.likelihood(0.4)
    prob.normalize_respons()
    P = prob.respons()
    prob.free()
    L.emc_hip_invalidate()
    nc = L.emc_hip_proposal_cells(None, None)
    off = np.zeros(nc + 1, np.int32)
    rot = np.zeros(W.R, np.int32)
    L.emc_hip_proposal_cells(E.iptr(off), E.iptr(rot))
    cell = np.empty(W.R, np.int64)
    for c in range(nc):
        cell[rot[off[c]:off[c + 1]]] = c
    cm = np.zeros((nc, M))
    np.add.at(cm, cell, P.astype(np.float64))
    mass = np.take_along_axis(cm, cand.T, axis=0).sum(axis=0)
    return {"patterns": int(M), "captured_min_mass_fraction": float(np.mean(mass >= min_mass)),
            "captured_mass_q1_q10_q50": np.percentile(mass, [1, 10, 50]).tolist(),
            "candidate_pairs_fraction": float(topk * np.mean(np.diff(off))) / W.R}


def ablation_summary(r, plain_total_s=None):
    st = r["per_step"]
    out = {"ms_per_iteration": r["dt"] * 1e3,
           "whole_table_iterations_timed": int(sum(s["whole_table_iteration"] for s in st)),
           "audit_below_min_mass_total": float(sum(s["audit_below_min_mass"] for s in st)),
           "audited_total": float(sum(s["audited"] for s in st)),
           "full_search_patterns_per_step": float(np.mean([s["full_search"] for s in st])),
           "pairs_scored_per_step": float(np.mean([s["pairs_scored"] for s in st])),
           "training_s": r["train_s"], "whole_schedule_s": r["total_s"],
           "network_topk_mass_q1_q10_q50": r["predicted_mass_q1_q10_q50"],
           "train_loss_per_dense_iteration": r["losses"]}
    if plain_total_s:
        out["whole_schedule_speedup_vs_plain_fp32"] = plain_total_s / r["total_s"]
    return out


def fp64_check(E, O, W, model, scaling, n_check):
    """One FP16 hi/lo iteration (Poisson, the whole rotation table, dense)
    of the first `n_check` patterns from the learned run's final model and
    scaling, against the FP64 oracle of the same equations on the same
    inputs: the errors of responsibilities (absolute), scaling and model
    (relative to their maxima)."""
    L = E.lib()
    Nc = min(n_check, W.N)
    for k, v in (("precision", "fp16"), ("likelihood", "poisson"), ("proposal", "off"), ("insert", "atomic"),
                 ("mstep", "auto"), ("sparse", "off")):
        E.set_option(k, v)
    imgs = np.ascontiguousarray(W.images[:Nc])
    sc = np.ascontiguousarray(scaling[:Nc], np.float32)
    prob = E.EmcProblem(W.cfg.side, imgs, W.mask, W.q, W.w, model, scaling=sc)
    prob.iteration(0.4)  # sigma is not used by the Poisson likelihood
    g = {"P": prob.respons(), "scaling": prob.scaling(), "model": prob.model()}
    prob.free()
    L.emc_hip_invalidate()
    E.check()
    log(f"fp64 check: FP64 oracle iteration on {Nc} patterns x {W.R} rotations ...")
    t0 = time.perf_counter()
    ref = O.o64_iteration_large(model, imgs, W.mask, sc, W.cfg.side, W.q, W.w, W.x, W.y, W.z, kind=O.POISSON)
    t_oracle = time.perf_counter() - t0

    def rel(a, b):
        return float(np.abs(np.asarray(a, np.float64) - b).max() / np.abs(b).max())

    return {"patterns": Nc, "rotations": W.R, "masked_pixels": W.n_m, "state": "final model and scaling of the run",
            "form": "FP16 hi/lo, Poisson, whole table",
            "responsibilities_max_abs_error": float(np.abs(g["P"] - ref["P"]).max()),
            "scaling_max_rel_error": rel(g["scaling"], ref["scaling"]),
            "model_max_rel_error": rel(g["model"], ref["model"]),
            "model_cc": cc(g["model"], ref["model"]), "oracle_s": t_oracle}


def cc(a, b):
    a = a.astype(np.float64) - a.mean()
    b = b.astype(np.float64) - b.mean()
    return float((a * b).sum() / np.sqrt((a * a).sum() * (b * b).sum()))


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed iterations (default: the rest of the workload's iteration count for the "
                         "learned form, 3 otherwise)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed iterations before them (default: 1 proposal iteration, 2 otherwise)")
    ap.add_argument("--config", type=int, default=3, help="workload preset (emc_amd/configs.py)")
    ap.add_argument("--patterns", type=int, default=0, help="override the pattern count")
    ap.add_argument("--likelihood", choices=["gaussian", "poisson"], default=None,
                    help="default: the workload's (poisson = count x log-slice contraction)")
    ap.add_argument("--sigma", type=float, default=0.4, help="noise level of the gaussian likelihood")
    ap.add_argument("--gemm-tile", choices=["256", "256w8", "128", "256x128"], default=None,
                    help="tile of the FP16 likelihood GEMM (default: the library's)")
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="library option (emc_hip_set_option) set before the run; repeatable")
    ap.add_argument("--max-fail", type=float, default=0.005,
                    help="audited fraction below min mass that sends an iteration to the whole table")
    ap.add_argument("--plain-steps", type=int, default=-1, help="timed FP32 iterations (default = steps)")
    ap.add_argument("--train-iters", type=int, default=5, help="dense iterations whose responsibilities train")
    ap.add_argument("--epochs", default="1,1,1,1,2",
                    help="training passes after each dense iteration (comma list; the last repeats)")
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--topk", type=int, default=32)
    ap.add_argument("--coarse-n", type=int, default=4)
    ap.add_argument("--train-batch", type=int, default=1024, help="patterns per network training step")
    ap.add_argument("--insert", choices=["atomic", "global_atomic", "deterministic", "deterministic_brick",
                                         "sorted_scatter"], default="atomic",
                    help="insertion mode (the library's option insert)")
    ap.add_argument("--fresh-patterns", type=int, default=1000,
                    help="patterns the run never saw, scored with the trained and the untrained network (0: off)")
    ap.add_argument("--no-ablation", action="store_true",
                    help="skip the two comparison runs (untrained network; network without tracking and shared "
                         "cells)")
    ap.add_argument("--cpu-patterns", type=int, default=8, help="pattern sample for the CPU baseline")
    ap.add_argument("--cpu-patterns-1t", type=int, default=1, help="pattern sample of the single-thread CPU baseline")
    ap.add_argument("--check-patterns", type=int, default=256,
                    help="pattern sample of the one-iteration check against the FP64 oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device three times (HBM counter calibration) and exit")
    args = ap.parse_args()
    INSERT[0] = args.insert
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              "(the pattern-sharded multi-GPU form is a design note only)", file=sys.stderr)
        sys.exit(2)

    import emc_amd as E
    from emc_amd import configs

    if args.gemm_tile:
        E.set_option("gemm_tile", args.gemm_tile)
    for kv in args.option:
        k, v = kv.split("=", 1)
        E.set_option(k, v)
    t0 = time.perf_counter()
    W = configs.build(args.config, args.patterns or None)
    N, R, n_m = W.N, W.R, W.n_m
    learned = W.cfg.proposal
    log(f"workload {W.cfg.name}: N={N} R={R} side={W.cfg.side} n_m={n_m} "
        f"(data {time.perf_counter() - t0:.1f}s)")

    like = args.likelihood or W.cfg.likelihood
    if args.warmup is None:
        args.warmup = 1 if learned else 2
    if args.steps is None:
        args.steps = max(W.cfg.iterations - args.train_iters - args.warmup, 1) if learned else 3
    plain_steps = args.steps if args.plain_steps < 0 else args.plain_steps
    pre = args.train_iters if learned else 0
    sigma = args.sigma if like == "gaussian" else None
    log(f"likelihood {like}; {pre} training + {args.warmup} untimed + {args.steps} timed iterations")
    epochs = [int(v) for v in args.epochs.split(",")]
    ablation = {}
    if learned:
        fast = run_learned(E, W, args.warmup, args.steps, args.sigma, like, args.train_iters, epochs, args.lr,
                           args.topk, args.coarse_n, args.max_fail, batch=args.train_batch)
        fresh = None
        if args.fresh_patterns > 0 and not W.cfg.csr:
            # the same object, mask and generator; with another pattern count
            # the generator draws other orientations and counts
            Wf = configs.build(args.config, args.fresh_patterns)
            assert np.array_equal(Wf.truth, W.truth)
            fresh = np.ascontiguousarray(Wf.images)
            del Wf
            ablation_fresh = {"trained": fresh_capture(E, W, fast["model_at_prediction"], fresh, args.topk)}
            log(f"fresh patterns, trained network: {ablation_fresh['trained']}")
            # the same with a freshly initialised (untrained) network
            if E.lib().emc_hip_proposal_init(E.fptr(W.q), R, args.coarse_n, W.cfg.side, args.topk,
                                             args.train_batch, 1) != 0:
                E.check()
            ablation_fresh["untrained"] = fresh_capture(E, W, fast["model_at_prediction"], fresh, args.topk)
            log(f"fresh patterns, untrained network: {ablation_fresh['untrained']}")
        if not args.no_ablation:
            # what the trained network adds: the same schedule with the
            # network left untrained, and the trained network without
            # posterior tracking and shared cells
            log("ablation: untrained network")
            ablation["untrained_network"] = run_learned(E, W, args.warmup, args.steps, args.sigma, like,
                                                        args.train_iters, [0], args.lr, args.topk, args.coarse_n,
                                                        args.max_fail, batch=args.train_batch, one_step=False)
            log("ablation: network only (no tracking, no shared cells)")
            ablation["network_only"] = run_learned(E, W, args.warmup, args.steps, args.sigma, like,
                                                   args.train_iters, epochs, args.lr, args.topk, args.coarse_n,
                                                   args.max_fail, batch=args.train_batch, track=False, shared=0,
                                                   one_step=False)
    else:
        fast = run_dense(E, W, "fp16", args.warmup, args.steps, args.sigma, like)
    plain = None
    if not args.no_plain and plain_steps > 0:
        plain = run_dense(E, W, "fp32", args.warmup, plain_steps, args.sigma, like, pre=pre)

    pairs = float(N) * R
    dt_fast = fast["dt"]
    like_ms, like_n = fast["phases"]["likelihood"]
    like_avg = like_ms / 1e3 / max(like_n, 1)
    kernel, kname = "likelihood GEMM (k_estep16)", "k_estep16"
    flop = 2.0 * N * R * n_m
    if learned:
        # the contraction kernel that took the most time over the timed
        # iterations (each in its own timing phase), with the pairs it
        # produces: group blocks (candidate pairs inside the blocks),
        # per-cell kernel (the other candidate pairs), shared rotations
        # (every pattern x the shared rotations), or the dense GEMM of the
        # iterations the audit sent to the whole table
        steps_st = fast["per_step"]
        cand_st = [st for st in steps_st if not st["whole_table_iteration"]]
        dense_st = [st for st in steps_st if st["whole_table_iteration"]]
        options = []
        if dense_st:
            options.append((sum(st["likelihood_ms"] for st in dense_st), len(dense_st), "likelihood GEMM (k_estep16)",
                            "k_estep16", 2.0 * N * R * n_m))
        if cand_st:
            mean = lambda k: float(np.mean([st[k] for st in cand_st]))
            options.append((sum(st["groups_ms"] for st in cand_st), len(cand_st),
                            "candidate E-step, group blocks (k_group_estep)", "k_group_estep",
                            2.0 * (mean("candidate_pairs") - mean("resid_pairs")) * n_m))
            options.append((sum(st["likelihood_ms"] for st in cand_st), len(cand_st),
                            "candidate E-step, per cell (k_cand_estep)", "k_cand_estep",
                            2.0 * mean("resid_pairs") * n_m))
            # the dense kernel's two launches per iteration: the audit
            # subset on the whole table and every pattern on the shared rotations
            launches = (mean("whole_table") > 0) + (mean("shared_rotations") > 0)
            options.append((sum(st["shared_ms"] + st["audit_gemm_ms"] for st in cand_st),
                            len(cand_st) * max(launches, 1),
                            "likelihood GEMM, audit subset + shared rotations (k_estep16)", "k_estep16",
                            2.0 * (mean("whole_table") * R + N * mean("shared_rotations")) * n_m / max(launches, 1)))
        tot_ms, cnt, kernel, kname, flop = max(options, key=lambda o: o[0])
        like_avg = tot_ms / 1e3 / max(cnt, 1)
    achieved = flop / like_avg

    out = {
        "metric": "pattern x rotation pairs/s per EMC iteration (FP16-MFMA + learned proposal vs plain FP32)"
        if learned else "pattern x rotation pairs/s per EMC iteration (FP16-MFMA hi/lo vs plain FP32)",
        "value": pairs / dt_fast,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_fast * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16 (hi/lo split) x fp32 accumulate; fp64 normalisation",
        "data": "synthetic (Gaussian-blob object, Poisson patterns), random initial model, inputs resident in HBM",
        "config": {"workload": W.cfg.name, "side": W.cfg.side, "patterns": N, "rotations": R,
                   "masked_pixels": n_m, "sampling_n": W.cfg.n_rot, "likelihood": like, "sigma": sigma,
                   "iterations": pre + args.warmup + args.steps, "insert": args.insert,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": kernel, "achieved": achieved / 1e12, "peak": PEAK_FP16 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP16,
                     "traffic": read_traffic(kname, W.cfg.name), "launch_ms": like_avg * 1e3,
                     "algorithmic_flop_per_launch": flop},
        "phases_ms_per_step": {k: v[0] / args.steps for k, v in fast["phases"].items()},
        "refined_responsibilities_per_step": fast["refined"],
    }
    if learned:
        steps_st = fast["per_step"]
        scored = float(np.mean([s["pairs_scored"] for s in steps_st]))
        out["proposal"] = {
            "schedule": f"{args.train_iters} dense FP16 iterations, network trained after each "
                        f"(epochs {args.epochs}, Adam lr {args.lr}); top-{args.topk} of the n={args.coarse_n} "
                        f"coarse cells; {args.warmup} untimed + {args.steps} timed proposal iterations",
            "whole_table_iterations_timed": float(sum(s["whole_table_iteration"] for s in steps_st)),
            "audit_max_fail": args.max_fail,
            "per_iteration": [{"iteration": pre + args.warmup + k + 1, "ms": fast["iteration_s"][k] * 1e3,
                               "whole_table": bool(st["whole_table_iteration"]),
                               "audit_below_min_mass": st["audit_below_min_mass"], "audited": st["audited"],
                               "audit_mean_mass": st["audit_mean_mass"], "full_search": st["full_search"],
                               "shared_cells": st["shared_cells"], "pairs_scored": st["pairs_scored"],
                               "recentre_host_ms": st["recentre_ms"]}
                              for k, st in enumerate(steps_st)],
            "pairs_scored_per_step": scored,
            "scored_pairs_per_s": scored / dt_fast,
            "whole_table_patterns_per_step": float(np.mean([s["whole_table"] for s in steps_st])),
            "full_search_patterns_per_step": float(np.mean([s["full_search"] for s in steps_st])),
            "audited_per_step": float(np.mean([s["audited"] for s in steps_st])),
            "audit_below_min_mass_per_step": float(np.mean([s["audit_below_min_mass"] for s in steps_st])),
            "audit_min_captured_mass": float(min(s["audit_min_mass"] for s in steps_st)),
            "audit_mean_captured_mass": float(np.mean([s["audit_mean_mass"] for s in steps_st])),
            "recentred_patterns_per_step": float(np.mean([s["recentred"] for s in steps_st])),
            "gated_at_prediction": fast["gated"],
            "train_loss_per_dense_iteration": fast["losses"],
            "dense_iterations_s": fast["dense_s"], "training_s": fast["train_s"], "prediction_s": fast["predict_s"],
            "whole_schedule_s": fast["total_s"],
            "network_topk_mass_q1_q10_q50": fast["predicted_mass_q1_q10_q50"],
            "full_audit_first_iteration": {k: fast["first_audit"][k] for k in
                                           ("audited", "audit_below_min_mass", "audit_min_mass", "audit_mean_mass")}
            if fast["first_audit"] else None,
            "model_cc_one_iteration_vs_plain_fp32": fast["one_step_cc"],
            "whole_schedule_speedup_vs_plain_fp32": (plain["total_s"] / fast["total_s"]) if plain else None,
            "network_on_fresh_patterns": ablation_fresh if fresh is not None else None,
            "ablation": {"full": ablation_summary(fast, plain["total_s"] if plain else None),
                         **{k: ablation_summary(v, plain["total_s"] if plain else None)
                            for k, v in ablation.items()}} if ablation else None,
            "one_iteration_ran_on_whole_table": bool(fast["one_step_whole_table"]),
            "sum_plogp_per_iteration": [h["likelihood"] for h in fast["history"]],
        }
    if plain is not None:
        out["plain_fp32"] = {"value": pairs / plain["dt"], "ms_per_step": plain["dt"] * 1e3,
                             "speedup_fast_vs_plain": plain["dt"] / dt_fast,
                             "phases_ms_per_step": {k: v[0] / plain_steps for k, v in plain["phases"].items()},
                             "sum_plogp_per_iteration": [h["likelihood"] for h in plain["history"]]}
        if plain_steps == args.steps:
            out["model_cc_fast_vs_plain"] = cc(fast["model"], plain["model"])

    if not args.no_cpu:
        from oracle import oracle as O
        Ns = min(args.cpu_patterns, N)
        log(f"cpu baseline: FP32 restatement on {Ns} patterns x {R} rotations ...")
        tr = O.time_reference_iteration(W.model0, W.images[:Ns], W.mask, W.cfg.side, W.x, W.y, W.z, W.q, W.w, N)
        N1 = min(args.cpu_patterns_1t, N)
        log(f"cpu baseline: the same on one thread, {N1} pattern(s) ...")
        t1 = O.time_reference_iteration(W.model0, W.images[:N1], W.mask, W.cfg.side, W.x, W.y, W.z, W.q, W.w, N,
                                        nthreads=1)
        out["cpu_baseline"] = {
            "value": pairs / tr["full_s"], "unit": "pairs/s", "cores": tr["threads"], "kind": "port",
            "sample": (f"one iteration of the oracle's FP32 restatement of the reference's CPU path "
                       f"(its absolute-difference likelihood, the same pattern x pixel x rotation "
                       f"contraction) on {Ns} of {N} patterns x all {R} rotations ({tr['sample_s']:.1f} s); "
                       f"per-pattern phases scaled by {N}/{Ns}, expand/insert/normalise counted once: "
                       f"{tr['full_s']:.1f} s per full iteration"),
            "phases_s": tr["phases"],
            "single_thread": {"value": pairs / t1["full_s"], "unit": "pairs/s", "cores": 1,
                              "sample": f"{N1} of {N} patterns x all {R} rotations ({t1['sample_s']:.1f} s), "
                                        f"scaled the same way: {t1['full_s']:.1f} s per full iteration"},
        }
        if learned and like == "poisson" and not W.cfg.csr:
            out["fp64_check"] = fp64_check(E, O, W, fast["model"], fast["scaling"], args.check_patterns)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
