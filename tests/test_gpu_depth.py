"""Parity at the benchmark workloads' own accumulation depth.

The likelihood contraction of the 128^3 workloads runs over n_m = 15,576
masked pixels per (pattern, rotation) pair and that of the 256^3 sparse
workload over ~51,000; the tests elsewhere stop at 32^3 (n_m ~ 1,000).  Here a
pattern subset of the real workload (the first patterns of the same
generator, same mask, the full rotation table) runs one iteration through the
C-ABI with the workload's own options and is compared with the FP64 oracle of
the same equations (oracle.o64_iteration_large):

  - the 128^3 / 128x128 / 50,100-rotation workload, Poisson likelihood,
    FP16 hi/lo form, dense over the whole table: responsibilities 1e-4
    absolute, scaling and model 1e-4 relative, and the log-likelihood
    differences to each pattern's maximum that decide the responsibilities;
  - the same with the learned proposal's candidate restriction (top-32 of
    the 3,240 coarse cells and a 2% audit on the whole table), against the
    oracle restricted to the same rotations;
  - the 256^3 / 256x256 / 25,680-rotation sparse workload (CSR counts,
    sparse-count x log-slice E-step).

The random initial model of every run and the GPU's own model after one
iteration are both used as starting states, so the second iteration starts
from a model with the structure EMC puts in it.

At this depth the matrix cores' FP32 sums and the FP32 storage of L are off
by up to ~5e-3 in L (asserted below for the raw log-likelihoods); the
responsibilities meet 1e-4 because every entry of at least 1e-9 is
recomputed from its FP64 likelihood after the softmax (option "refine",
threshold 1e-9).
"""
import numpy as np
import pytest

import emc_amd as E
from emc_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# tolerances (the benchmark plan's bounds): responsibilities absolute,
# scaling and model relative to their maxima
TOL_P, TOL_PHI, TOL_M = 1e-4, 1e-4, 1e-4


@pytest.fixture(autouse=True)
def _reset():
    E.lib().emc_hip_clear_error()
    yield
    for k, v in (("likelihood", "gaussian"), ("proposal", "off"), ("precision", "fp32"), ("insert", "atomic"),
                 ("sparse", "off"), ("mstep", "auto"), ("audit_fraction", "0")):
        E.set_option(k, v)
    E.lib().emc_hip_clear_error()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def gpu_step(W, model, scaling, host_L=False):
    """One iteration in the driver order; optionally the raw log-likelihoods
    the reference's cuda_calculate_responsabilities copies to the host."""
    prob = E.EmcProblem(W.cfg.side, W.images, W.mask, W.q, W.w, model, scaling=scaling)
    Lh = np.zeros((W.R, W.N), np.float32) if host_L else None
    prob.model_max()
    prob.get_slices()
    prob.likelihood(0.4, Lh)
    prob.normalize_respons()
    prob.total_respons()
    prob.reset_model()
    prob.update_scaling()
    prob.update_slices()
    prob.normalize_model()
    out = {"P": prob.respons(), "scaling": prob.scaling(), "model": prob.model(), "L": Lh}
    prob.free()
    E.check()
    return out


def l_gap_error(Lg, L64, window=30.0):
    """max over (r, i) with L64[r,i] >= max_r L64[:,i] - window of
    |(Lg - max Lg) - (L64 - max L64)|: the error in what the softmax sees."""
    dg = Lg.astype(np.float64) - Lg.max(axis=0, keepdims=True)
    d64 = L64 - L64.max(axis=0, keepdims=True)
    near = d64 >= -window
    return float(np.abs(dg - d64)[near].max()), int(near.sum())


def compare(tag, g, ref):
    dP = float(np.abs(g["P"] - ref["P"]).max())
    ds = rel(g["scaling"], ref["scaling"])
    dm = rel(g["model"], ref["model"])
    print(f"{tag}: P {dP:.2e}, scaling {ds:.2e}, model {dm:.2e}; P max median "
          f"{np.median(ref['P'].max(axis=0)):.4f}, patterns with max P < 0.99: {(ref['P'].max(axis=0) < 0.99).sum()}")
    return dP, ds, dm


@pytest.fixture(scope="module")
def w128():
    return configs.build(2, n_patterns=256)


def options_128():
    E.set_option("precision", "fp16")
    E.set_option("likelihood", "poisson")
    E.set_option("insert", "atomic")
    E.set_option("mstep", "auto")
    E.set_option("proposal", "off")


def test_depth_128_poisson_fp16_two_states(w128):
    W = w128
    options_128()
    phi0 = np.ones(W.N, np.float32)
    state = (W.model0, phi0)
    for it in range(2):
        g = gpu_step(W, state[0], state[1], host_L=True)
        ref = O.o64_iteration_large(state[0], W.images, W.mask, state[1], W.cfg.side, W.q, W.w, W.x, W.y, W.z,
                                    kind=O.POISSON)
        dP, ds, dm = compare(f"128^3 Poisson FP16 iteration {it + 1} (n_m {W.n_m}, R {W.R}, N {W.N})", g, ref)
        dl, nnear = l_gap_error(g["L"], ref["L"])
        lrel = rel(g["L"], ref["L"])
        print(f"  L: max |L| {np.abs(ref['L']).max():.3e}, relative error {lrel:.2e}; error of L - max L within 30 "
              f"of the max: {dl:.2e} over {nnear} entries")
        assert dP < TOL_P and ds < TOL_PHI and dm < TOL_M
        # FP32 storage of |L| ~ 1e4..1e5 alone allows ~4e-3 in the gaps
        assert lrel < 1e-6 and dl < 2e-2
        state = (g["model"], g["scaling"])


def test_depth_128_proposal_restricted(w128):
    """The learned proposal's restriction at depth: candidates = the oracle's
    32 coarse cells of largest posterior mass per pattern (n = 4 coarse table),
    2% audit on the whole table; the GPU's candidate E-step (group blocks and
    per-cell kernel), softmax over the candidates and the M-step over the
    candidate posteriors against the FP64 oracle restricted to the same
    rotations."""
    W = w128
    options_128()
    L = E.lib()
    topk, seed = 32, 11
    assert L.emc_hip_proposal_init(E.fptr(W.q), W.R, 4, W.cfg.side, topk, 64, seed) == 0
    nc = L.emc_hip_proposal_cells(None, None)
    off = np.zeros(nc + 1, np.int32)
    rot = np.zeros(W.R, np.int32)
    assert L.emc_hip_proposal_cells(E.iptr(off), E.iptr(rot)) == nc
    phi0 = np.ones(W.N, np.float32)
    full = O.o64_iteration_large(W.model0, W.images, W.mask, phi0, W.cfg.side, W.q, W.w, W.x, W.y, W.z,
                                 kind=O.POISSON)
    cell = np.empty(W.R, np.int64)
    for c in range(nc):
        cell[rot[off[c]:off[c + 1]]] = c
    cm = np.zeros((nc, W.N))
    np.add.at(cm, cell, full["P"])
    # the 2nd..33rd cells: the best cell is left out so that the restriction
    # changes the posterior of every pattern outside the audit
    cand = np.argsort(-cm, axis=0, kind="stable")[1:1 + topk].T.astype(np.int32).copy()
    assert L.emc_hip_proposal_set_candidates(E.iptr(cand), W.N, topk) == 0
    for k, v in (("proposal", "on"), ("audit_fraction", "0.02"), ("min_mass", "0.999"), ("proposal_gate", "audit"),
                 ("proposal_track", "off"), ("audit_max_fail", "1"), ("proposal_shared", "0"),
                 ("proposal_group_min", "16")):
        E.set_option(k, v)
    g = gpu_step(W, W.model0, phi0)
    sub = np.array([i for i in range(W.N) if L.emc_uniform(seed, 0xA0D17, 1, i) < 0.02], np.int64)
    allowed = O.allowed_from_candidates(cand, off, rot, W.R, sub)
    ref = O.o64_iteration_large(W.model0, W.images, W.mask, phi0, W.cfg.side, W.q, W.w, W.x, W.y, W.z,
                                kind=O.POISSON, allowed=allowed)
    st = (E.C.c_double * 22)()
    L.emc_hip_proposal_stats(st, 22)
    print(f"candidates: {len(sub)} audited, pairs scored {st[6]:.0f} (allowed {allowed.sum()}), "
          f"group-block pairs {st[17]:.0f}")
    dP, ds, dm = compare("128^3 Poisson FP16, top-32 candidates + 2% audit", g, ref)
    assert st[2] == len(sub) and st[6] == allowed.sum()
    assert dP < TOL_P and ds < TOL_PHI and dm < TOL_M


def test_depth_256_sparse_csr():
    W = configs.build(4, n_patterns=96)
    E.set_option("precision", "fp16")
    E.set_option("likelihood", "poisson")
    E.set_option("sparse", "on")
    E.set_option("insert", "atomic")
    E.set_option("proposal", "off")
    phi0 = np.ones(W.N, np.float32)
    state = (W.model0, phi0)
    for it in range(2):
        g = gpu_step(W, state[0], state[1], host_L=True)
        ref = O.o64_iteration_large(state[0], W.images, W.mask, state[1], W.cfg.side, W.q, W.w, W.x, W.y, W.z,
                                    kind=O.POISSON)
        dP, ds, dm = compare(f"256^3 sparse CSR Poisson iteration {it + 1} (n_m {W.n_m}, R {W.R}, N {W.N})", g, ref)
        dl, nnear = l_gap_error(g["L"], ref["L"])
        print(f"  L: max |L| {np.abs(ref['L']).max():.3e}, relative error {rel(g['L'], ref['L']):.2e}; "
              f"error of L - max L within 30 of the max: {dl:.2e} over {nnear} entries")
        assert dP < TOL_P and ds < TOL_PHI and dm < TOL_M
        state = (g["model"], g["scaling"])
