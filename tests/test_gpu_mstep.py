"""The M-step over the nonzero responsibilities (option "mstep", FP16 form)
against the dense matrix-core M-step and the FP64 oracle, through the C-ABI:
  - forced "sparse" and "dense" from the same state: same responsibilities
    (same E-step), model and weights within 1e-5 relative of each other (the
    dropped terms are exact zeros; P in FP32 instead of the hi/lo operand),
    and within 1e-4 relative of the FP64 oracle;
  - the lists hold exactly the nonzero responsibilities;
  - "auto" takes the lists when at most 1 in 200 responsibilities is
    nonzero and the dense form otherwise;
  - counts that FP16 does not hold exactly (hi + lo halves) and the Poisson
    likelihood.
"""
import numpy as np
import pytest

import emc_amd as E
from oracle import oracle as O
from tests import problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sharp():
    # truth-model start, 2,000 photons: concentrated posteriors
    return problems.make(side=32, n_patterns=300, n_rot=4, n_blobs=20, radius=6.0, photons=2000.0,
                         phi_sigma=0.2, beamstop=3.0, seed=7, model="truth")


@pytest.fixture(autouse=True)
def _reset():
    E.lib().emc_hip_clear_error()
    yield
    for k, v in (("mstep", "auto"), ("likelihood", "gaussian"), ("precision", "fp32"), ("insert", "atomic")):
        E.set_option(k, v)
    E.lib().emc_hip_clear_error()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def iterate(P, mode, likelihood="gaussian", images=None):
    E.set_option("precision", "fp16")
    E.set_option("insert", "deterministic")
    E.set_option("likelihood", likelihood)
    E.set_option("mstep", mode)
    prob = E.EmcProblem(P.side, P.images if images is None else images, P.mask, P.q, P.w, P.model)
    it = prob.iteration(0.4)
    out = {"P": prob.respons(), "scaling": prob.scaling(), "model": prob.model(), "weight": prob.weight(),
           "slices": prob.slices(), "overall": it["overall_respons"], "nnz": E.lib().emc_hip_mstep_nnz()}
    prob.free()
    E.check()
    return out


@pytest.mark.parametrize("likelihood", ["gaussian", "poisson"])
def test_sparse_matches_dense_and_fp64(sharp, likelihood):
    P = sharp
    d = iterate(P, "dense", likelihood)
    s = iterate(P, "sparse", likelihood)
    kind = O.POISSON if likelihood == "poisson" else O.GAUSSIAN
    ref = O.o64_iteration(P.model.astype(np.float64), P.images, P.mask, np.ones(P.N), P.side, P.q, P.w, P.x, P.y,
                          P.z, kind=kind)
    nz = int((s["P"] != 0).sum())
    dm, dw, dsl = rel(s["model"], d["model"]), rel(s["weight"], d["weight"]), rel(s["slices"], d["slices"])
    om, os_ = rel(s["model"], ref["model"]), rel(s["scaling"], ref["scaling"])
    print(f"{likelihood}: nonzero P {nz} of {P.N * P.R}; sparse vs dense model {dm:.2e} weight {dw:.2e} "
          f"slices {dsl:.2e}; sparse vs FP64 model {om:.2e} scaling {os_:.2e}")
    assert d["nnz"] == -1 and s["nnz"] == nz
    assert np.array_equal(s["P"], d["P"]) and np.array_equal(s["scaling"], d["scaling"])
    assert dm < 1e-5 and dw < 1e-6 and dsl < 1e-5
    assert abs(s["overall"] - d["overall"]) <= 1e-6 * abs(d["overall"])
    # Poisson at 2,000 photons: |L| reaches 10^4, so the FP32 accumulation
    # and FP32 storage of L alone would move P by ~4e-4; the significant
    # responsibilities are recomputed from FP64 likelihoods (option "refine")
    assert np.abs(s["P"] - ref["P"]).max() < 1e-4 and om < 1e-4 and os_ < 1e-4


def test_auto_picks_by_density(sharp):
    P = sharp
    a = iterate(P, "auto", "poisson")
    nz = int((a["P"] != 0).sum())
    sparse_expected = nz * 200 <= P.N * P.R
    print(f"auto: {nz} nonzero of {P.N * P.R} -> {'lists' if sparse_expected else 'dense'}")
    assert a["nnz"] == (nz if sparse_expected else -1)
    # a flat posterior (random model, Gaussian): the dense form
    flat = problems.make(side=16, n_patterns=48, n_rot=2, seed=1, model="random")
    b = iterate(flat, "auto", "gaussian")
    nzb = int((b["P"] != 0).sum())
    assert nzb * 200 > flat.N * flat.R and b["nnz"] == -1


def test_sparse_inexact_counts(sharp):
    """Non-integer counts: the packed operand is split into hi + lo halves,
    and the list M-step reads both."""
    P = sharp
    imgs = P.images * np.float32(1.37)
    d = iterate(P, "dense", images=imgs)
    s = iterate(P, "sparse", images=imgs)
    dm = rel(s["model"], d["model"])
    print(f"inexact counts: sparse vs dense model {dm:.2e}")
    assert s["nnz"] >= 0 and dm < 1e-5


def test_sparse_matches_dense_side_80():
    """80 x 80 patterns (n_mp = 4,864 packed pixels): the list M-step's zero
    fill for rotations without a nonzero responsibility runs over more than
    one row of its 2D grid; sparse and dense M-steps agree as at 32^3."""
    P = problems.make(side=80, n_patterns=64, n_rot=2, n_blobs=20, radius=14.0, photons=3000.0,
                      phi_sigma=0.2, beamstop=4.0, seed=11, model="truth")
    d = iterate(P, "dense", "poisson")
    s = iterate(P, "sparse", "poisson")
    nz = int((s["P"] != 0).sum())
    rows_without = int(((s["P"] != 0).sum(axis=1) == 0).sum())
    dm, dsl = rel(s["model"], d["model"]), rel(s["slices"], d["slices"])
    print(f"nonzero P {nz} of {P.N * P.R}, rotations without one {rows_without}; model {dm:.2e} slices {dsl:.2e}")
    assert rows_without > 0 and s["nnz"] == nz
    assert np.array_equal(s["P"], d["P"]) and np.array_equal(s["scaling"], d["scaling"])
    assert dm < 1e-5 and dsl < 1e-5
