"""The FP64 refinement of the significant responsibilities (option "refine").

After the softmax every responsibility of at least refine_tau is recomputed
from its likelihood summed in FP64; a pattern with more than 1,024 such
entries refines those >= 1e-6 instead, or those >= 1e-3 (at most 1,000
entries of a distribution reach that).  These tests run the 32^3 / 3,240-
rotation workload (Gaussian likelihood) at noise levels that make the
posterior of every pattern sharp, broad, or nearly flat, so that all three
levels are taken, and check:

  - responsibilities against the FP64 oracle iteration (1e-4 absolute),
    scaling and model (1e-4 relative to their maxima);
  - that no pattern refines more than 1,024 entries;
  - that two runs of the same iteration give the same bits (the per-pattern
    sums run in one fixed order, whatever order the collection's atomics
    filled the lists in), with the deterministic insertion.
"""
import numpy as np
import pytest

import emc_amd as E
from emc_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL_P, TOL_PHI, TOL_M = 1e-4, 1e-4, 1e-4
N_PATTERNS = 300


@pytest.fixture(autouse=True)
def _reset():
    E.lib().emc_hip_clear_error()
    yield
    for k, v in (("likelihood", "gaussian"), ("precision", "fp32"), ("insert", "atomic"), ("refine", "on"),
                 ("refine_tau", "1e-9")):
        E.set_option(k, v)
    E.lib().emc_hip_clear_error()


@pytest.fixture(scope="module")
def workload():
    return configs.build(0, n_patterns=N_PATTERNS)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def step(W, sigma):
    prob = E.EmcProblem(W.cfg.side, W.images, W.mask, W.q, W.w, W.model0)
    prob.iteration(sigma)
    out = {"P": prob.respons(), "scaling": prob.scaling(), "model": prob.model(),
           "refined": E.lib().emc_hip_refine_count()}
    prob.free()
    E.check()
    return out


@pytest.mark.parametrize("sigma", [0.05, 0.1, 0.4])
@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_refinement_levels_vs_fp64(workload, sigma, precision):
    W = workload
    E.set_option("likelihood", "gaussian")
    E.set_option("precision", precision)
    E.set_option("insert", "deterministic")
    g = step(W, sigma)
    ref = O.o64_iteration_large(W.model0, W.images, W.mask, np.ones(W.N), W.cfg.side, W.q, W.w, W.x, W.y, W.z,
                                sigma=sigma, kind=O.GAUSSIAN)
    P64 = ref["P"]
    at_tau = (P64 >= 1e-9).sum(axis=0)
    print(f"\n{precision} sigma {sigma}: entries >= 1e-9 per pattern min {at_tau.min()} max {at_tau.max()}; "
          f"refined {g['refined']}")
    # every refined entry belongs to a pattern that refines at most 1,024
    assert 0 <= g["refined"] <= 1024 * W.N
    dP = float(np.abs(g["P"] - P64).max())
    dphi = rel(g["scaling"], ref["scaling"])
    dm = rel(g["model"], ref["model"])
    print(f"  P {dP:.2e} scaling {dphi:.2e} model {dm:.2e}")
    assert dP <= TOL_P and dphi <= TOL_PHI and dm <= TOL_M
    # the same iteration again: the same bits
    h = step(W, sigma)
    assert h["refined"] == g["refined"]
    for k in ("P", "scaling", "model"):
        assert np.array_equal(g[k].view(np.uint32), h[k].view(np.uint32)), k


def test_levels_are_taken(workload):
    """From the random initial model, sigma 0.05 leaves every pattern at most
    79 entries >= 1e-9 (all refined at tau) and sigma 0.4 gives every
    pattern all 3,240 entries >= 1e-6 (each refines only its entries
    >= 1e-3); sigma 0.1 mixes the three levels."""
    W = workload
    E.set_option("likelihood", "gaussian")
    E.set_option("precision", "fp32")
    P = {}
    for sigma in (0.05, 0.1, 0.4):
        P[sigma] = O.o64_iteration_large(W.model0, W.images, W.mask, np.ones(W.N), W.cfg.side, W.q, W.w, W.x,
                                         W.y, W.z, sigma=sigma, kind=O.GAUSSIAN)["P"]
    c9 = {s: (p >= 1e-9).sum(axis=0) for s, p in P.items()}
    c6 = {s: (p >= 1e-6).sum(axis=0) for s, p in P.items()}
    c3 = {s: (p >= 1e-3).sum(axis=0) for s, p in P.items()}
    assert c9[0.05].max() <= 1024
    assert c6[0.4].min() > 1024
    lv = np.where(c9[0.1] <= 1024, 0, np.where(c6[0.1] <= 1024, 1, 2))
    assert set(lv.tolist()) == {0, 1, 2}
    expect = {0.05: int(c9[0.05].sum()), 0.4: int(c3[0.4].sum()),
              0.1: int(np.choose(lv, [c9[0.1], c6[0.1], c3[0.1]]).sum())}
    for sigma, n in expect.items():
        g = step(W, sigma)
        print(f"sigma {sigma}: refined {g['refined']}, FP64 posterior gives {n}")
        # entries within rounding of a threshold may fall either side
        assert abs(g["refined"] - n) <= max(3, n // 100)
