"""Insertion modes through the C-ABI.

"atomic" (the default) is brick-owned: each workgroup owns an 8^3 brick of
the model (and a share of the active rotations when there are too few
bricks to fill the GPU) and finds, for every active rotation, the detector
pixels that can reach it from the brick's box mapped back onto the
detector.  It must add exactly the contributions the per-pixel kernels add
(the same voxels, the same values, in another order): against
"global_atomic" (one float atomic per pixel) and "deterministic", nearest
and trilinear, dense and sparse posteriors; a model side that is not a
multiple of 8 falls back to the per-pixel kernel.

"deterministic_brick" runs 4^3 bricks in the order of the reference's
sequential insertion (emc.c:491-539: rotations ascending, pixels x-major),
and "sorted_scatter" sums the same contributions in the same order from a
voxel-sorted map (a wave per voxel where the list is long): the two give the
same bits, on every run, and "deterministic" (either, by the number of
active rotations) gives them too."""
import numpy as np
import pytest

import emc_amd as E
from oracle import oracle as O
from tests import problems

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _reset():
    E.lib().emc_hip_clear_error()
    yield
    for k, v in (("insert", "atomic"), ("interp", "nearest"), ("precision", "fp32"), ("likelihood", "gaussian")):
        E.set_option(k, v)


def run(P, insert, interp, likelihood="gaussian", precision="fp32"):
    E.set_option("insert", insert)
    E.set_option("interp", interp)
    E.set_option("likelihood", likelihood)
    E.set_option("precision", precision)
    prob = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
    prob.model_max()
    prob.get_slices()
    prob.likelihood(0.4)
    prob.normalize_respons()
    prob.total_respons()
    prob.reset_model()
    prob.update_scaling()
    prob.update_slices()
    out = {"raw": prob.model(), "weight": prob.weight()}
    prob.normalize_model()
    out["model"] = prob.model()
    prob.free()
    E.check()
    return out


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
@pytest.mark.parametrize("likelihood", ["gaussian", "poisson"])
def test_brick_insert_matches_per_pixel_kernels(interp, likelihood):
    P = problems.make(side=32, n_patterns=200, n_rot=4, n_blobs=20, radius=6.0, photons=2000.0, beamstop=3.0,
                      gap_cross=1, seed=9, model="truth")
    b = run(P, "atomic", interp, likelihood)
    g = run(P, "global_atomic", interp, likelihood)
    d = run(P, "deterministic", interp, likelihood)
    # the same set of voxels reached, the same sums to FP32 reordering
    for k in ("raw", "weight"):
        assert np.array_equal(b[k] != 0, d[k] != 0), k
    eb, eg = rel(b["raw"], d["raw"]), rel(g["raw"], d["raw"])
    wb, wg = rel(b["weight"], d["weight"]), rel(g["weight"], d["weight"])
    print(f"{interp} {likelihood}: brick vs deterministic raw {eb:.2e} weight {wb:.2e}; "
          f"global atomics vs deterministic raw {eg:.2e} weight {wg:.2e}")
    assert eb < 1e-5 and wb < 1e-5 and eg < 1e-5 and wg < 1e-5
    assert rel(b["model"], d["model"]) < 1e-5


def test_brick_insert_larger_model_vs_oracle():
    """64^3 with the 2-px horizontal gap of the 64^3 workload: every
    contribution found (weight equal to the deterministic gather's to FP32
    reordering) and the model against the FP64 oracle."""
    P = problems.make(side=64, n_patterns=64, n_rot=4, n_blobs=30, radius=10.0, photons=4000.0, beamstop=4.0,
                      gap_h=2, seed=4, model="truth")
    b = run(P, "atomic", "nearest")
    d = run(P, "deterministic", "nearest")
    assert np.array_equal(b["weight"] != 0, d["weight"] != 0)
    assert rel(b["weight"], d["weight"]) < 1e-5 and rel(b["raw"], d["raw"]) < 1e-5
    ref = O.o64_iteration_large(P.model, P.images, P.mask, np.ones(P.N), P.side, P.q, P.w, P.x, P.y, P.z)
    assert rel(b["weight"], ref["weight"]) < 1e-5


def test_side_not_multiple_of_brick_falls_back():
    P = problems.make(side=20, n_patterns=40, n_rot=2, n_blobs=8, radius=3.0, photons=800.0, beamstop=2.0, seed=2,
                      model="truth")
    b = run(P, "atomic", "nearest")
    d = run(P, "deterministic", "nearest")
    assert np.array_equal(b["weight"] != 0, d["weight"] != 0) and rel(b["raw"], d["raw"]) < 1e-5


@pytest.mark.parametrize("side,gap_h,gap_cross", [(32, 0, 1), (64, 2, 0), (36, 0, 0), (18, 0, 0)])
@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
@pytest.mark.parametrize("likelihood", ["gaussian", "poisson"])
def test_deterministic_brick_equals_sorted_scatter(side, gap_h, gap_cross, interp, likelihood):
    """Same bits from the deterministic bricks and the sorted-scatter map
    (side 18: not a multiple of 4, both take the map), and again on a second
    run; the atomic bricks (with rotation groups at these sizes) within FP32
    reordering of them."""
    P = problems.make(side=side, n_patterns=120, n_rot=4, n_blobs=12, radius=side / 6.0, photons=2000.0,
                      beamstop=2.0, gap_h=gap_h, gap_cross=gap_cross, seed=side, model="truth")
    d = run(P, "deterministic_brick", interp, likelihood)
    m = run(P, "sorted_scatter", interp, likelihood)
    d2 = run(P, "deterministic_brick", interp, likelihood)
    dd = run(P, "deterministic", interp, likelihood)
    a = run(P, "atomic", interp, likelihood)
    for k in ("raw", "weight", "model"):
        assert np.array_equal(d[k].view(np.uint32), m[k].view(np.uint32)), k
        assert np.array_equal(d[k].view(np.uint32), d2[k].view(np.uint32)), k
        assert np.array_equal(d[k].view(np.uint32), dd[k].view(np.uint32)), k
    assert np.array_equal(a["weight"] != 0, d["weight"] != 0)
    assert rel(a["raw"], d["raw"]) < 1e-5 and rel(a["weight"], d["weight"]) < 1e-5
