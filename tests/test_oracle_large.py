"""The large-problem FP64 iteration (oracle.o64_iteration_large: masked
pixels only, count lists, zero responsibilities skipped, the Gaussian
likelihood by its expansion) against the dense FP64 functions it stands in
for at depth, on problems small enough for both: every output to FP64
rounding, with and without a restriction of the responsibilities."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import problems


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("kind", [O.GAUSSIAN, O.POISSON])
@pytest.mark.parametrize("start", ["random", "truth"])
def test_large_iteration_equals_dense(kind, start):
    P = problems.make(side=16, n_patterns=40, n_rot=2, seed=3, photons=300.0, model=start, gap_h=1)
    phi = np.linspace(0.7, 1.3, P.N)
    m64 = P.model.astype(np.float64)
    ref = O.o64_iteration(m64, P.images, P.mask, phi, P.side, P.q, P.w, P.x, P.y, P.z, kind=kind)
    g = O.o64_iteration_large(P.model, P.images, P.mask, phi, P.side, P.q, P.w, P.x, P.y, P.z, kind=kind)
    assert rel(g["L"], ref["L"]) < 1e-12
    assert np.abs(g["P"] - ref["P"]).max() < 1e-12
    assert rel(g["scaling"], ref["scaling"]) < 1e-12
    assert rel(g["model"], ref["model"]) < 1e-12
    assert rel(g["model_raw"], ref["model_raw"]) < 1e-12 and rel(g["weight"], ref["weight"]) < 1e-12
    assert abs(g["total"] - ref["total"]) < 1e-9 * abs(ref["total"])
    allowed = np.random.default_rng(1).random((P.R, P.N)) < 0.5
    ra = O.o64_proposal_iteration(m64, P.images, P.mask, phi, P.side, P.q, P.w, P.x, P.y, P.z, allowed, kind=kind)
    ga = O.o64_iteration_large(P.model, P.images, P.mask, phi, P.side, P.q, P.w, P.x, P.y, P.z, kind=kind,
                               allowed=allowed)
    assert np.abs(ga["P"] - ra["P"]).max() < 1e-12 and not ga["P"][~allowed].any()
    assert rel(ga["scaling"], ra["scaling"]) < 1e-12
    assert rel(ga["model"], ra["model"]) < 1e-12
