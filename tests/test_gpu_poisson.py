"""The Poisson likelihood (low-photon option) through the C-ABI against the
FP64 oracle of the same equations, dense (FP32 and FP16 log-slice
operand) and candidate-restricted.

Tolerances: responsibilities 1e-4 absolute, scaling and model 1e-4
relative.  (The matrix cores' FP32 sums of L ~ 10^3 here are off by ~1e-4
in L; the responsibilities of at least 1e-3 are recomputed from FP64
likelihoods after the softmax, option "refine".)"""
import numpy as np
import pytest

import emc_amd as E
from oracle import oracle as O
from tests import problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def low():
    # 32^3, 32x32 patterns, 300 photons per pattern, n=4 (3,240 rotations)
    return problems.make(side=32, n_patterns=200, n_rot=4, n_blobs=20, radius=6.0, photons=300.0,
                         phi_sigma=0.3, beamstop=3.0, seed=5, model="truth")


@pytest.fixture(autouse=True)
def _reset():
    E.lib().emc_hip_clear_error()
    yield
    for k, v in (("likelihood", "gaussian"), ("proposal", "off"), ("precision", "fp32"), ("insert", "atomic"),
                 ("gemm_tail", "split")):
        E.set_option(k, v)
    E.lib().emc_hip_clear_error()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def gpu_iteration(P, precision):
    E.set_option("precision", precision)
    E.set_option("likelihood", "poisson")
    E.set_option("insert", "deterministic")
    prob = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
    it = prob.iteration(0.4)
    out = {"P": prob.respons(), "scaling": prob.scaling(), "model": prob.model(), "total": it["likelihood"]}
    prob.free()
    return out


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_poisson_iteration_vs_fp64(low, precision):
    P = low
    g = gpu_iteration(P, precision)
    ref = O.o64_iteration(P.model.astype(np.float64), P.images, P.mask, np.ones(P.N), P.side, P.q, P.w, P.x, P.y,
                          P.z, kind=O.POISSON)
    dP = np.abs(g["P"] - ref["P"]).max()
    ds = rel(g["scaling"], ref["scaling"])
    dm = rel(g["model"], ref["model"])
    dt = abs(g["total"] - ref["total"]) / abs(ref["total"])
    print(f"Poisson {precision} vs FP64: P {dP:.2e}, scaling {ds:.2e}, model {dm:.2e}, sum PlogP {dt:.2e}; "
          f"P max {ref['P'].max():.3f}")
    assert dP < 1e-4 and ds < 1e-4 and dm < 1e-4


def test_poisson_candidates_all_cells_reproduce_dense(low):
    """Same operands and k order (the dense GEMM unsplit along k)."""
    P = low
    E.set_option("gemm_tail", "whole")
    dense = gpu_iteration(P, "fp16")
    L = E.lib()
    assert L.emc_hip_proposal_init(E.fptr(P.q), P.R, 2, P.side, 4, 64, 3) == 0
    nc = L.emc_hip_proposal_cells(None, None)
    cand = np.tile(np.arange(nc, dtype=np.int32), (P.N, 1))
    assert L.emc_hip_proposal_set_candidates(E.iptr(cand), P.N, nc) == 0
    E.set_option("proposal", "on")
    E.set_option("audit_fraction", "0")
    E.set_option("proposal_gate", "audit")
    c = gpu_iteration(P, "fp16")
    dP = np.abs(c["P"] - dense["P"]).max()
    dm = rel(c["model"], dense["model"])
    ds = rel(c["scaling"], dense["scaling"])
    print(f"Poisson all cells vs dense FP16: P {dP:.2e}, scaling {ds:.2e}, model {dm:.2e}")
    assert dP < 1e-6 and ds < 1e-6 and dm < 1e-5


def to_csr(images):
    rows, cols = np.nonzero(images)
    off = np.zeros(images.shape[0] + 1, np.int32)
    np.add.at(off, rows + 1, 1)
    off = np.cumsum(off).astype(np.int32)
    return off, cols.astype(np.int32), images[rows, cols].astype(np.uint16)


def test_images_from_csr_roundtrip(low):
    P = low
    off, pix, cnt = to_csr(P.images)
    d = E.DeviceArray(4 * P.images.size)
    assert E.lib().emc_hip_images_from_csr(d.vp, E.iptr(off), E.iptr(pix),
                                           cnt.ctypes.data_as(E.C.POINTER(E.C.c_ushort)), P.N, P.side ** 2) == 0
    back = d.download(np.float32, P.images.size).reshape(P.images.shape)
    d.free()
    E.check()
    print(f"CSR: {len(pix)} nonzeros of {P.images.size} ({len(pix) / P.images.size:.4f})")
    assert np.array_equal(back, P.images)


def test_sparse_estep_matches_dense_and_fp64(low):
    P = low
    dense = gpu_iteration(P, "fp16")
    E.set_option("sparse", "on")
    try:
        sp = gpu_iteration(P, "fp16")
    finally:
        E.set_option("sparse", "off")
    ref = O.o64_iteration(P.model.astype(np.float64), P.images, P.mask, np.ones(P.N), P.side, P.q, P.w, P.x, P.y,
                          P.z, kind=O.POISSON)
    dPd = np.abs(sp["P"] - dense["P"]).max()
    dP = np.abs(sp["P"] - ref["P"]).max()
    dm = rel(sp["model"], ref["model"])
    ds = rel(sp["scaling"], ref["scaling"])
    print(f"sparse Poisson E-step: vs dense FP16 P {dPd:.2e}; vs FP64 P {dP:.2e}, scaling {ds:.2e}, model {dm:.2e}")
    assert dPd < 1e-4 and dP < 1e-4 and ds < 1e-4 and dm < 1e-4
