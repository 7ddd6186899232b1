"""CPU pins of the proposal's oracle and host maths: the fine -> coarse cell
assignment and the candidate-restricted responsibilities."""
import numpy as np

import emc_amd as E
from oracle import oracle as O


def test_assign_cells_matches_brute_force():
    qf, _ = E.rotation_table(4, half=True)
    qc, _ = E.rotation_table(2, half=True)
    cell = np.zeros(len(qf), np.int32)
    assert E.lib().emc_assign_cells(E.fptr(qf), len(qf), E.fptr(qc), len(qc), E.iptr(cell)) == 0
    ref = O.o_assign_cells(qf, qc)
    assert np.array_equal(cell, ref)
    # the sampling-2 points are among the sampling-4 points: each coarse
    # rotation lies in its own cell
    d = np.abs(qf.astype(np.float64) @ qc.T.astype(np.float64))
    own = d.argmax(axis=0)
    assert np.all(d[own, np.arange(len(qc))] > 1 - 1e-6)
    assert np.array_equal(cell[own], np.arange(len(qc)))
    # every cell is non-empty and no fine rotation is nearer another cell
    assert np.bincount(cell, minlength=len(qc)).min() >= 1
    best = d.max(axis=1)
    assert np.all(d[np.arange(len(qf)), cell] == best)


def test_restricted_softmax_all_allowed_is_dense():
    rng = np.random.default_rng(0)
    L = rng.normal(size=(50, 7)) * 3
    assert np.allclose(O.o64_restricted_softmax(L, np.ones_like(L, bool)), O.o64_softmax(L), atol=1e-14)


def test_restricted_softmax_is_conditional_posterior():
    rng = np.random.default_rng(1)
    L = rng.normal(size=(40, 9)) * 4
    A = rng.random((40, 9)) < 0.3
    A[0] = True  # every pattern has at least one allowed rotation
    P = O.o64_restricted_softmax(L, A)
    D = O.o64_softmax(L)
    cond = np.where(A, D, 0.0) / np.where(A, D, 0.0).sum(axis=0, keepdims=True)
    assert np.allclose(P, cond, atol=1e-13)
    assert np.allclose(P.sum(axis=0), 1.0)
    assert np.all(P[~A] == 0.0)
    one = np.zeros_like(A)
    one[3] = True
    assert np.allclose(O.o64_restricted_softmax(L, one)[3], 1.0)


def test_allowed_from_candidates():
    off = np.array([0, 2, 3, 6])
    rot = np.array([4, 1, 0, 2, 3, 5])
    cand = np.array([[0, 2], [1, 0]])
    A = O.allowed_from_candidates(cand, off, rot, 6, sub=[1])
    assert sorted(np.nonzero(A[:, 0])[0]) == [1, 2, 3, 4, 5]
    assert A[:, 1].all()


# ---- Poisson likelihood (the low-photon option) ----
def _small(seed=0, N=6, R=5, side=8):
    rng = np.random.default_rng(seed)
    mask = np.ones(side * side, np.int32)
    mask[:3] = 0
    slices = rng.random((R, side * side)) * 3
    slices[:, 5] = 0.0  # an empty voxel: the 1e-6 floor
    images = rng.poisson(1.5, size=(N, side * side)).astype(np.float32)
    phi = rng.random(N) + 0.5
    return mask, slices, images, phi


def test_poisson_likelihood_brute_force():
    mask, sl, img, phi = _small()
    L = O.o64_likelihood(sl, img, mask, 0.4, phi, kind=O.POISSON)
    m = mask != 0
    for i in range(img.shape[0]):
        for r in range(sl.shape[0]):
            s = np.maximum(sl[r, m], 1e-6)
            ref = (img[i, m] * (np.log(phi[i]) + np.log(s)) - phi[i] * sl[r, m]).sum()
            assert abs(L[r, i] - ref) < 1e-9 * max(1.0, abs(ref))


def test_poisson_likelihood_maximised_by_matching_slice():
    # for fixed K and phi, sum K log(phi S) - phi S is largest at S = K / phi
    mask, sl, img, phi = _small(1)
    img = np.maximum(img, 1.0)  # keep the optimum away from the floor
    best = img[0] / phi[0]
    cand = np.stack([best, best * 1.05, best * 0.95, best + 0.1])
    L = O.o64_likelihood(cand, img[:1], mask, 0.4, phi[:1], kind=O.POISSON)
    assert L[0, 0] > L[1:, 0].max()


def test_poisson_scaling_and_mstep_closed_forms():
    # one-hot P on r0 and counts proportional to that slice: phi = c exactly,
    # and the M-step returns K / phi on the masked pixels
    mask, sl, _, _ = _small(2)
    R, N2d = sl.shape
    c = np.array([2.0, 0.5, 4.0])
    r0 = 3
    img = (c[:, None] * sl[r0][None, :]).astype(np.float32)
    P = np.zeros((R, 3))
    P[r0] = 1.0
    phi = O.o64_update_scaling(img, sl, mask, P, kind=O.POISSON)
    assert np.allclose(phi, c, rtol=1e-6)
    sl2, T = O.o64_mstep(img, mask, P, phi, sl, kind=O.POISSON)
    m = mask != 0
    assert np.allclose(sl2[r0, m], sl[r0, m], rtol=1e-6)
    assert T[r0] == 3.0 and np.all(T[np.arange(R) != r0] == 0)
    assert np.all(sl2[np.arange(R) != r0][:, m] == 0.0)  # no responsibility: 0
