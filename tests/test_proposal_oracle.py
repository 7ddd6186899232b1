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
