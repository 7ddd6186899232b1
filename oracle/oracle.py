"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker.  The product never
imports it.  See oracle.h for what each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NEAREST, TRILINEAR = 0, 1
GAUSSIAN, POISSON = 0, 1

_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)

_SIGS = [
    ("o_n_to_samples", C.c_int, [C.c_int]),
    ("o_rotation_list", C.c_int, [C.c_int, _f, _f]),
    ("o_rotation_list_half", C.c_int, [C.c_int, _f, _f]),
    ("o_coordinates", None, [C.c_int, C.c_float, C.c_float, C.c_float, _f, _f, _f]),
    ("o_voxel_index", None, [_f, C.c_float, C.c_float, C.c_float, C.c_int, _i, _i, _i]),
    ("r32_expand", None, [_f, C.c_int, _f, C.c_int, _f, _f, _f, _f]),
    ("r32_likelihood", None, [_f, _f, _i, C.c_float, _f, _f, C.c_int, C.c_int, C.c_int]),
    ("r32_softmax", None, [_f, C.c_int, C.c_int]),
    ("r32_total_respons", C.c_double, [_f, C.c_size_t]),
    ("r32_update_scaling", None, [_f, _f, _i, _f, _f, C.c_int, C.c_int, C.c_int]),
    ("r32_update_slices", C.c_float, [_f, _f, _i, _f, _f, C.c_int, C.c_int, C.c_int, C.c_int, _f, _f, _f, _f,
                                      _f, _f, _f, _f]),
    ("r32_mstep", None, [_f, _f, _i, _f, _f, C.c_int, C.c_int, C.c_int, _f]),
    ("r32_insert_all", None, [_f, _f, C.c_int, _f, _i, _f, _f, _f, C.c_int, _f, _f, _f]),
    ("r32_insert_all_par", None, [_f, _f, C.c_int, _f, _i, _f, _f, _f, C.c_int, _f, _f, _f]),
    ("r32_normalize_model", None, [_f, _f, C.c_int]),
    ("r32_model_max", C.c_float, [_f, C.c_int]),
    ("r32_insert_slice", None, [_f, _f, C.c_int, _f, _i, C.c_float, _f, _f, _f, _f]),
    ("o64_expand", None, [_d, C.c_int, _f, C.c_int, _f, _f, _f, C.c_int, _d]),
    ("o64_likelihood", None, [_d, _f, _i, C.c_double, _d, C.c_int, C.c_int, C.c_int, C.c_int, _d]),
    ("o64_softmax", None, [_d, C.c_int, C.c_int, _d]),
    ("o64_total_respons", C.c_double, [_d, C.c_size_t]),
    ("o64_update_scaling", None, [_f, _d, _i, _d, C.c_int, C.c_int, C.c_int, C.c_int, _d]),
    ("o64_mstep", None, [_f, _i, _d, _d, C.c_int, C.c_int, C.c_int, C.c_int, _d, _d, _d]),
    ("o64_insert", None, [_d, _i, _d, _f, C.c_int, C.c_int, C.c_int, _f, _f, _f, _f, C.c_int, _d, _d]),
    ("o64_normalize_model", None, [_d, _d, C.c_int]),
    ("r32_iteration", C.c_double, [_f, _f, _f, _f, _i, _f, _f, C.c_int, C.c_int, C.c_int, _f, _f, _f, _f, _f,
                                   C.c_float, C.c_int]),
    ("o_max_threads", C.c_int, []),
    ("o_set_threads", None, [C.c_int]),
    ("o64_iteration_large", C.c_double, [_d, _f, _i, _d, C.c_int, C.c_int, C.c_int, _f, _f, _f, _f, _f,
                                         C.c_double, C.c_int, C.c_void_p, _d, _d, _d, _d, _d, _d]),
]

_LIB = None


def build() -> None:
    subprocess.run(["make", "-B", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_f)


def d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_d)


def i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_i)


def c32(a):
    return np.ascontiguousarray(a, np.float32)


def c64(a):
    return np.ascontiguousarray(a, np.float64)


def ci(a):
    return np.ascontiguousarray(a, np.int32)


# ---- host maths ----
def rotation_list(n, half=False):
    L = lib()
    ns = L.o_n_to_samples(n)
    q = np.zeros((ns, 4), np.float32)
    w = np.zeros(ns, np.float32)
    m = (L.o_rotation_list_half if half else L.o_rotation_list)(n, f(q), f(w))
    assert m > 0
    return q[:m].copy(), w[:m].copy()


def coordinates(side, wavelength=5.6e-9, pixel_size=16e-6, distance=0.15):
    n2 = side * side
    x, y, z = (np.zeros(n2, np.float32) for _ in range(3))
    lib().o_coordinates(side, wavelength, pixel_size, distance, f(x), f(y), f(z))
    return x, y, z


def voxel_index(q, xc, yc, zc, S):
    vx, vy, vz = C.c_int(), C.c_int(), C.c_int()
    lib().o_voxel_index(f(c32(q)), xc, yc, zc, S, C.byref(vx), C.byref(vy), C.byref(vz))
    return vx.value, vy.value, vz.value


# ---- FP32 restatement ----
def r32_expand(model, S, q, x, y, z):
    R = q.shape[0]
    out = np.zeros((R, S * S), np.float32)
    lib().r32_expand(f(c32(model)), S, f(c32(q)), R, f(x), f(y), f(z), f(out))
    return out


def r32_likelihood(slices, images, mask, sigma, scaling):
    R, N2d = slices.shape
    N = images.shape[0]
    out = np.zeros((R, N), np.float32)
    lib().r32_likelihood(f(c32(slices)), f(c32(images)), i(ci(mask)), sigma, f(c32(scaling)), f(out), N2d, N, R)
    return out


def r32_softmax(L):
    P = c32(L).copy()
    R, N = P.shape
    lib().r32_softmax(f(P), R, N)
    return P


def r32_total_respons(P):
    P = c32(P)
    return lib().r32_total_respons(f(P), P.size)


def r32_update_scaling(images, slices, mask, P):
    N = images.shape[0]
    R, N2d = slices.shape
    out = np.zeros(N, np.float32)
    lib().r32_update_scaling(f(c32(images)), f(c32(slices)), i(ci(mask)), f(c32(P)), f(out), N, R, N2d)
    return out


def r32_update_slices(images, slices, mask, P, scaling, S, model, weight, x, y, z, q, weights):
    """Returns (overall, slices_out, T); model/weight are updated in place."""
    sl = c32(slices).copy()
    R, N2d = sl.shape
    N = images.shape[0]
    T = np.zeros(R, np.float32)
    ov = lib().r32_update_slices(f(c32(images)), f(sl), i(ci(mask)), f(c32(P)), f(c32(scaling)), N, R, N2d, S,
                                 f(model), f(weight), f(x), f(y), f(z), f(c32(q)), f(c32(weights)), f(T))
    return ov, sl, T


def r32_normalize_model(model, weight):
    m = c32(model).copy()
    lib().r32_normalize_model(f(m), f(c32(weight)), m.size)
    return m


def r32_model_max(model):
    """max(0, max M) as cuda_model_max returns it (emc_cuda.cu:364-368)."""
    m = c32(model).reshape(-1)
    return float(lib().r32_model_max(f(m), m.size))


def r32_iteration(model, images, mask, scaling, S, x, y, z, q, weights, sigma=0.4, nthreads=0):
    """One full iteration; returns (model, weight, scaling, P, total)."""
    m = c32(model).copy()
    wgt = np.zeros_like(m)
    R = q.shape[0]
    N = images.shape[0]
    sl = np.zeros((R, S * S), np.float32)
    P = np.zeros((R, N), np.float32)
    sc = c32(scaling).copy()
    tot = lib().r32_iteration(f(m), f(wgt), f(sl), f(c32(images)), i(ci(mask)), f(sc), f(P), N, R, S, f(x), f(y),
                              f(z), f(c32(q)), f(c32(weights)), sigma, nthreads)
    return m, wgt, sc, P, tot


def time_reference_iteration(model, images, mask, S, x, y, z, q, weights, n_full, sigma=0.4, nthreads=0):
    """Wall time of one FP32-restatement iteration (emc.c:989-1084 order) on
    the pattern sample `images`, phase by phase, and its extrapolation to
    n_full patterns: the per-pattern phases (likelihood, softmax, sum P log P,
    scaling, M-step) scale with the pattern count, expand, insertion and
    normalisation do not."""
    import time

    L = lib()
    R = q.shape[0]
    Ns = images.shape[0]
    N2d = S * S
    images = c32(images)
    q = c32(q)
    weights = c32(weights)
    mask = ci(mask)
    sl = np.zeros((R, N2d), np.float32)
    P = np.zeros((R, Ns), np.float32)
    sc = np.ones(Ns, np.float32)
    m = c32(model).copy()
    wg = np.zeros_like(m)
    T = np.zeros(R, np.float32)
    t = {}

    def run(name, fn):
        t0 = time.perf_counter()
        fn()
        t[name] = time.perf_counter() - t0

    saved = L.o_max_threads()
    if nthreads > 0:
        L.o_set_threads(nthreads)
    threads = L.o_max_threads()
    run("expand", lambda: L.r32_expand(f(m), S, f(q), R, f(x), f(y), f(z), f(sl)))
    run("likelihood", lambda: L.r32_likelihood(f(sl), f(images), i(mask), sigma, f(sc), f(P), N2d, Ns, R))
    run("softmax", lambda: L.r32_softmax(f(P), R, Ns))
    run("total", lambda: L.r32_total_respons(f(P), P.size))
    m[:] = 0
    run("scaling", lambda: L.r32_update_scaling(f(images), f(sl), i(mask), f(P), f(sc), Ns, R, N2d))
    run("mstep", lambda: L.r32_mstep(f(images), f(sl), i(mask), f(P), f(sc), Ns, R, N2d, f(T)))
    run("insert", lambda: L.r32_insert_all_par(f(m), f(wg), S, f(sl), i(mask), f(T), f(q), f(weights), R, f(x), f(y),
                                           f(z)))
    run("normalize", lambda: L.r32_normalize_model(f(m), f(wg), m.size))
    dep = t["likelihood"] + t["softmax"] + t["total"] + t["scaling"] + t["mstep"]
    indep = t["expand"] + t["insert"] + t["normalize"]
    L.o_set_threads(saved)
    return {"phases": t, "sample_s": dep + indep, "full_s": indep + dep * n_full / Ns, "threads": threads}


# ---- FP64 oracle ----
def o64_expand(model, S, q, x, y, z, interp=NEAREST):
    R = q.shape[0]
    out = np.zeros((R, S * S), np.float64)
    lib().o64_expand(d(c64(model)), S, f(c32(q)), R, f(x), f(y), f(z), interp, d(out))
    return out


def o64_likelihood(slices, images, mask, sigma, scaling, kind=GAUSSIAN):
    R, N2d = slices.shape
    N = images.shape[0]
    out = np.zeros((R, N), np.float64)
    lib().o64_likelihood(d(c64(slices)), f(c32(images)), i(ci(mask)), sigma, d(c64(scaling)), N2d, N, R, kind,
                         d(out))
    return out


def o64_softmax(L):
    R, N = L.shape
    out = np.zeros((R, N), np.float64)
    lib().o64_softmax(d(c64(L)), R, N, d(out))
    return out


def o64_total_respons(P):
    P = c64(P)
    return lib().o64_total_respons(d(P), P.size)


def o64_update_scaling(images, slices, mask, P, kind=GAUSSIAN):
    N = images.shape[0]
    R, N2d = slices.shape
    out = np.zeros(N, np.float64)
    lib().o64_update_scaling(f(c32(images)), d(c64(slices)), i(ci(mask)), d(c64(P)), N, R, N2d, kind, d(out))
    return out


def o64_mstep(images, mask, P, scaling, slices_in, kind=GAUSSIAN):
    R, N = P.shape
    N2d = images.shape[1]
    out = np.zeros((R, N2d), np.float64)
    T = np.zeros(R, np.float64)
    lib().o64_mstep(f(c32(images)), i(ci(mask)), d(c64(P)), d(c64(scaling)), N, R, N2d, kind, d(c64(slices_in)),
                    d(out), d(T))
    return out, T


def o64_insert(slices, mask, T, weights, S, q, x, y, z, interp=NEAREST, model=None, weight=None):
    R, N2d = slices.shape
    m = np.zeros(S ** 3, np.float64) if model is None else c64(model).copy()
    wg = np.zeros(S ** 3, np.float64) if weight is None else c64(weight).copy()
    lib().o64_insert(d(c64(slices)), i(ci(mask)), d(c64(T)), f(c32(weights)), R, N2d, S, f(c32(q)), f(x), f(y),
                     f(z), interp, d(m), d(wg))
    return m, wg


def o64_normalize_model(model, weight):
    m = c64(model).copy()
    lib().o64_normalize_model(d(m), d(c64(weight)), m.size)
    return m


def o64_iteration(model, images, mask, scaling, S, q, weights, x, y, z, sigma=0.4, interp=NEAREST, kind=GAUSSIAN):
    """One FP64 iteration in the driver order; returns a dict of every output."""
    sl = o64_expand(model, S, q, x, y, z, interp)
    L = o64_likelihood(sl, images, mask, sigma, scaling, kind)
    P = o64_softmax(L)
    tot = o64_total_respons(P)
    phi = o64_update_scaling(images, sl, mask, P, kind)
    sl2, T = o64_mstep(images, mask, P, phi, sl, kind)
    m, wg = o64_insert(sl2, mask, T, weights, S, q, x, y, z, interp)
    mn = o64_normalize_model(m, wg)
    return {"slices": sl, "L": L, "P": P, "total": tot, "scaling": phi, "slices_new": sl2, "T": T,
            "model_raw": m, "weight": wg, "model": mn}


def o64_iteration_large(model, images, mask, scaling, S, q, weights, x, y, z, sigma=0.4, kind=GAUSSIAN,
                        allowed=None):
    """One FP64 iteration (nearest voxel) for problems too large for
    o64_iteration: the same equations over the masked pixels, with sparse
    count lists and zero responsibilities skipped.  `allowed` ([R][N] bool)
    restricts each pattern's softmax as o64_restricted_softmax does."""
    images = c32(images)
    N = images.shape[0]
    R = q.shape[0]
    L = np.zeros((R, N), np.float64)
    P = np.zeros((R, N), np.float64)
    phi = np.zeros(N, np.float64)
    m = np.zeros(S ** 3, np.float64)
    raw = np.zeros(S ** 3, np.float64)
    wgt = np.zeros(S ** 3, np.float64)
    al = None
    if allowed is not None:
        allowed = np.ascontiguousarray(allowed, np.uint8)
        assert allowed.shape == (R, N)
        al = allowed.ctypes.data_as(C.c_void_p)
    tot = lib().o64_iteration_large(d(c64(model)), f(images), i(ci(mask)), d(c64(scaling)), N, R, S, f(c32(q)),
                                    f(c32(weights)), f(c32(x)), f(c32(y)), f(c32(z)), sigma, kind, al, d(L), d(P),
                                    d(phi), d(m), d(raw), d(wgt))
    return {"L": L, "P": P, "total": tot, "scaling": phi, "model": m, "model_raw": raw, "weight": wgt}


# ---- learned proposal: candidate-restricted iteration (numpy, FP64) ----
def o_assign_cells(q_fine, q_coarse):
    """Coarse cell of each fine rotation: argmax_c |q_fine[r] . q_coarse[c]|,
    lowest c on ties, products summed in the order q0..q3 in double."""
    a = np.asarray(q_fine, np.float64)
    b = np.asarray(q_coarse, np.float64)
    out = np.empty(a.shape[0], np.int64)
    for s in range(0, a.shape[0], 2048):
        blk = a[s:s + 2048]
        dots = np.abs(((blk[:, None, 0] * b[None, :, 0] + blk[:, None, 1] * b[None, :, 1])
                       + blk[:, None, 2] * b[None, :, 2]) + blk[:, None, 3] * b[None, :, 3])
        out[s:s + 2048] = dots.argmax(axis=1)
    return out


def o64_restricted_softmax(L, allowed):
    """Responsibilities of each pattern over its allowed rotations only, with
    the reference's rule (emc_cuda.cu:445-476): m = max(-1e10, max L),
    P = exp(L - m) where L - m > -1e10, normalised; zero where not allowed."""
    L = np.asarray(L, np.float64)
    allowed = np.asarray(allowed, bool)
    Lm = np.where(allowed, L, -np.inf)
    m = np.maximum(-1e10, Lm.max(axis=0))
    d = Lm - m[None, :]
    e = np.where(allowed & (d > -1e10), np.exp(np.where(allowed, d, 0.0)), 0.0)
    return e / e.sum(axis=0, keepdims=True)


def allowed_from_candidates(cand, cell_off, cell_rot, R, sub=None):
    """[R][N] mask of the rotations a pattern is scored on: those of its
    candidate cells, or every rotation for the patterns in `sub`."""
    N = cand.shape[0]
    A = np.zeros((R, N), bool)
    for i in range(N):
        for c in cand[i]:
            A[cell_rot[cell_off[c]:cell_off[c + 1]], i] = True
    if sub is not None:
        A[:, np.asarray(sub, np.int64)] = True
    return A


def o64_proposal_iteration(model, images, mask, scaling, S, q, weights, x, y, z, allowed, sigma=0.4,
                           interp=NEAREST, kind=GAUSSIAN):
    """One FP64 iteration in the driver order with the responsibilities
    restricted to `allowed` ([R][N] bool)."""
    sl = o64_expand(model, S, q, x, y, z, interp)
    L = o64_likelihood(sl, images, mask, sigma, scaling, kind)
    P = o64_restricted_softmax(L, allowed)
    tot = o64_total_respons(P)
    phi = o64_update_scaling(images, sl, mask, P, kind)
    sl2, T = o64_mstep(images, mask, P, phi, sl, kind)
    m, wg = o64_insert(sl2, mask, T, weights, S, q, x, y, z, interp)
    mn = o64_normalize_model(m, wg)
    return {"slices": sl, "L": L, "P": P, "total": tot, "scaling": phi, "slices_new": sl2, "T": T,
            "model_raw": m, "weight": wg, "model": mn}
