"""Pins of the CPU oracle (oracle/) against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors and cannot be built
here (emc.c needs spimage.h and GSL; the .cu files need nvcc), so every check
below is a closed form, an invariant, a special case or a brute-force count
computed independently of the oracle's code.  Each is chosen so that a
plausible mistake (a transposed rotation, a swapped layout, a dropped term, a
wrong bound) fails it.
"""
import math

import numpy as np
import pytest

import emc_amd as E
from oracle import oracle as O
from tests import problems

PHI = (1 + math.sqrt(5)) / 2


# ---------------------------------------------------------------------------
# rotation table (emc.c:43-363)
# ---------------------------------------------------------------------------
def icosian_vertices():
    """The 120 unit quaternions of the binary icosahedral group, built from
    its textbook description (not from the reference's loops)."""
    v = []
    for s in range(16):
        v.append([0.5 * (1 if (s >> k) & 1 else -1) for k in range(4)])
    for k in range(4):
        for sg in (1, -1):
            e = [0.0] * 4
            e[k] = sg
            v.append(e)
    even = [p for p in __import__("itertools").permutations(range(4))
            if sum(1 for a in range(4) for b in range(a + 1, 4) if p[a] > p[b]) % 2 == 0]
    for p in even:
        for s in range(8):
            vals = [0.5 * (1 if s & 1 else -1), 0.5 * PHI * (1 if s & 2 else -1),
                    0.5 / PHI * (1 if s & 4 else -1), 0.0]
            e = [0.0] * 4
            for k in range(4):
                e[p[k]] = vals[k]
            v.append(e)
    v = np.array(v)
    # remove duplicates created by the zero entry's sign (none: sign bits only on nonzero)
    return np.unique(np.round(v, 12), axis=0)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_table_counts(n):
    q, w = O.rotation_list(n)
    assert q.shape[0] == 20 * (n + 5 * n ** 3)
    qh, wh = O.rotation_list(n, half=True)
    assert qh.shape[0] == 10 * (n + 5 * n ** 3)


def test_table_vertices_are_the_600_cell():
    q, _ = O.rotation_list(1)
    ref = icosian_vertices()
    assert ref.shape == (120, 4)
    got = np.unique(np.round(q.astype(np.float64), 5), axis=0)
    assert got.shape == (120, 4)
    # every table vertex matches one textbook vertex to float precision
    d = np.abs(q[:, None, :].astype(np.float64) - ref[None, :, :]).max(axis=2)
    assert (d.min(axis=1) < 2e-7).all()


@pytest.mark.parametrize("n", [2, 4])
def test_table_unit_norm_weights_symmetry(n):
    q, w = O.rotation_list(n)
    assert np.abs(np.linalg.norm(q.astype(np.float64), axis=1) - 1).max() < 3e-7
    assert (w > 0).all()
    assert abs(w.astype(np.float64).sum() - 1.0) < 1e-4
    # vertices have |v| = 1 before normalisation: equal weights
    assert np.ptp(w[:120]) < 1e-9
    # centrally symmetric: -q present for every q
    qd = q.astype(np.float64)
    dots = qd @ qd.T
    assert (dots.min(axis=1) < -1 + 1e-6).all()


def test_table_edge_midpoints_n2():
    """n = 2: the 720 points after the vertices are edge midpoints; each sits
    at cos(18 deg) from exactly two vertices (600-cell edges span 36 deg)."""
    q, _ = O.rotation_list(2)
    v = q[:120].astype(np.float64)
    mids = q[120:840].astype(np.float64)
    c = mids @ v.T
    near = (np.abs(c - math.cos(math.radians(18))) < 1e-6).sum(axis=1)
    assert (near == 2).all()


@pytest.mark.parametrize("n", [2, 4])
def test_half_table_distinct_and_weights(n):
    q, w = O.rotation_list(n)
    qh, wh = O.rotation_list(n, half=True)
    d = np.abs(qh.astype(np.float64) @ qh.astype(np.float64).T)
    np.fill_diagonal(d, 0)
    assert d.max() < 1 - 1e-6  # no two entries are the same rotation
    assert abs(wh.astype(np.float64).sum() - w.astype(np.float64).sum()) < 1e-5
    # each half entry is in the full table
    dd = np.abs(qh.astype(np.float64) @ q.astype(np.float64).T)
    assert (dd.max(axis=1) > 1 - 1e-6).all()


def _cells_of_600_cell():
    """The 600 tetrahedral cells of the 600-cell, found from the n = 1 table
    (whose vertices are pinned above) as the 4-cliques of its edge graph
    (neighbours at cos 36 deg = phi / 2), and the distance h of every cell's
    flat hyperplane from the origin."""
    v = O.rotation_list(1)[0].astype(np.float64)
    adj = np.abs(v @ v.T - PHI / 2) < 1e-5
    cells = []
    for a in range(120):
        nb = [b for b in np.nonzero(adj[a])[0] if b > a]
        for b, c, d in __import__("itertools").combinations(nb, 3):
            if adj[b, c] and adj[b, d] and adj[c, d]:
                cells.append((a, b, c, d))
    cen = v[np.array(cells)].mean(axis=1)
    h = np.linalg.norm(cen, axis=1)
    return cen / h[:, None], h


def _unnormalised_norms(q, normals, h):
    """|q_un| of every table point before it was projected onto the sphere
    (emc.c:355-360): the table interpolates linearly inside the flat cells of
    the 600-cell, so q_un lies in the hyperplane {x : x . c = h} of the cell
    containing it, the one whose unit normal c makes the largest q . c."""
    return h / (q @ normals.T).max(axis=1)


def _classes(n, count):
    """0 vertex, 1 edge, 2 face, 3 cell interior: by position in the table."""
    ev, fv = n - 1, (n - 1) * (n - 2) // 2
    cls = np.full(count, 3)
    cls[:120] = 0
    cls[120:120 + 720 * ev] = 1
    cls[120 + 720 * ev:120 + 720 * ev + 1200 * fv] = 2
    return cls


def test_600_cell_has_600_equidistant_cells():
    normals, h = _cells_of_600_cell()
    assert normals.shape == (600, 4)
    # float32 vertices: equal to 1e-7
    assert np.ptp(h) < 1e-7
    # a regular tetrahedron with vertices on the unit sphere and edge 1/phi
    assert abs(h[0] - math.sqrt(1 - 3 * (1 / PHI) ** 2 / 8)) < 1e-7


@pytest.mark.parametrize("n", [2, 3, 4, 6])
def test_table_weights_follow_inverse_cube_law(n):
    """Weights are f_class / (N |q_un|^3) (emc.c:212-217,229-334): within
    each class of points w * |q_un|^3 is one constant.  |q_un| comes from
    the cell geometry, not from the table code; an exponent 2 or 4, or a
    weight computed after normalising q, makes w * |q_un|^3 vary within the
    edge class (n >= 4) or the face and cell classes."""
    normals, h = _cells_of_600_cell()
    q, w = O.rotation_list(n)
    q = q.astype(np.float64)
    r = _unnormalised_norms(q, normals, h[0])
    assert r.max() <= 1 + 1e-6 and r.min() > 0.9
    k = w.astype(np.float64) * r ** 3
    cls = _classes(n, q.shape[0])
    for c in range(4):
        if (cls == c).any():
            kc = k[cls == c]
            assert np.ptp(kc) / kc.mean() < 2e-6, (c, np.ptp(kc) / kc.mean())
    if n >= 4:
        # the test discriminates: |q_un| varies along the edges
        re = r[cls == 1]
        assert np.ptp(re) > 1e-3


def _haar_mean_exp(a):
    """E[exp(a (u.q)^2)] for q uniform on S^3 (the Haar measure of SO(3)):
    q0 has density (2/pi) sqrt(1 - t^2), which gives
    exp(a/2) (I0(a/2) - I1(a/2))."""
    from scipy import special
    return math.exp(a / 2) * (special.i0(a / 2) - special.i1(a / 2))


def test_haar_closed_form():
    from scipy import integrate
    a = 10.0
    num = integrate.quad(lambda t: 2 / math.pi * math.exp(a * t * t) * math.sqrt(1 - t * t), -1, 1)[0]
    assert abs(num - _haar_mean_exp(a)) < 1e-9 * num


@pytest.mark.parametrize("half", [False, True])
def test_table_weights_are_a_haar_quadrature(half):
    """emc.c:212-345: the weighted n = 4 table integrates exp(10 (u.q)^2)
    over SO(3) to within 5e-5 of the Haar mean for six directions u.  Equal
    weights miss by up to 2.8e-3; swapping the vertex and edge factors, using
    1 for both, or an exponent 2 or 4 misses by more than 6e-4."""
    a = 10.0
    exact = _haar_mean_exp(a)
    q, w = O.rotation_list(4, half=half)
    q = q.astype(np.float64)
    w = w.astype(np.float64)
    w = w / w.sum()  # the table's float running sum leaves sum(w) = 1 +- 1e-5
    rng = np.random.default_rng(1)
    u = rng.normal(size=(6, 4))
    u[0] = (1, 0, 0, 0)
    u /= np.linalg.norm(u, axis=1)[:, None]
    f = np.exp(a * (q @ u.T) ** 2)
    err = np.abs(w @ f - exact) / exact
    assert err.max() < 5e-5, err
    eq = np.abs(f.mean(axis=0) - exact) / exact
    assert eq.max() > 1e-3  # equal weights fail the same check


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 10])
def test_product_table_matches_oracle_bitwise(n):
    for half in (False, True):
        qo, wo = O.rotation_list(n, half=half)
        qp, wp = E.rotation_table(n, half=half)
        assert qp.shape == qo.shape
        assert np.array_equal(qp.view(np.uint32), qo.view(np.uint32))
        assert np.array_equal(wp.view(np.uint32), wo.view(np.uint32))


# ---------------------------------------------------------------------------
# detector coordinates (emc.c:365-393)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("side", [16, 64])
def test_coordinates_closed_form(side):
    x, y, z = O.coordinates(side)
    px, py = np.meshgrid(np.arange(side), np.arange(side), indexing="xy")  # [y][x]
    ex = (px - side // 2) + 0.5
    ey = (py - side // 2) + 0.5
    r = np.hypot(ex, ey)
    th = np.arctan(r * 16e-6 / 0.15)
    ez = r * np.tan(th / 2)
    assert np.array_equal(x.reshape(side, side), ex.astype(np.float32))
    assert np.array_equal(y.reshape(side, side), ey.astype(np.float32))
    assert np.allclose(z.reshape(side, side), ez, rtol=1e-5, atol=1e-9)
    xp, yp, zp = E.coordinates(side)
    for a, b in ((x, xp), (y, yp), (z, zp)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---------------------------------------------------------------------------
# expand (emc_cuda.cu:51-111)
# ---------------------------------------------------------------------------
def index_model(S):
    return np.arange(S ** 3, dtype=np.float64)


def test_expand_identity_off_by_one():
    S = 16
    x, y, z = O.coordinates(S)
    q = np.array([[1, 0, 0, 0]], np.float32)
    sl = O.r32_expand(index_model(S).astype(np.float32), S, q, x, y, z)[0]
    for yy in range(S):
        for xx in range(S):
            p = yy * S + xx
            vx, vy, vz = xx + 1, yy + 1, S // 2 + 1
            exp = vz * S * S + vy * S + vx if (vx < S and vy < S) else 0.0
            assert sl[p] == exp, (xx, yy)


def test_expand_rotation_about_z():
    """q = (0,0,0,1) is a half turn about z: (x, y) -> voxel (S-x, S-y)."""
    S = 16
    x, y, z = O.coordinates(S)
    q = np.array([[0, 0, 0, 1]], np.float32)
    sl = O.o64_expand(index_model(S), S, q, x, y, z)[0]
    for yy in range(S):
        for xx in range(S):
            vx, vy, vz = S - xx, S - yy, S // 2 + 1
            exp = vz * S * S + vy * S + vx if (0 < vx < S and 0 < vy < S) else 0.0
            assert sl[yy * S + xx] == exp


def rotmat(q):
    a, b, c, d = (float(t) for t in q)
    return np.array([[a * a + b * b - c * c - d * d, 2 * b * c - 2 * a * d, 2 * b * d + 2 * a * c],
                     [2 * b * c + 2 * a * d, a * a - b * b + c * c - d * d, 2 * c * d - 2 * a * b],
                     [2 * b * d - 2 * a * c, 2 * c * d + 2 * a * b, a * a - b * b - c * c + d * d]])


def test_expand_linear_ramp_nearest_and_trilinear():
    S = 24
    x, y, z = O.coordinates(S)
    vz, vy, vx = np.meshgrid(np.arange(S), np.arange(S), np.arange(S), indexing="ij")
    a, b, c = 0.7, -1.3, 2.1
    ramp = (a * vx + b * vy + c * vz).reshape(-1) + 100.0
    q, _ = O.rotation_list(1, half=True)
    q = q[:12]
    nn = O.o64_expand(ramp, S, q, x, y, z, O.NEAREST)
    tri = O.o64_expand(ramp, S, q, x, y, z, O.TRILINEAR)
    for r in range(q.shape[0]):
        u = rotmat(q[r]) @ np.stack([x, y, z]).astype(np.float64) + (S / 2 + 0.5)
        exact = a * u[0] + b * u[1] + c * u[2] + 100.0
        inb_nn = ((np.round(u) > 0) & (np.round(u) < S)).all(axis=0)
        # nearest: within the rounding bound of the exact value
        assert (np.abs(nn[r][inb_nn] - exact[inb_nn]) <= 0.5 * (abs(a) + abs(b) + abs(c)) + 1e-6).all()
        assert (nn[r][~inb_nn] == 0).all()
        # trilinear reproduces a linear function exactly inside the grid
        inb_t = ((np.floor(u) >= 0) & (np.floor(u) + 1 < S)).all(axis=0)
        assert np.abs(tri[r][inb_t] - exact[inb_t]).max() < 1e-9


def test_expand_r32_equals_o64_nearest():
    P = problems.tiny()
    sl32 = O.r32_expand(P.model, P.side, P.q, P.x, P.y, P.z)
    sl64 = O.o64_expand(P.model.astype(np.float64), P.side, P.q, P.x, P.y, P.z)
    assert np.array_equal(sl32.astype(np.float64), sl64)


# ---------------------------------------------------------------------------
# likelihood (emc_cuda.cu:113-150)
# ---------------------------------------------------------------------------
def test_likelihood_closed_forms():
    P = problems.tiny(n_patterns=6)
    N, n2 = P.images.shape
    phi = np.linspace(0.5, 2.0, N)
    R = 5
    delta = np.array([0.0, 0.1, -0.3, 1.0, 2.5])
    # slice r of pattern 0 = K_0/phi_0 + delta_r on every pixel
    sl = P.images[0][None, :] / phi[0] + delta[:, None]
    sigma = 0.4
    L = O.o64_likelihood(sl, P.images, P.mask, sigma, phi)
    assert L.shape == (R, N)
    assert np.allclose(L[:, 0], -delta ** 2 / (2 * sigma ** 2), rtol=1e-12, atol=1e-12)
    L32 = O.r32_likelihood(sl.astype(np.float32), P.images, P.mask, sigma, phi.astype(np.float32))
    assert np.allclose(L32[:, 0], -delta ** 2 / (2 * sigma ** 2), rtol=1e-5, atol=1e-5)


def test_likelihood_brute_force_and_r32():
    P = problems.tiny(n_patterns=10)
    sl = O.r32_expand(P.model * 3, P.side, P.q[:40], P.x, P.y, P.z)
    phi = np.linspace(0.6, 1.7, P.N)
    L = O.o64_likelihood(sl, P.images, P.mask, 0.4, phi)
    m = P.mask != 0
    for r in (0, 7, 39):
        for i in (0, 9):
            d = sl[r][m].astype(np.float64) - P.images[i][m] / phi[i]
            assert abs(L[r, i] - (-(d * d).sum() / 2 / m.sum() / 0.16)) <= 1e-10 * abs(L[r, i])
    L32 = O.r32_likelihood(sl, P.images, P.mask, 0.4, phi.astype(np.float32))
    assert np.allclose(L32, L, rtol=2e-5)


# ---------------------------------------------------------------------------
# softmax (emc_cuda.cu:445-476) and sum P log P (emc_cuda.cu:484-505)
# ---------------------------------------------------------------------------
def test_softmax_invariants():
    rng = np.random.default_rng(3)
    R, N = 300, 17
    L = rng.normal(scale=30, size=(R, N))
    Pm = O.o64_softmax(L)
    assert np.allclose(Pm.sum(axis=0), 1, atol=1e-12)
    assert (Pm >= 0).all()
    shift = rng.normal(size=N) * 100
    assert np.allclose(O.o64_softmax(L + shift[None, :]), Pm, atol=1e-12)
    # uniform
    assert np.allclose(O.o64_softmax(np.zeros((R, N))), 1.0 / R)
    assert abs(O.o64_total_respons(np.full((R, N), 1.0 / R)) - (-N * math.log(R))) < 1e-9
    # one dominant rotation
    Lh = np.full((R, N), -1000.0)
    Lh[np.arange(N) % R, np.arange(N)] = 0
    P1 = O.o64_softmax(Lh)
    assert np.array_equal(P1 > 0.5, Lh == 0)
    # FP32 restatement agrees
    P32 = O.r32_softmax(L.astype(np.float32))
    assert np.abs(P32 - Pm).max() < 2e-6
    assert abs(O.r32_total_respons(P32) - O.o64_total_respons(Pm)) < 1e-3 * abs(O.o64_total_respons(Pm))


def test_softmax_sentinel_semantics():
    # every L far below the -1e10 floor: the reference divides 0 by 0 (NaN)
    L = np.full((4, 2), -3e10, np.float32)
    P = O.r32_softmax(L)
    assert np.isnan(P).all()


# ---------------------------------------------------------------------------
# scaling (emc_cuda.cu:195-228)
# ---------------------------------------------------------------------------
def test_scaling_known_answer():
    P_ = problems.tiny(n_patterns=5)
    R = 7
    sl = O.o64_expand(P_.truth.astype(np.float64), P_.side, P_.q[:R], P_.x, P_.y, P_.z)
    c = np.array([0.5, 1.0, 2.0, 3.5, 0.25])
    r0 = np.array([1, 3, 0, 6, 2])
    images = (c[:, None] * sl[r0]).astype(np.float32)
    Pm = np.zeros((R, 5))
    Pm[r0, np.arange(5)] = 1.0
    phi = O.o64_update_scaling(images, sl, P_.mask, Pm)
    # K = c S_r0 (rounded to float32) with one-hot P at r0 gives phi = c
    assert np.allclose(phi, c, rtol=1e-6)
    phi32 = O.r32_update_scaling(images, sl.astype(np.float32), P_.mask, Pm.astype(np.float32))
    assert np.allclose(phi32, c, rtol=1e-5)


# ---------------------------------------------------------------------------
# M-step, insertion, normalisation (emc_atomic.cu:17-133, emc_cuda.cu:414-433)
# ---------------------------------------------------------------------------
def test_mstep_insert_round_trip():
    """One-hot responsibilities and K_i = phi_i * expand(M, r_i): the M-step
    recovers the slices and insertion + normalisation gives back M (up to the
    mean factor) on every voxel the slices cover."""
    Pr = problems.tiny(n_patterns=30, side=16)
    S, R = Pr.side, 30
    q = Pr.q[::7][:R]
    M = Pr.truth.astype(np.float64) + 0.5
    sl = O.o64_expand(M, S, q, Pr.x, Pr.y, Pr.z)
    phi = np.linspace(0.5, 1.5, R)
    images = (sl * phi[:, None]).astype(np.float32)  # pattern i <- rotation i
    Pm = np.eye(R)
    w = np.linspace(1, 2, R).astype(np.float32)
    sl2, T = O.o64_mstep(images, Pr.mask, Pm, phi, sl)
    m = Pr.mask != 0
    assert np.allclose(T, 1.0)
    assert np.allclose(sl2[:, m], images[:, m] / phi[:, None], rtol=1e-12)
    raw, wgt = O.o64_insert(sl2, Pr.mask, T, w, S, q, Pr.x, Pr.y, Pr.z)
    hit = wgt > 0
    assert hit.sum() > 100
    norm = O.o64_normalize_model(raw, wgt)
    assert abs(norm.mean() - 1) < 1e-12
    assert (norm[~hit] == 0).all()
    # voxels of plane 0 are inserted with the value 0 (expand never reads them)
    vz, vy, vx = np.unravel_index(np.arange(S ** 3), (S, S, S))
    interior = hit & (vx > 0) & (vy > 0) & (vz > 0)
    ratio = norm[interior] / M[interior]
    # exact up to the float32 rounding of K = phi * S
    assert np.ptp(ratio) < 1e-6 * ratio.mean()
    # the FP32 restatement reproduces this within float precision
    m32 = np.zeros(S ** 3, np.float32)
    w32 = np.zeros(S ** 3, np.float32)
    ov, sl32, T32 = O.r32_update_slices(images, sl.astype(np.float32), Pr.mask, Pm.astype(np.float32),
                                        phi.astype(np.float32), S, m32, w32, Pr.x, Pr.y, Pr.z, q, w)
    assert abs(ov - R) < 1e-4
    assert np.allclose(m32, raw, rtol=1e-5, atol=1e-6)
    assert np.allclose(w32, wgt, rtol=1e-5)


def test_insert_trilinear_partition_of_unity():
    """Trilinear insertion spreads each contribution with weights summing to
    one: total model mass = sum of w_r T_r S' over in-grid footprints."""
    Pr = problems.tiny(side=16)
    S = Pr.side
    q = Pr.q[:3]
    sl = np.ones((3, S * S))
    T = np.ones(3)
    w = np.ones(3, np.float32)
    mask = np.zeros(S * S, np.int32)
    # central pixels only, so every footprint is fully inside the grid
    for yy in range(S // 2 - 3, S // 2 + 3):
        for xx in range(S // 2 - 3, S // 2 + 3):
            mask[yy * S + xx] = 1
    raw, wgt = O.o64_insert(sl, mask, T, w, S, q, Pr.x, Pr.y, Pr.z, O.TRILINEAR)
    assert abs(wgt.sum() - 3 * mask.sum()) < 1e-9
    assert abs(raw.sum() - 3 * mask.sum()) < 1e-9


def test_normalize_model_closed_form():
    rng = np.random.default_rng(0)
    m = rng.random(4096)
    w = rng.random(4096)
    w[::7] = 0
    out = O.o64_normalize_model(m, w)
    exp = np.where(w > 0, m / np.where(w > 0, w, 1), 0)
    exp /= exp.mean()
    assert np.allclose(out, exp, rtol=1e-12)
    o32 = O.r32_normalize_model(m.astype(np.float32), w.astype(np.float32))
    assert np.allclose(o32, exp, rtol=1e-5)


# ---------------------------------------------------------------------------
# whole iteration
# ---------------------------------------------------------------------------
def test_iteration_r32_vs_o64():
    Pr = problems.tiny()
    sc = np.ones(Pr.N, np.float32)
    m32, w32, s32, P32, t32 = O.r32_iteration(Pr.model, Pr.images, Pr.mask, sc, Pr.side, Pr.x, Pr.y, Pr.z,
                                               Pr.q, Pr.w)
    ref = O.o64_iteration(Pr.model.astype(np.float64), Pr.images, Pr.mask, sc.astype(np.float64), Pr.side,
                          Pr.q, Pr.w, Pr.x, Pr.y, Pr.z)
    assert np.abs(P32 - ref["P"]).max() < 1e-4
    assert np.allclose(s32, ref["scaling"], rtol=1e-4)
    big = ref["model"] > 1e-3 * ref["model"].max()
    assert np.abs(m32[big] / ref["model"][big] - 1).max() < 1e-4


def test_dedupe_equivalence():
    """q and -q give identical slices; the half table's responsibilities are
    the sums over antipodal pairs and the normalised model is unchanged."""
    Pr = problems.tiny(n_patterns=20)
    qf, wf = O.rotation_list(2)
    qh, wh = O.rotation_list(2, half=True)
    sc = np.ones(Pr.N)
    M = Pr.model.astype(np.float64)
    full = O.o64_iteration(M, Pr.images, Pr.mask, sc, Pr.side, qf, wf, Pr.x, Pr.y, Pr.z)
    half = O.o64_iteration(M, Pr.images, Pr.mask, sc, Pr.side, qh, wh, Pr.x, Pr.y, Pr.z)
    # match half entries to their pair in the full table
    dots = qh.astype(np.float64) @ qf.astype(np.float64).T
    same = np.argmax(dots, axis=1)
    anti = np.argmin(dots, axis=1)
    assert np.allclose(half["P"], full["P"][same] + full["P"][anti], atol=1e-9)
    assert np.allclose(half["scaling"], full["scaling"], rtol=1e-9)
    assert np.allclose(half["model"], full["model"], rtol=1e-5, atol=1e-9)


# ---------------------------------------------------------------------------
# product host code
# ---------------------------------------------------------------------------
def test_synthetic_data_statistics():
    Pr = problems.make(side=32, n_patterns=200, photons=2000.0, beamstop=3.0, n_blobs=20, radius=6.0)
    assert abs(Pr.truth.astype(np.float64).mean() - 1) < 1e-5
    assert (Pr.truth >= 0).all()
    counts = Pr.images.sum(axis=1)
    assert abs(counts.mean() / 2000.0 - 1) < 0.05
    assert np.array_equal(Pr.images, np.round(Pr.images))
    assert (Pr.images[:, Pr.mask == 0] == 0).all()
    assert abs(Pr.phi_true.mean() - 1) < 1e-5
    # reproducible
    Pr2 = problems.make(side=32, n_patterns=200, photons=2000.0, beamstop=3.0, n_blobs=20, radius=6.0)
    assert np.array_equal(Pr.images, Pr2.images)


def test_mask_counts():
    # configuration 0: 32x32 with a radius-3 beamstop
    assert E.make_mask(32, 3.0).sum() == 1024 - 32  # 32 pixel centres within r < 3
    m = E.make_mask(64, 4.0, gap_h=2)
    assert m.reshape(64, 64)[31:33].sum() == 0
    assert 3900 < m.sum() < 3960
    m2 = E.make_mask(128, 6.0, gap_cross=3)
    assert 15400 < m2.sum() < 15800


def test_host_insert_slice_identity():
    """emc.h insert_slice (CPU, emc.c:491-539): identity rotation puts pixel
    (x, y) into voxel (x+1, y+1, S/2+1); bounds 0 <= v < S."""
    import ctypes as C
    S = 8
    x, y, z = E.coordinates(S)
    model = np.zeros(S ** 3, np.float32)
    weight = np.zeros(S ** 3, np.float32)
    sl = np.arange(S * S, dtype=np.float32) + 1
    mask = np.ones(S * S, np.int32)
    sm, sw = E.sp3(S, model), E.sp3(S, weight)
    ms = E.SpMatrix(S, S, E.fptr(sl))
    mk = E.SpIMatrix(S, S, E.iptr(mask))
    xs, ys, zs = (E.SpMatrix(S, S, E.fptr(a)) for a in (x, y, z))
    qq = E.Quaternion((C.c_float * 4)(1, 0, 0, 0))
    E.lib().insert_slice(C.byref(sm), C.byref(sw), C.byref(ms), C.byref(mk), 2.0, C.byref(qq), C.byref(xs),
                         C.byref(ys), C.byref(zs))
    for yy in range(S):
        for xx in range(S):
            vx, vy, vz = xx + 1, yy + 1, S // 2 + 1
            if vx < S and vy < S:
                v = (vz * S + vy) * S + vx
                assert model[v] == 2.0 * sl[yy * S + xx] and weight[v] == 2.0
    assert weight.sum() == 2.0 * (S - 1) * (S - 1)


def test_model_max_is_clamped_maximum():
    """cuda_model_max (emc_cuda.cu:364-368) reduces with max starting from 0:
    max(0, max M); an all-negative model gives 0."""
    rng = np.random.default_rng(4)
    m = rng.normal(size=5000).astype(np.float32)
    assert O.r32_model_max(m) == float(m.max())
    assert O.r32_model_max(-np.abs(m) - 1.0) == 0.0
    m[1234] = 7.5
    assert O.r32_model_max(m) == 7.5
