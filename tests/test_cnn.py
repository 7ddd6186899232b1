"""The rotation-proposal network: oracle pinned by finite differences and an
adjoint identity (CPU); the HIP forward, backward and Adam step checked
against the oracle (GPU)."""
import ctypes as C

import numpy as np
import pytest

import emc_amd as E
from oracle import cnn as O
from tests import problems


def test_im2col_col2im_adjoint():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(2, 8, 8, 3))
    c = rng.normal(size=(2 * 4 * 4, 27))
    assert abs((O.im2col(a) * c).sum() - (a * O.col2im(c, 2, 8, 8, 3)).sum()) < 1e-10


def test_im2col_matches_direct_convolution():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(1, 6, 6, 2))
    W = rng.normal(size=(3, 3, 3, 2))  # [OC][ky][kx][IC]
    out = (O.im2col(a) @ W.reshape(3, -1).T).reshape(1, 3, 3, 3)
    pad = np.pad(a, ((0, 0), (1, 1), (1, 1), (0, 0)))
    for oy in range(3):
        for ox in range(3):
            for oc in range(3):
                s = (pad[0, 2 * oy:2 * oy + 3, 2 * ox:2 * ox + 3, :] * W[oc]).sum()
                assert abs(out[0, oy, ox, oc] - s) < 1e-12


def test_oracle_gradient_finite_differences():
    rng = np.random.default_rng(2)
    lay = O.Layout(16, 7)
    p = rng.normal(size=lay.n) * 0.2
    x = rng.random(size=(3, 16, 16, 1))
    t = rng.random(size=(7, 3))
    t /= t.sum(axis=0)
    loss, g, _ = O.loss_and_grad(p, lay, x, t)
    h = 1e-6
    for k in rng.choice(lay.n, 60, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (O.loss_and_grad(pp, lay, x, t)[0] - O.loss_and_grad(pm, lay, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(g[k]), (k, fd, g[k])


def test_bf16_rounding():
    # bf16 spacing at 1 is 2^-7; 1 + 2^-8 and 1 + 3 * 2^-8 are halfway cases
    v = np.array([1.0, 1.00390625, 1.01171875, 3.0e-3, -2.5], np.float32)
    r = O.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.015625  # ties to even
    assert abs(r[3] - 3.0e-3) <= 3.0e-3 * 2 ** -8
    assert r[4] == -2.5


@pytest.fixture(scope="module")
def net_setup():
    P = problems.tiny(side=16, n_patterns=24)
    d_img = E.DeviceArray.from_host(P.images.reshape(-1))
    d_mask = E.DeviceArray.from_host(P.mask)
    yield P, d_img, d_mask
    d_img.free()
    d_mask.free()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(net_setup):
    P, d_img, d_mask = net_setup
    L = E.lib()
    ncell, B = 11, 8
    assert L.emc_hip_cnn_init(16, ncell, 16, 5) == 0
    E.check()
    n = L.emc_hip_cnn_param_count()
    lay = O.Layout(16, ncell)
    assert n == lay.n
    p = np.zeros(n, np.float32)
    assert L.emc_hip_cnn_get(E.fptr(p), None) == 0
    idx = np.arange(3, 3 + B, dtype=np.int32)
    logits = np.zeros((ncell, B), np.float32)
    assert L.emc_hip_cnn_forward(d_img.vp, d_mask.vp, 256, E.iptr(idx), B, E.fptr(logits)) == 0
    x = O.net_input(P.images, P.mask, idx, 16)
    ref_bf, _ = O.forward(p.astype(np.float64), lay, x, bf=True)
    ref, _ = O.forward(p.astype(np.float64), lay, x, bf=False)
    print("logits rel err vs bf16 oracle", rel(logits, ref_bf), "vs fp64", rel(logits, ref))
    assert rel(logits, ref_bf) < 1e-5  # same BF16 operands; FP32 vs FP64 accumulation
    assert rel(logits, ref) < 3e-2


@pytest.mark.gpu
def test_gpu_gradients_match_oracle(net_setup):
    P, d_img, d_mask = net_setup
    L = E.lib()
    ncell, B = 11, 8
    assert L.emc_hip_cnn_init(16, ncell, 16, 6) == 0
    lay = O.Layout(16, ncell)
    p = np.zeros(lay.n, np.float32)
    L.emc_hip_cnn_get(E.fptr(p), None)
    idx = np.arange(B, dtype=np.int32)
    rng = np.random.default_rng(3)
    t = rng.random((ncell, B)).astype(np.float32)
    t /= t.sum(axis=0)
    loss = C.c_double()
    assert L.emc_hip_cnn_step(d_img.vp, d_mask.vp, 256, E.iptr(idx), B, E.fptr(t), 0.0, C.byref(loss)) == 0
    E.check()
    g = np.zeros(lay.n, np.float32)
    p2 = np.zeros(lay.n, np.float32)
    L.emc_hip_cnn_get(E.fptr(p2), E.fptr(g))
    assert np.array_equal(p, p2)  # lr = 0 leaves the weights
    x = O.net_input(P.images, P.mask, idx, 16)
    l_ref, g_ref, _ = O.loss_and_grad(p.astype(np.float64), lay, x, t.astype(np.float64), bf=True)
    assert abs(loss.value - l_ref) < 1e-5 * abs(l_ref)
    for k, (w0, b0, (r, c)) in enumerate(lay.off):
        ew = rel(g[w0:w0 + r * c], g_ref[w0:w0 + r * c])
        eb = rel(g[b0:b0 + r], g_ref[b0:b0 + r])
        print(f"layer {k}: dW rel {ew:.2e}, db rel {eb:.2e}")
        assert ew < 1e-5 and eb < 1e-5


@pytest.mark.gpu
def test_gpu_training_reduces_loss(net_setup):
    P, d_img, d_mask = net_setup
    L = E.lib()
    ncell, B = 11, 16
    assert L.emc_hip_cnn_init(16, ncell, 16, 7) == 0
    idx = np.arange(B, dtype=np.int32)
    t = np.zeros((ncell, B), np.float32)
    t[np.arange(B) % ncell, np.arange(B)] = 1.0
    losses = []
    for _ in range(60):
        loss = C.c_double()
        assert L.emc_hip_cnn_step(d_img.vp, d_mask.vp, 256, E.iptr(idx), B, E.fptr(t), 3e-3, C.byref(loss)) == 0
        losses.append(loss.value)
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.3 * losses[0]
