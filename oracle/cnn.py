"""FP64 CPU oracle of the rotation-proposal network (test infrastructure only).

Same parameter layout and conventions as the HIP network: activations NHWC,
conv weights [OC][3][3][IC] (flattened with IC fastest), 3x3 / stride 2 /
pad 1 convolutions with ReLU (16, 32, 64, 64 channels), dense 256 + ReLU,
dense n_cells, softmax cross-entropy averaged over the batch.  With
bf16=True every GEMM operand is rounded to BF16 (round to nearest even)
exactly where the GPU rounds it; accumulation stays FP64.
"""
from __future__ import annotations

import numpy as np

IC = [1, 16, 32, 64]
OC = [16, 32, 64, 64]
HIDDEN = 256


def bf16(x):
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


class Layout:
    def __init__(self, side, ncell):
        self.side, self.ncell = side, ncell
        self.feat = OC[3] * (side // 16) ** 2
        self.off = []
        o = 0
        shapes = [(OC[l], 9 * IC[l]) for l in range(4)] + [(HIDDEN, self.feat), (ncell, HIDDEN)]
        for (r, c) in shapes:
            self.off.append((o, o + r * c, (r, c)))
            o += r * c + r
        self.n = o

    def unpack(self, p):
        Ws, bs = [], []
        for (w0, b0, (r, c)) in self.off:
            Ws.append(np.asarray(p[w0:w0 + r * c], np.float64).reshape(r, c))
            bs.append(np.asarray(p[b0:b0 + r], np.float64))
        return Ws, bs

    def pack(self, Ws, bs):
        p = np.zeros(self.n, np.float64)
        for (w0, b0, (r, c)), W, b in zip(self.off, Ws, bs):
            p[w0:w0 + r * c] = W.reshape(-1)
            p[b0:b0 + r] = b
        return p


def im2col(a):
    """a [B][H][W][C] -> col [B*OH*OW][9*C] with (ky, kx, c) ordering."""
    B, H, W, C = a.shape
    OH, OW = H // 2, W // 2
    pad = np.zeros((B, H + 2, W + 2, C), a.dtype)
    pad[:, 1:H + 1, 1:W + 1, :] = a
    cols = np.zeros((B, OH, OW, 3, 3, C), a.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky, kx, :] = pad[:, ky:ky + 2 * OH:2, kx:kx + 2 * OW:2, :]
    return cols.reshape(B * OH * OW, 9 * C)


def col2im(col, B, H, W, C):
    OH, OW = H // 2, W // 2
    cols = col.reshape(B, OH, OW, 3, 3, C)
    pad = np.zeros((B, H + 2, W + 2, C), col.dtype)
    for ky in range(3):
        for kx in range(3):
            pad[:, ky:ky + 2 * OH:2, kx:kx + 2 * OW:2, :] += cols[:, :, :, ky, kx, :]
    return pad[:, 1:H + 1, 1:W + 1, :]


def net_input(images, mask, idx, side):
    x = np.log1p(np.maximum(images[idx].astype(np.float64), 0)) * (mask != 0)
    return x.reshape(len(idx), side, side, 1)


def forward(p, lay, x, bf=False):
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    Ws, bs = lay.unpack(p)
    cache = {"col": [], "act": []}
    a = x
    for l in range(4):
        B, H, W, C = a.shape
        col = im2col(a)
        z = r(col) @ r(Ws[l]).T + bs[l]
        a = np.maximum(z, 0).reshape(B, H // 2, W // 2, OC[l])
        cache["col"].append(col)
        cache["act"].append(a)
    a3 = a.reshape(a.shape[0], -1)
    h = np.maximum(r(a3) @ r(Ws[4]).T + bs[4], 0)
    logits = r(Ws[5]) @ r(h).T + bs[5][:, None]  # [C][B]
    cache["a3"], cache["h"] = a3, h
    return logits, cache


def loss_and_grad(p, lay, x, targets, bf=False):
    """targets [C][B]; returns (mean loss, gradient vector, logits)."""
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    Ws, bs = lay.unpack(p)
    logits, cache = forward(p, lay, x, bf)
    B = x.shape[0]
    z = logits - logits.max(axis=0, keepdims=True)
    lp = z - np.log(np.exp(z).sum(axis=0, keepdims=True))
    loss = -(targets * lp).sum() / B
    dlog = (np.exp(lp) - targets) / B
    if bf:
        dlog = dlog.astype(np.float32).astype(np.float64)  # the GPU stores it in float
    dW, db = [None] * 6, [None] * 6
    h, a3 = cache["h"], cache["a3"]
    dW[5] = r(dlog) @ r(h)
    db[5] = dlog.sum(axis=1)
    dh = (r(dlog).T @ r(Ws[5])) * (h > 0)
    dW[4] = r(dh).T @ r(a3)
    db[4] = dh.sum(axis=0)
    da = (r(dh) @ r(Ws[4])).reshape(cache["act"][3].shape)
    for l in range(3, -1, -1):
        act = cache["act"][l]
        Bn, oh, ow, oc = act.shape
        dz = (da * (act > 0)).reshape(-1, oc)
        dW[l] = r(dz).T @ r(cache["col"][l])
        db[l] = dz.sum(axis=0)
        if l > 0:
            dcol = r(dz) @ r(Ws[l])
            if bf:
                dcol = dcol.astype(np.float32).astype(np.float64)
            H = oh * 2
            da = col2im(dcol, Bn, H, H, IC[l])
    return loss, lay.pack(dW, db), logits
