"""The allocation-free normalisation (emc_hip_normalize_model_noalloc)
against cuda_normalize_model, through the C-ABI: two calls in a row, the
second on a smaller model, each bit for bit equal to the allocating entry
point's result on the same input."""
import numpy as np
import pytest

import emc_amd as E

pytestmark = pytest.mark.gpu


def normalized(side, model, weight, noalloc):
    L = E.lib()
    n = side ** 3
    dm = E.DeviceArray.from_host(model)
    dw = E.DeviceArray.from_host(weight)
    if noalloc:
        assert L.emc_hip_normalize_model_noalloc(dm.vp, dw.vp, n) == 0
    else:
        host = model.copy()
        L.cuda_normalize_model(E.C.byref(E.sp3(side, host)), dm.vp, dw.vp)
    E.check()
    out = dm.download(np.float32, n)
    dm.free()
    dw.free()
    return out


def test_noalloc_normalize_matches_cuda_normalize_model():
    rng = np.random.default_rng(7)
    for side in (64, 24):  # the second call on a smaller model
        n = side ** 3
        model = (rng.random(n) * 10).astype(np.float32)
        weight = rng.random(n).astype(np.float32)
        weight[rng.random(n) < 0.2] = 0.0  # voxels no slice reached
        ref = normalized(side, model, weight, noalloc=False)
        got = normalized(side, model, weight, noalloc=True)
        print(f"side {side}: {n} voxels, mean {got.mean():.6f}, max |diff| {np.abs(got - ref).max()}")
        assert np.array_equal(got, ref)
        assert got[weight == 0].max() == 0.0 and abs(float(got.mean()) - 1.0) < 1e-4


def test_noalloc_normalize_rejects_bad_arguments():
    L = E.lib()
    L.emc_hip_clear_error()
    assert L.emc_hip_normalize_model_noalloc(None, None, 8) == -1
    buf = E.C.create_string_buffer(256)
    assert L.emc_hip_last_error(buf, 256) != 0 and b"noalloc" in buf.value
    L.emc_hip_clear_error()
    d = E.DeviceArray(64)
    assert L.emc_hip_normalize_model_noalloc(d.vp, d.vp, -1) == -1
    L.emc_hip_clear_error()
    d.free()
