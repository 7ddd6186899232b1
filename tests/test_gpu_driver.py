"""The C driver (emc.c's main() over the C-ABI) runs end to end and agrees
with the same sequence of entry points driven from Python."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import emc_amd as E
from emc_amd import configs

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(E.LIB_PATH), "emc")


def read_volume(path, side, name="real"):
    L = E.lib()
    out = np.zeros(side ** 3, np.float32)
    assert L.emc_h5_read(str(path).encode(), name.encode(), 0, out.ctypes.data_as(C.c_void_p), out.size, None,
                         None) == out.size, L.emc_io_error()
    return out


def read_h5(path, name, dtype=np.float32):
    L = E.lib()
    rank = C.c_int()
    dims = (C.c_longlong * 3)()
    n = L.emc_h5_read(str(path).encode(), name.encode(), 0, None, 0, C.byref(rank), dims)
    assert n >= 0, L.emc_io_error()
    out = np.zeros(n, np.float32)
    L.emc_h5_read(str(path).encode(), name.encode(), 0, out.ctypes.data_as(C.c_void_p), n, None, None)
    return out.reshape([dims[k] for k in range(rank.value)]).astype(dtype)


def test_driver_config0_deterministic(tmp_path):
    out = tmp_path / "run"
    r = subprocess.run([EXE, "--config", "0", "--patterns", "300", "--iterations", "3", "--insert", "deterministic",
                        "--out", str(out), "--quiet"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    lik = np.loadtxt(out / "likelihood.data")
    err = np.loadtxt(out / "scale_error.data")
    assert lik.shape == (3,) and err.shape == (3,)
    assert np.isfinite(lik).all() and np.isfinite(err).all()
    model = read_volume(out / "model_final.h5", 32)
    assert model.size == 32 ** 3 and abs(model.astype(np.float64).mean() - 1) < 1e-5
    for it in range(3):  # the reference writes a model file every iteration (emc.c:1081-1082)
        assert read_volume(out / f"model_{it:04d}.h5", 32).size == 32 ** 3
    assert (out / "model_init.h5").exists() and (out / "model_init_weight.h5").exists()
    # the same calls from Python give the identical model (deterministic insertion)
    W = configs.build(0, n_patterns=300)
    E.set_option("precision", "fp32")
    E.set_option("insert", "deterministic")
    prob = E.EmcProblem(32, W.images, W.mask, W.q, W.w, W.model0)
    liks = [prob.iteration()["likelihood"] for _ in range(3)]
    assert np.array_equal(prob.model(), model)
    assert np.allclose(liks, lik, rtol=1e-5)  # the file holds %g (6 digits)
    prob.free()
    E.set_option("insert", "atomic")


def test_driver_reads_hdf5_data_and_writes_viewer_outputs(tmp_path):
    """emc.c's file path end to end: patterns and mask read from HDF5 image
    files (libspimage layout) with the read stride and circular crop, the
    model files of every iteration, and the viewer's outputs; the final model
    equals the same calls driven from Python on the same data."""
    W = configs.build(0, n_patterns=120)
    side = 32
    L = E.lib()
    data = tmp_path / "Data"
    data.mkdir()
    for i in range(W.N):
        img = np.ascontiguousarray(W.images[i])
        assert L.emc_write_spimage(str(data / f"image_{i:04d}.h5").encode(), E.fptr(img), None, side, side, 1) == 0
    m = W.mask.astype(np.float32)
    assert L.emc_write_spimage(str(tmp_path / "mask.h5").encode(), E.fptr(m), None, side, side, 1) == 0
    out = tmp_path / "run"
    r = subprocess.run([EXE, "--data", str(data / "image_%04d.h5"), "--mask", str(tmp_path / "mask.h5"),
                        "--read-stride", "1", "--side", str(side), "--patterns", str(W.N), "--iterations", "2",
                        "--sampling", "4", "--insert", "deterministic", "--viewer-output", "--out", str(out),
                        "--quiet"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    mask = np.zeros(side * side, np.int32)
    assert L.emc_read_mask(str(tmp_path / "mask.h5").encode(), side, 1, E.iptr(mask)) == 0
    assert mask.sum() <= W.mask.sum()  # the circular crop only removes pixels
    R = W.R
    assert read_h5(out / "state.h5", "iteration", np.int32).reshape(-1)[0] == 1
    bs = read_h5(out / "best_scaling.h5", "scaling")
    assert bs.shape == (2, W.N)
    avg = read_h5(out / "average_resp_0001.h5", "data")
    assert avg.shape == (R,) and abs(float(avg.astype(np.float64).sum()) - 1.0) < 1e-4
    bq = np.loadtxt(out / "best_quaternion_0001.data")
    assert bq.shape == (W.N, 4) and np.allclose(np.linalg.norm(bq, axis=1), 1.0, atol=1e-4)
    resp = read_h5(out / "responsabilities_0001.h5", "data")
    assert resp.shape == (W.N, R) and np.allclose(resp.astype(np.float64).sum(axis=1), 1.0, atol=1e-4)
    assert read_h5(out / "scaling_0001.h5", "data").shape == (R, W.N)
    final = read_volume(out / "model_final.h5", side)
    # the same calls from Python on the same (read) data
    images = W.images * mask[None, :].astype(np.float32)
    E.set_option("precision", "fp32")
    E.set_option("insert", "deterministic")
    prob = E.EmcProblem(side, np.ascontiguousarray(images), mask, W.q, W.w, E.random_model(side, 12345))
    for _ in range(2):
        prob.iteration()
    py = prob.model()
    best = np.zeros(W.N, np.int32)
    assert L.emc_hip_respons_summary(prob.d_respons.vp, R, W.N, E.iptr(best), None) == 0
    prob.free()
    E.set_option("insert", "atomic")
    assert np.array_equal(final, py)
    assert np.allclose(bq, W.q[best], atol=1e-6)
    assert np.allclose(resp.argmax(axis=1), best)
    # the rotation table written as the viewer's rotations file and read back
    # through --rotations-file: the same run, the same bits
    rot = read_h5(out / "rotations.h5", "rotations")
    assert rot.shape == (R, 5)
    assert np.array_equal(rot[:, :4], W.q) and np.array_equal(rot[:, 4], W.w.astype(np.float32))
    out2 = tmp_path / "run2"
    r = subprocess.run([EXE, "--data", str(data / "image_%04d.h5"), "--mask", str(tmp_path / "mask.h5"),
                        "--read-stride", "1", "--side", str(side), "--patterns", str(W.N), "--iterations", "2",
                        "--rotations-file", str(out / "rotations.h5"), "--insert", "deterministic", "--out",
                        str(out2), "--quiet"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-500:], r.stderr[-500:])
    assert r.returncode == 0
    assert np.array_equal(read_volume(out2 / "model_final.h5", side), final)
