"""The C driver (emc.c's main() over the C-ABI) runs end to end and agrees
with the same sequence of entry points driven from Python."""
import os
import subprocess

import numpy as np
import pytest

import emc_amd as E
from emc_amd import configs

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(E.LIB_PATH), "emc")


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
    model = np.fromfile(out / "model_final.raw", np.float32)
    assert model.size == 32 ** 3 and abs(model.astype(np.float64).mean() - 1) < 1e-5
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
