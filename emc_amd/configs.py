"""The five workloads of the benchmark plan (synthetic single-particle data).

Each builds the inputs of one EMC run with the product's host generator:
object of Gaussian blobs, patterns from uniformly random orientations with
lognormal scale factors, Poisson counts, a beamstop (and detector gaps),
the antipodally deduplicated rotation table, a uniform(0,1) initial model.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

import emc_amd as E


@dataclasses.dataclass(frozen=True)
class Config:
    name: str
    side: int
    n_blobs: int
    blob_radius: float
    n_patterns: int
    phi_sigma: float
    photons: float
    beamstop: float
    gap_h: int
    gap_cross: int
    n_rot: int          # sampling n of the rotation table (half table used)
    iterations: int
    precision: str = "fp32"
    # "poisson" where the workload's likelihood is the count x log-slice
    # contraction; "gaussian" is the reference's absolute-difference form
    likelihood: str = "gaussian"
    proposal: bool = False  # the learned rotation proposal after the training iterations
    csr: bool = False       # patterns held as CSR, sparse-count E-step


CONFIGS = {
    0: Config("s32-n1k-r3240", 32, 20, 6.0, 1000, 0.2, 2000.0, 3.0, 0, 0, 4, 10),
    1: Config("s64-n10k-r25680", 64, 40, 12.0, 10000, 0.3, 5000.0, 4.0, 2, 0, 8, 20),
    2: Config("s128-n50k-r50100", 128, 80, 24.0, 50000, 0.3, 10000.0, 6.0, 0, 3, 10, 20, "fp16", "poisson"),
    3: Config("s128-n50k-r50100-proposal", 128, 80, 24.0, 50000, 0.3, 10000.0, 6.0, 0, 3, 10, 20, "fp16", "poisson",
              True),
    4: Config("s256-n20k-r25680-sparse", 256, 150, 48.0, 20000, 0.5, 300.0, 12.0, 0, 8, 8, 30, "fp16", "poisson",
              True, True),
}


@dataclasses.dataclass
class Workload:
    cfg: Config
    images: np.ndarray
    mask: np.ndarray
    q: np.ndarray
    w: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    model0: np.ndarray
    truth: np.ndarray
    phi_true: np.ndarray
    q_true: np.ndarray

    @property
    def N(self):
        return self.images.shape[0]

    @property
    def R(self):
        return self.q.shape[0]

    @property
    def n_m(self):
        return int((self.mask != 0).sum())


def build(k: int, n_patterns: Optional[int] = None, seed: int = 0) -> Workload:
    c = CONFIGS[k]
    N = n_patterns or c.n_patterns
    x, y, z = E.coordinates(c.side)
    mask = E.make_mask(c.side, c.beamstop, c.gap_h, c.gap_cross)
    p = E.synth_params(c.side, c.n_blobs, c.blob_radius, N, c.phi_sigma, c.photons, seed)
    truth = E.synth_intensity(p)
    images, qt, phi = E.synth_patterns(p, truth, x, y, z, mask)
    q, w = E.rotation_table(c.n_rot, half=True)
    model0 = E.random_model(c.side, seed + 12345)
    return Workload(c, images, mask, q, w, x, y, z, model0, truth, phi, qt)
