"""Python bindings of libemc_hip.so, the MI355X EMC C-ABI.

The product is the C-ABI shared library (include/emc.h, emc_hip.h,
emc_host.h) and the C driver; this module only loads the library with ctypes
and gives tests and the benchmark typed access to the same entry points a C
caller uses.  There is no Python or CPU fallback: if the library has not been
built, importing `lib()` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemc_hip.so")

c_float_p = C.POINTER(C.c_float)
c_int_p = C.POINTER(C.c_int)
c_void_pp = C.POINTER(C.c_void_p)


class Setup(C.Structure):
    """emc.h:12-18"""
    _fields_ = [("side", C.c_int), ("wavelength", C.c_float), ("pixel_size", C.c_float),
                ("detector_size", C.c_int), ("detector_distance", C.c_float)]


class SpMatrix(C.Structure):
    _fields_ = [("rows", C.c_uint), ("cols", C.c_uint), ("data", c_float_p)]


class SpIMatrix(C.Structure):
    _fields_ = [("rows", C.c_uint), ("cols", C.c_uint), ("data", c_int_p)]


class Sp3Matrix(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint), ("data", c_float_p)]


class H5Dataset(C.Structure):
    """emc_h5_dataset (emc_io.h)."""
    _fields_ = [("name", C.c_char_p), ("type", C.c_int), ("rank", C.c_int), ("dims", C.c_longlong * 3),
                ("data", C.c_void_p)]


class Quaternion(C.Structure):
    _fields_ = [("q", C.c_float * 4)]


class SynthParams(C.Structure):
    """emc_host.h emc_synth_params"""
    _fields_ = [("side", C.c_int), ("n_blobs", C.c_int), ("blob_radius", C.c_float),
                ("sigma_min", C.c_float), ("sigma_max", C.c_float), ("n_patterns", C.c_int),
                ("phi_sigma", C.c_float), ("photons", C.c_float), ("seed", C.c_uint64)]


# The reference's Setup constants (emc.c:766-771).
def reference_setup(side: int) -> Setup:
    return Setup(side=side, wavelength=5.6e-9, pixel_size=16e-6, detector_size=512,
                 detector_distance=0.15)


_LIB: Optional[C.CDLL] = None

# (name, restype, argtypes) of every exported function, as declared in
# include/emc.h, include/emc_hip.h and include/emc_host.h.
_SIGS = [
    # emc.h
    ("insert_slice", None, [C.POINTER(Sp3Matrix), C.POINTER(Sp3Matrix), C.POINTER(SpMatrix),
                            C.POINTER(SpIMatrix), C.c_float, C.POINTER(Quaternion), C.POINTER(SpMatrix),
                            C.POINTER(SpMatrix), C.POINTER(SpMatrix)]),
    ("cuda_update_slices", C.c_float, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                       C.c_int, C.c_int, C.POINTER(Sp3Matrix), C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, c_float_p, C.c_void_p, Setup,
                                       C.c_void_p]),
    ("cuda_update_scaling", None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                   C.c_int, C.c_int, c_float_p]),
    ("cuda_calculate_responsabilities", None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p,
                                               C.c_void_p, C.c_int, C.c_int, C.c_int, c_float_p]),
    ("cuda_get_slices", None, [C.POINTER(Sp3Matrix), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_int]),
    ("cuda_allocate_slices", None, [c_void_pp, Setup, C.c_int]),
    ("cuda_model_max", C.c_float, [C.c_void_p, C.c_int]),
    ("cuda_allocate_model", None, [c_void_pp, C.POINTER(Sp3Matrix)]),
    ("cuda_allocate_mask", None, [c_void_pp, C.POINTER(SpIMatrix)]),
    ("cuda_reset_model", None, [C.POINTER(Sp3Matrix), C.c_void_p]),
    ("cuda_normalize_model", None, [C.POINTER(Sp3Matrix), C.c_void_p, C.c_void_p]),
    ("cuda_allocate_rotations", None, [c_void_pp, C.POINTER(C.POINTER(Quaternion)), C.c_int]),
    ("cuda_allocate_images", None, [c_void_pp, C.POINTER(C.POINTER(SpMatrix)), C.c_int]),
    ("cuda_allocate_coords", None, [c_void_pp, c_void_pp, c_void_pp, C.POINTER(SpMatrix),
                                    C.POINTER(SpMatrix), C.POINTER(SpMatrix)]),
    ("cuda_allocate_real", None, [c_void_pp, C.c_int]),
    ("cuda_allocate_scaling", None, [c_void_pp, C.c_int]),
    ("cuda_normalize_responsabilities", None, [C.c_void_p, C.c_int, C.c_int]),
    ("cuda_total_respons", C.c_float, [C.c_void_p, c_float_p, C.c_int]),
    # emc_hip.h
    ("emc_hip_last_error", C.c_int, [C.c_char_p, C.c_size_t]),
    ("emc_hip_clear_error", None, []),
    ("emc_hip_free", None, [C.c_void_p]),
    ("emc_hip_synchronize", None, []),
    ("emc_hip_shutdown", None, []),
    ("emc_hip_total_respons_n", C.c_float, [C.c_void_p, C.c_size_t]),
    ("emc_hip_normalize_model_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("emc_hip_respons_summary", C.c_int, [C.c_void_p, C.c_int, C.c_int, c_int_p, c_float_p]),
    ("emc_hip_allocate_bytes", None, [c_void_pp, C.c_size_t]),
    ("emc_hip_host_alloc", C.c_int, [c_void_pp, C.c_size_t]),
    ("emc_hip_host_free", None, [C.c_void_p]),
    ("emc_hip_memcpy_to_host", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("emc_hip_memcpy_to_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("emc_hip_set_option", C.c_int, [C.c_char_p, C.c_char_p]),
    ("emc_hip_invalidate", None, []),
    ("emc_hip_phase_times", C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_longlong), C.c_int]),
    ("emc_hip_reset_phase_times", None, []),
    ("emc_hip_phase_name", C.c_char_p, [C.c_int]),
    ("emc_hip_pack_info", C.c_int, [c_int_p, C.c_int]),
    ("emc_hip_mstep_nnz", C.c_longlong, []),
    ("emc_hip_refine_count", C.c_longlong, []),
    ("emc_hip_cnn_init", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_ulonglong]),
    ("emc_hip_cnn_param_count", C.c_long, []),
    ("emc_hip_cnn_get", C.c_int, [c_float_p, c_float_p]),
    ("emc_hip_cnn_set", C.c_int, [c_float_p]),
    ("emc_hip_cnn_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, c_int_p, C.c_int, c_float_p]),
    ("emc_hip_cnn_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, c_int_p, C.c_int, c_float_p, C.c_float,
                                   C.POINTER(C.c_double)]),
    ("emc_hip_proposal_init", C.c_int, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong]),
    ("emc_hip_proposal_cells", C.c_int, [c_int_p, c_int_p]),
    ("emc_hip_proposal_train", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                         C.c_float, C.POINTER(C.c_double)]),
    ("emc_hip_proposal_predict", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
    ("emc_hip_proposal_set_candidates", C.c_int, [c_int_p, C.c_int, C.c_int]),
    ("emc_hip_proposal_get", C.c_int, [c_int_p, c_int_p, c_float_p, C.c_int]),
    ("emc_hip_proposal_stats", C.c_int, [C.POINTER(C.c_double), C.c_int]),
    ("emc_hip_proposal_shared", C.c_int, [c_int_p, C.c_int]),
    ("emc_hip_images_from_csr", C.c_int, [C.c_void_p, c_int_p, c_int_p, C.POINTER(C.c_ushort), C.c_int, C.c_int]),
    # emc_host.h
    ("emc_n_to_samples", C.c_int, [C.c_int]),
    ("emc_rotation_table", C.c_int, [C.c_int, C.c_int, c_float_p, c_float_p]),
    ("emc_assign_cells", C.c_int, [c_float_p, C.c_int, c_float_p, C.c_int, c_int_p]),
    ("generate_rotation_list", C.c_int, [C.c_int, C.c_void_p, C.c_void_p]),
    ("emc_coordinates", None, [Setup, c_float_p, c_float_p, c_float_p]),
    ("calculate_coordinates", None, [Setup, C.POINTER(SpMatrix), C.POINTER(SpMatrix), C.POINTER(SpMatrix)]),
    ("emc_make_mask", C.c_int, [C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, c_int_p]),
    ("emc_synth_intensity", C.c_int, [C.POINTER(SynthParams), c_float_p]),
    ("emc_synth_patterns", C.c_int, [C.POINTER(SynthParams), c_float_p, c_float_p, c_float_p, c_float_p,
                                     c_int_p, c_float_p, c_float_p, c_float_p]),
    ("emc_synth_random_model", None, [C.c_int, C.c_uint64, c_float_p]),
    ("emc_quaternion_from_uniforms", None, [C.c_double, C.c_double, C.c_double, c_float_p]),
    ("emc_uniform", C.c_double, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]),
    # emc_io.h
    ("emc_h5_write", C.c_int, [C.c_char_p, C.c_void_p, C.c_int]),
    ("emc_h5_read", C.c_longlong, [C.c_char_p, C.c_char_p, C.c_int, C.c_void_p, C.c_longlong, c_int_p,
                                   C.POINTER(C.c_longlong)]),
    ("emc_io_error", C.c_char_p, []),
    ("emc_write_spimage", C.c_int, [C.c_char_p, c_float_p, c_int_p, C.c_int, C.c_int, C.c_int]),
    ("emc_read_spimage", C.c_int, [C.c_char_p, c_float_p, c_int_p, C.c_int, C.c_int, C.c_int]),
    ("emc_spimage_shape", C.c_int, [C.c_char_p, c_int_p, c_int_p, c_int_p]),
    ("emc_read_patterns", C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, c_float_p]),
    ("emc_read_mask", C.c_int, [C.c_char_p, C.c_int, C.c_int, c_int_p]),
]

EXPORTED = [s[0] for s in _SIGS]


def lib() -> C.CDLL:
    """Load libemc_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C emc_amd` "
                               "(or __graft_entry__.build()); there is no fallback path")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def fptr(a: np.ndarray) -> c_float_p:
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_float_p)


def iptr(a: np.ndarray) -> c_int_p:
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_int_p)


def check() -> None:
    """Raise if the library recorded an error."""
    buf = C.create_string_buffer(1024)
    if lib().emc_hip_last_error(buf, 1024):
        msg = buf.value.decode()
        lib().emc_hip_clear_error()
        raise RuntimeError(f"emc_hip: {msg}")


def set_option(key: str, value: str) -> None:
    if lib().emc_hip_set_option(key.encode(), value.encode()) != 0:
        raise ValueError(f"option {key}={value} rejected")


# ---- host maths -------------------------------------------------------------
def rotation_table(n: int, half: bool = True):
    L = lib()
    ns = L.emc_n_to_samples(n)
    q = np.zeros((ns, 4), np.float32)
    w = np.zeros(ns, np.float32)
    m = L.emc_rotation_table(n, 1 if half else 0, fptr(q), fptr(w))
    if m < 0:
        raise RuntimeError("rotation table construction failed")
    return np.ascontiguousarray(q[:m]), np.ascontiguousarray(w[:m])


def coordinates(side: int, setup: Optional[Setup] = None):
    s = setup or reference_setup(side)
    n2 = side * side
    x, y, z = (np.zeros(n2, np.float32) for _ in range(3))
    lib().emc_coordinates(s, fptr(x), fptr(y), fptr(z))
    return x, y, z


def make_mask(side: int, beamstop: float, gap_h: int = 0, gap_cross: int = 0, circle: bool = False):
    m = np.zeros(side * side, np.int32)
    lib().emc_make_mask(side, beamstop, gap_h, gap_cross, 1 if circle else 0, iptr(m))
    return m


def synth_params(side, n_blobs, blob_radius, n_patterns, phi_sigma, photons, seed=0):
    return SynthParams(side=side, n_blobs=n_blobs, blob_radius=blob_radius, sigma_min=1.0, sigma_max=2.0,
                       n_patterns=n_patterns, phi_sigma=phi_sigma, photons=photons, seed=seed)


def synth_intensity(p: SynthParams) -> np.ndarray:
    v = np.zeros(p.side ** 3, np.float32)
    if lib().emc_synth_intensity(C.byref(p), fptr(v)) != 0:
        raise RuntimeError("emc_synth_intensity failed")
    return v


def synth_patterns(p: SynthParams, volume, x, y, z, mask, out: Optional[np.ndarray] = None):
    n2 = p.side * p.side
    images = out if out is not None else np.zeros((p.n_patterns, n2), np.float32)
    q = np.zeros((p.n_patterns, 4), np.float32)
    phi = np.zeros(p.n_patterns, np.float32)
    if lib().emc_synth_patterns(C.byref(p), fptr(volume), fptr(x), fptr(y), fptr(z), iptr(mask),
                                fptr(images), fptr(q), fptr(phi)) != 0:
        raise RuntimeError("emc_synth_patterns failed")
    return images, q, phi


def random_model(side: int, seed: int = 0) -> np.ndarray:
    m = np.zeros(side ** 3, np.float32)
    lib().emc_synth_random_model(side, seed, fptr(m))
    return m


# ---- device buffers ------------------------------------------------------------
class DeviceArray:
    """A device buffer allocated by the library (hipMalloc underneath)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        lib().emc_hip_allocate_bytes(C.byref(p), C.c_size_t(nbytes))
        check()
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if lib().emc_hip_memcpy_to_device(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes):
            check()

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        if lib().emc_hip_memcpy_to_host(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes):
            check()
        return out

    def free(self) -> None:
        if self.ptr:
            lib().emc_hip_free(C.c_void_p(self.ptr))
            self.ptr = None

    @property
    def vp(self) -> C.c_void_p:
        return C.c_void_p(self.ptr)


def sp3(side: int, data: np.ndarray) -> Sp3Matrix:
    assert data.dtype == np.float32 and data.size == side ** 3
    return Sp3Matrix(x=side, y=side, z=side, data=fptr(data))


class EmcProblem:
    """Device-resident inputs of one EMC run, set up through the C-ABI the way
    the reference driver does (emc.c:961-985)."""

    def __init__(self, side: int, images: np.ndarray, mask: np.ndarray, q: np.ndarray, w: np.ndarray,
                 model: np.ndarray, x=None, y=None, z=None, scaling: Optional[np.ndarray] = None,
                 keep_images_host: bool = False):
        L = lib()
        self.side = side
        self.N = int(images.shape[0])
        self.R = int(q.shape[0])
        self.N2d = side * side
        # the model array cuda_update_slices copies back every iteration, in
        # page-locked memory (copied at the link's rate)
        self._model_pin = C.c_void_p()
        nb = 4 * side ** 3
        if L.emc_hip_host_alloc(C.byref(self._model_pin), C.c_size_t(nb)) == 0 and self._model_pin.value:
            buf = (C.c_float * side ** 3).from_address(self._model_pin.value)
            self.model_host = np.ctypeslib.as_array(buf)
            self.model_host[:] = np.ascontiguousarray(model, np.float32).reshape(-1)
        else:
            L.emc_hip_clear_error()
            self._model_pin = C.c_void_p()
            self.model_host = np.ascontiguousarray(model, np.float32).reshape(-1).copy()
        self.weights = np.ascontiguousarray(w, np.float32)
        self.q = np.ascontiguousarray(q, np.float32)
        if x is None:
            x, y, z = coordinates(side)
        self.x, self.y, self.z = x, y, z
        self.d_images = DeviceArray.from_host(np.ascontiguousarray(images, np.float32).reshape(-1))
        self.d_mask = DeviceArray.from_host(np.ascontiguousarray(mask, np.int32))
        self.d_rot = DeviceArray.from_host(self.q.reshape(-1))
        self.d_x = DeviceArray.from_host(x)
        self.d_y = DeviceArray.from_host(y)
        self.d_z = DeviceArray.from_host(z)
        self.d_model = DeviceArray.from_host(self.model_host)
        self.d_weight = DeviceArray.from_host(np.zeros(side ** 3, np.float32))
        self.d_slices = DeviceArray(4 * self.R * self.N2d)
        self.d_respons = DeviceArray(4 * self.R * self.N)
        ph = np.ones(self.N, np.float32) if scaling is None else np.ascontiguousarray(scaling, np.float32)
        self.d_scaling = DeviceArray.from_host(ph)
        self.scaling_host = ph.copy()
        self.sm = sp3(side, self.model_host)
        self.setup = reference_setup(side)
        check()

    def free(self):
        for d in (self.d_images, self.d_mask, self.d_rot, self.d_x, self.d_y, self.d_z, self.d_model,
                  self.d_weight, self.d_slices, self.d_respons, self.d_scaling):
            d.free()
        if self._model_pin.value:
            self.model_host = None
            self.sm = None
            lib().emc_hip_host_free(self._model_pin)
            self._model_pin = C.c_void_p()

    # ---- the reference's per-iteration calls (emc.c:989-1084) ----
    def get_slices(self):
        lib().cuda_get_slices(C.byref(self.sm), self.d_model.vp, self.d_slices.vp, self.d_rot.vp, self.d_x.vp,
                              self.d_y.vp, self.d_z.vp, self.R)
        check()

    def likelihood(self, sigma: float, host_respons: Optional[np.ndarray] = None):
        lib().cuda_calculate_responsabilities(self.d_slices.vp, self.d_images.vp, self.d_mask.vp, sigma,
                                              self.d_scaling.vp, self.d_respons.vp, self.N2d, self.N, self.R,
                                              fptr(host_respons) if host_respons is not None else None)
        check()

    def normalize_respons(self):
        lib().cuda_normalize_responsabilities(self.d_respons.vp, self.R, self.N)
        check()

    def total_respons(self) -> float:
        n = self.N * self.R
        if n < 2 ** 31:
            v = lib().cuda_total_respons(self.d_respons.vp, None, n)
        else:  # the reference's int argument cannot hold it
            v = lib().emc_hip_total_respons_n(self.d_respons.vp, n)
        check()
        return v

    def reset_model(self):
        lib().cuda_reset_model(C.byref(self.sm), self.d_model.vp)
        lib().cuda_reset_model(C.byref(self.sm), self.d_weight.vp)
        check()

    def update_scaling(self):
        lib().cuda_update_scaling(self.d_images.vp, self.d_slices.vp, self.d_mask.vp, self.d_respons.vp,
                                  self.d_scaling.vp, self.N, self.R, self.N2d, fptr(self.scaling_host))
        check()

    def update_slices(self) -> float:
        v = lib().cuda_update_slices(self.d_images.vp, self.d_slices.vp, self.d_mask.vp, self.d_respons.vp,
                                     self.d_scaling.vp, self.N, self.R, self.N2d, C.byref(self.sm),
                                     self.d_model.vp, self.d_x.vp, self.d_y.vp, self.d_z.vp, self.d_rot.vp,
                                     fptr(self.weights), self.d_weight.vp, self.setup, None)
        check()
        return v

    def normalize_model(self):
        lib().cuda_normalize_model(C.byref(self.sm), self.d_model.vp, self.d_weight.vp)
        check()

    def model_max(self) -> float:
        v = lib().cuda_model_max(self.d_model.vp, self.side ** 3)
        check()
        return v

    def iteration(self, sigma: float = 0.4) -> dict:
        """One EMC iteration in the driver's order (emc.c:989-1084)."""
        mx = self.model_max()
        self.get_slices()
        self.likelihood(sigma)
        self.normalize_respons()
        tot = self.total_respons()
        self.reset_model()
        self.update_scaling()
        overall = self.update_slices()
        self.normalize_model()
        return {"model_max": mx, "likelihood": tot, "overall_respons": overall}

    # ---- downloads ----
    def slices(self) -> np.ndarray:
        return self.d_slices.download(np.float32, self.R * self.N2d).reshape(self.R, self.N2d)

    def respons(self) -> np.ndarray:
        return self.d_respons.download(np.float32, self.R * self.N).reshape(self.R, self.N)

    def model(self) -> np.ndarray:
        return self.d_model.download(np.float32, self.side ** 3)

    def weight(self) -> np.ndarray:
        return self.d_weight.download(np.float32, self.side ** 3)

    def scaling(self) -> np.ndarray:
        return self.d_scaling.download(np.float32, self.N)
