"""The driver's HDF5 files (emc_io.h), on the CPU.

The writer's output is checked by a second, independent reading of the
HDF5 format written here in Python from the format specification
(superblock v0, symbol-table group, v1 B-tree, local heap, version-1 object
headers): every structural field and every value, plus the writer/reader
round trip of the C library, the libspimage layout (emc.c:804,840,917,944,
1082) and the error paths.  libhdf5 and h5py are absent, so whether the
files open in them is not tested here ("parity unpinned" for that)."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

import emc_amd as E

F32, I32, F64 = 0, 1, 2
UNDEF = 0xFFFFFFFFFFFFFFFF


def write(path, sets):
    arr = (E.H5Dataset * len(sets))()
    keep = []
    for k, (name, a) in enumerate(sets):
        a = np.ascontiguousarray(a)
        keep.append(a)
        t = {np.dtype(np.float32): F32, np.dtype(np.int32): I32, np.dtype(np.float64): F64}[a.dtype]
        dims = (C.c_longlong * 3)(*(list(a.shape) + [0, 0, 0])[:3])
        arr[k] = E.H5Dataset(name.encode(), t, a.ndim, dims, a.ctypes.data_as(C.c_void_p))
    return E.lib().emc_h5_write(str(path).encode(), arr, len(sets))


def read(path, name, dtype):
    t = {np.float32: F32, np.int32: I32, np.float64: F64}[dtype]
    rank = C.c_int()
    dims = (C.c_longlong * 3)()
    n = E.lib().emc_h5_read(str(path).encode(), name.encode(), t, None, 0, C.byref(rank), dims)
    assert n >= 0, E.lib().emc_io_error()
    out = np.zeros(n, dtype)
    assert E.lib().emc_h5_read(str(path).encode(), name.encode(), t, out.ctypes.data_as(C.c_void_p), n, None,
                               None) == n
    return out.reshape([dims[k] for k in range(rank.value)])


# ---- an independent reading of the format, from the specification ----
class Spec:
    def __init__(self, raw):
        self.b = raw

    def u(self, at, n):
        return int.from_bytes(self.b[at:at + n], "little")

    def superblock(self):
        b = self.b
        assert b[:8] == b"\x89HDF\r\n\x1a\n"
        assert b[8] == 0 and b[9] == 0 and b[10] == 0 and b[12] == 0  # versions
        assert b[13] == 8 and b[14] == 8  # sizes of offsets and lengths
        leaf_k, node_k = self.u(16, 2), self.u(18, 2)
        assert self.u(24, 8) == 0 and self.u(32, 8) == UNDEF and self.u(48, 8) == UNDEF
        assert self.u(40, 8) == len(b)  # end-of-file address
        root = 56
        assert self.u(root + 16, 4) == 1  # cache type: scratch holds the symbol table
        return leaf_k, node_k, self.u(root + 8, 8), self.u(root + 24, 8), self.u(root + 32, 8)

    def messages(self, oh):
        assert self.b[oh] == 1  # object header version 1
        n, size = self.u(oh + 2, 2), self.u(oh + 8, 4)
        pos, out = oh + 16, []
        for _ in range(n):
            t, s = self.u(pos, 2), self.u(pos + 2, 2)
            assert s % 8 == 0 and pos + 8 + s <= oh + 16 + size
            out.append((t, pos + 8, s))
            pos += 8 + s
        assert pos == oh + 16 + size
        return out

    def group(self):
        leaf_k, node_k, root_oh, btree, heap = self.superblock()
        msgs = self.messages(root_oh)
        assert [(t, s) for t, _, s in msgs] == [(0x11, 16)]
        assert (self.u(msgs[0][1], 8), self.u(msgs[0][1] + 8, 8)) == (btree, heap)
        assert self.b[heap:heap + 4] == b"HEAP"
        hsize, free, hdata = self.u(heap + 8, 8), self.u(heap + 16, 8), self.u(heap + 24, 8)
        assert free == UNDEF and self.b[hdata] == 0
        name = lambda off: self.b[hdata + off:self.b.index(b"\0", hdata + off)].decode()
        assert self.b[btree:btree + 4] == b"TREE" and self.b[btree + 4] == 0 and self.b[btree + 5] == 0
        used = self.u(btree + 6, 2)
        assert self.u(btree + 8, 8) == UNDEF and self.u(btree + 16, 8) == UNDEF
        entries, keys = [], [name(self.u(btree + 24, 8))]
        for c in range(used):
            child = self.u(btree + 32 + 16 * c, 8)
            keys.append(name(self.u(btree + 40 + 16 * c, 8)))
            assert self.b[child:child + 4] == b"SNOD" and self.b[child + 4] == 1
            ns = self.u(child + 6, 2)
            assert 0 < ns <= 2 * leaf_k
            node = [(name(self.u(child + 8 + 40 * k, 8)), self.u(child + 16 + 40 * k, 8)) for k in range(ns)]
            assert keys[-1] == node[-1][0]  # key after a child = its last name
            entries += node
        names = [n for n, _ in entries]
        assert keys[0] == "" and names == sorted(names) and len(set(names)) == len(names)
        return dict(entries)

    def dataset(self, oh):
        msgs = {t: (at, s) for t, at, s in self.messages(oh)}
        assert set(msgs) == {0x1, 0x3, 0x5, 0x8}
        at, _ = msgs[0x1]
        assert self.b[at] == 1
        rank = self.b[at + 1]
        dims = [self.u(at + 8 + 8 * k, 8) for k in range(rank)]
        at, _ = msgs[0x3]
        cls, size = self.b[at] & 15, self.u(at + 4, 4)
        assert self.b[at] >> 4 == 1
        if cls == 1:
            assert self.b[at + 1] == 0x20 and self.b[at + 2] == 8 * size - 1
            prec, eloc, esz, mloc, msz, bias = struct.unpack("<xxHBBBBI", self.b[at + 8:at + 20])
            assert (prec, eloc, esz, mloc, msz, bias) == ((32, 23, 8, 0, 23, 127) if size == 4 else
                                                          (64, 52, 11, 0, 52, 1023))
            dt = "<f4" if size == 4 else "<f8"
        else:
            assert cls == 0 and self.b[at + 1] == 0x08 and size == 4 and self.u(at + 10, 2) == 32
            dt = "<i4"
        at, _ = msgs[0x8]
        assert self.b[at] == 3 and self.b[at + 1] == 1
        addr, nbytes = self.u(at + 2, 8), self.u(at + 10, 8)
        assert nbytes == size * int(np.prod(dims)) and addr % 8 == 0 and addr + nbytes <= len(self.b)
        return np.frombuffer(self.b[addr:addr + nbytes], dt).reshape(dims)


def test_writer_layout_and_values_by_the_specification(tmp_path):
    rng = np.random.default_rng(0)
    sets = [("volume", rng.random((3, 4, 5)).astype(np.float32)),
            ("mask", rng.integers(-5, 5, (6, 7)).astype(np.int32)),
            ("curve", rng.random(11)),
            ("one", np.array([42], np.int32))]
    sets += [(f"extra_{k:02d}", np.full(k + 1, k, np.float32)) for k in range(12)]  # several symbol nodes
    path = tmp_path / "f.h5"
    assert write(path, sets) == 0
    spec = Spec(path.read_bytes())
    entries = spec.group()
    assert sorted(entries) == sorted(n for n, _ in sets)
    for name, a in sets:
        got = spec.dataset(entries[name])
        assert got.shape == a.shape and np.array_equal(got, a), name
        back = read(path, name, a.dtype.type)
        assert back.shape == a.shape and np.array_equal(back, a), name


def test_reader_converts_types(tmp_path):
    path = tmp_path / "t.h5"
    a = np.array([1.5, -2.25, 3e6], np.float32)
    b = np.array([7, -9], np.int32)
    assert write(path, [("a", a), ("b", b)]) == 0
    assert np.array_equal(read(path, "a", np.float64), a.astype(np.float64))
    assert np.array_equal(read(path, "b", np.float32), b.astype(np.float32))
    assert np.array_equal(read(path, "a", np.int32), a.astype(np.int32))


def test_errors(tmp_path):
    L = E.lib()
    path = tmp_path / "e.h5"
    assert write(path, [("x", np.zeros(4, np.float32))]) == 0
    assert L.emc_h5_read(str(path).encode(), b"y", F32, None, 0, None, None) == -1
    assert b"no dataset y" in L.emc_io_error()
    small = np.zeros(2, np.float32)
    assert L.emc_h5_read(str(path).encode(), b"x", F32, small.ctypes.data_as(C.c_void_p), 2, None, None) == -1
    bad = tmp_path / "bad.h5"
    bad.write_bytes(b"not hdf5 at all" * 10)
    assert L.emc_h5_read(str(bad).encode(), b"x", F32, None, 0, None, None) == -1
    assert write(path, [("x", np.zeros(1, np.float32)), ("x", np.zeros(1, np.float32))]) == -1
    assert write(path, [("a/b", np.zeros(1, np.float32))]) == -1


@pytest.mark.parametrize("shape", [(4, 6, 5), (1, 8, 8)])
def test_spimage_round_trip(tmp_path, shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(1)
    real = rng.random(nx * ny * nz).astype(np.float32)
    mask = (rng.random(nx * ny * nz) > 0.3).astype(np.int32)
    path = str(tmp_path / "img.h5").encode()
    L = E.lib()
    assert L.emc_write_spimage(path, E.fptr(real), E.iptr(mask), nx, ny, nz) == 0
    sx, sy, sz = C.c_int(), C.c_int(), C.c_int()
    assert L.emc_spimage_shape(path, C.byref(sx), C.byref(sy), C.byref(sz)) == 0
    assert (sx.value, sy.value, sz.value) == (nx, ny, nz)
    r2 = np.zeros_like(real)
    m2 = np.zeros_like(mask)
    assert L.emc_read_spimage(path, E.fptr(r2), E.iptr(m2), nx, ny, nz) == 0
    assert np.array_equal(r2, real) and np.array_equal(m2, mask)
    spec = Spec(open(path, "rb").read())
    entries = spec.group()
    assert {"real", "mask", "num_dimensions", "image_center", "phased", "shifted", "scaled", "detector_distance",
            "lambda", "pixel_size", "version"} == set(entries)
    dims = (nz, ny, nx) if nz > 1 else (ny, nx)
    assert spec.dataset(entries["real"]).shape == dims
    assert spec.dataset(entries["num_dimensions"])[0] == len(dims)


def test_spimage_modulus_of_complex(tmp_path):
    """A phased (complex) image is read as |real + i imag|, the modulus the
    reference takes with sp_cabs (emc.c:811-815)."""
    re_ = np.array([3.0, -1.0, 0.0, 2.0], np.float32)
    im = np.array([4.0, 0.0, -2.0, 0.0], np.float32)
    path = tmp_path / "c.h5"
    assert write(path, [("real", re_.reshape(2, 2)), ("imag", im.reshape(2, 2)),
                        ("mask", np.ones((2, 2), np.int32))]) == 0
    out = np.zeros(4, np.float32)
    assert E.lib().emc_read_spimage(str(path).encode(), E.fptr(out), None, 2, 2, 1) == 0
    assert np.allclose(out, [5.0, 1.0, 2.0, 2.0])


def test_pattern_and_mask_ingest(tmp_path):
    """emc.c:777-864 with read_stride 2 on 10x10 source images for a 4x4
    detector: pixel x comes from source column (int)(2 ((x - 2) + 0.5) + 5 -
    0.5) = (int)(2x + 1.5) (3, 5, 7 for x = 1, 2, 3), rows likewise; the mask
    is 0 where the source mask is 0 and outside radius 2 of (1.5, 1.5)."""
    L = E.lib()
    side, stride, n = 4, 2, 3
    rng = np.random.default_rng(5)
    src = []
    for i in range(n):
        img = rng.integers(0, 50, (10, 10)).astype(np.float32) * (1 if i % 2 else -1)  # |.| is taken
        src.append(img)
        assert L.emc_write_spimage(str(tmp_path / f"image_{i:04d}.h5").encode(), E.fptr(img.reshape(-1)), None,
                                   10, 10, 1) == 0
    out = np.zeros((n, side * side), np.float32)
    assert L.emc_read_patterns(str(tmp_path / "image_%04d.h5").encode(), n, side, stride, E.fptr(out)) == 0
    for i in range(n):
        for x in range(side):
            for y in range(side):
                sx = int(stride * ((x - side // 2) + 0.5) + 10 // 2 - 0.5)
                sy = int(stride * ((y - side // 2) + 0.5) + 10 // 2 - 0.5)
                assert out[i, y * side + x] == abs(src[i][sy, sx])
    m_src = np.ones((10, 10), np.float32)
    m_src[5, 5] = 0.0  # source (x=5, y=5) -> detector (x=2, y=2) at stride 2
    assert L.emc_write_spimage(str(tmp_path / "mask.h5").encode(), E.fptr(m_src.reshape(-1)), None, 10, 10, 1) == 0
    mask = np.zeros(side * side, np.int32)
    assert L.emc_read_mask(str(tmp_path / "mask.h5").encode(), side, stride, E.iptr(mask)) == 0
    expect = np.zeros(side * side, np.int32)
    for x in range(side):
        for y in range(side):
            inside = np.hypot(x - side / 2 + 0.5, y - side / 2 + 0.5) <= side / 2
            expect[y * side + x] = int(inside and not (x == 2 and y == 2))
    assert np.array_equal(mask, expect)
    assert L.emc_read_mask(None, side, stride, E.iptr(mask)) == 0
    expect[2 * side + 2] = 1
    assert np.array_equal(mask, expect)
    small = np.zeros(16, np.float32)
    assert L.emc_read_patterns(str(tmp_path / "image_%04d.h5").encode(), 1, 8, 2, E.fptr(np.zeros(64, np.float32))) \
        == -1  # an 8-pixel detector at stride 2 reaches outside a 10-pixel image
    del small
