"""The C-ABI library loads and exports every function the public headers
declare (no GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re

import pytest

import emc_amd as E

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")


def declared_functions():
    names = set()
    for h in ("emc.h", "emc_hip.h", "emc_host.h", "emc_io.h"):
        src = open(os.path.join(INCLUDE, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"static inline[^{]*\{[^}]*\}", "", src)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", src):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return sorted(names)


def test_headers_parse():
    names = declared_functions()
    for must in ("cuda_get_slices", "cuda_calculate_responsabilities", "cuda_update_slices",
                 "cuda_total_respons", "insert_slice", "emc_hip_last_error", "emc_rotation_table"):
        assert must in names


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    L = ctypes.CDLL(E.LIB_PATH)
    assert hasattr(L, name), name


def test_binding_table_covers_headers():
    assert set(declared_functions()) == set(E.EXPORTED)


def test_error_api_without_gpu_work():
    L = E.lib()
    buf = ctypes.create_string_buffer(64)
    L.emc_hip_clear_error()
    assert L.emc_hip_last_error(buf, 64) == 0
    assert L.emc_hip_phase_name(0) == b"expand"
    assert L.emc_hip_phase_name(99) is None
