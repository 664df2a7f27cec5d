"""The C-ABI library loads on a machine without a GPU and exports every
entry point that include/cgfd3d.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cgfd3d.h")
LIB = os.path.join(ROOT, "copara-cgfd3d_amd", "libcgfd3d.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cgfd_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_expected_surface():
    names = declared_symbols()
    for required in ("cgfd_create", "cgfd_destroy", "cgfd_set_metrics", "cgfd_set_media", "cgfd_set_pml",
                     "cgfd_set_moments", "cgfd_set_forces", "cgfd_set_receivers", "cgfd_init", "cgfd_run",
                     "cgfd_get_seismo", "cgfd_get_pgv", "cgfd_last_error"):
        assert required in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libcgfd3d.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="libcgfd3d.so not built")
def test_python_binding_covers_the_header():
    import cgfd3d
    assert set(declared_symbols()) <= set(cgfd3d.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libcgfd3d.so not built")
def test_create_rejects_bad_dims_without_touching_the_gpu():
    import cgfd3d
    lib = cgfd3d.load_library()
    h = ctypes.c_void_p()
    bad = cgfd3d.Dims(0, 10, 10, 100.0, 0.01, 10)
    assert lib.cgfd_create(ctypes.byref(bad), 0, 0, ctypes.byref(h)) == -1
    bad = cgfd3d.Dims(10, 10, 3, 100.0, 0.01, 10)  # nk < 4: no room for the surface rows
    assert lib.cgfd_create(ctypes.byref(bad), 0, 0, ctypes.byref(h)) == -1
    bad = cgfd3d.Dims(10, 10, 10, 100.0, 0.01, 10)
    assert lib.cgfd_create(ctypes.byref(bad), 7, 0, ctypes.byref(h)) == -1
    # one padded field past the 4 GiB range of the kernels' 32-bit offsets:
    # 1024 x 1024 x 1000 cells fit in HBM but not in one field's offsets
    big = cgfd3d.Dims(1024, 1024, 1000, 100.0, 0.01, 10)
    assert lib.cgfd_create(ctypes.byref(big), 0, 0, ctypes.byref(h)) == -1
    # FP64 halves the cells that fit: a 900^3 field is 3.2 GB in FP32, 6.3 GB in FP64
    mid = cgfd3d.Dims(900, 900, 900, 100.0, 0.01, 10)
    assert lib.cgfd_create(ctypes.byref(mid), 1, 0, ctypes.byref(h)) == -1
    assert not h.value
