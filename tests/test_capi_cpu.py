"""The C-ABI libraries load on a machine without a GPU and export every
entry point that include/cgfd3d.h and include/cgfd_fno.h declare (no compute
calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cgfd3d.h")
LIB = os.path.join(ROOT, "copara-cgfd3d_amd", "libcgfd3d.so")


def declared_symbols(header=HEADER):
    text = open(header).read()
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


@pytest.mark.skipif(not os.path.exists(LIB), reason="libcgfd3d.so not built")
def test_batch_entry_points_validate_without_the_gpu():
    import cgfd3d
    lib = cgfd3d.load_library()
    assert lib.cgfd_batch_fusable(None, 2) == -1                # no list
    assert lib.cgfd_run_batch(None, 2, 1) == -1
    empty = (ctypes.c_void_p * 2)(None, None)
    assert lib.cgfd_batch_fusable(empty, 2) == -1               # null contexts
    assert lib.cgfd_run_batch(empty, 2, 1) == -1
    assert lib.cgfd_batch_fusable(empty, 0) == 0                # an empty list is not fusable
    assert lib.cgfd_batch_fusable(empty, -1) == -1
    # the shared medium and tile width entry points refuse null contexts
    assert lib.cgfd_share_medium(None, None) == -1
    assert lib.cgfd_set_tile_width(None, 32) == -1
    assert lib.cgfd_tile_width(None) == -1


FNO_HEADER = os.path.join(ROOT, "include", "cgfd_fno.h")
FNO_LIB = os.path.join(ROOT, "copara-cgfd3d_amd", "libcgfd_fno.so")


@pytest.mark.skipif(not os.path.exists(FNO_LIB), reason="libcgfd_fno.so not built")
def test_fno_library_exports_and_validates_without_the_gpu():
    names = declared_symbols(FNO_HEADER)
    assert sorted(names) == ["cgfd_fno_adam", "cgfd_fno_gelu_bwd", "cgfd_fno_gemm_bf16", "cgfd_fno_gemm_bf16_ex"]
    lib = ctypes.CDLL(FNO_LIB)
    assert all(hasattr(lib, n) for n in names)
    f = lib.cgfd_fno_gemm_bf16
    L, P, I = ctypes.c_long, ctypes.c_void_p, ctypes.c_int
    f.argtypes = [P, L, L, L, P, L, L, L, P, L, L, I, I, I, I, I, P]
    dummy = ctypes.c_void_p(16)
    # rejected before any launch: empty sizes, null operands, too many row tiles
    assert f(dummy, 0, 1, 1, dummy, 0, 1, 1, dummy, 0, 1, 0, 4, 4, 1, 0, None) == -1
    assert f(None, 0, 1, 1, dummy, 0, 1, 1, dummy, 0, 1, 4, 4, 4, 1, 0, None) == -1
    assert f(dummy, 0, 1, 1, dummy, 0, 1, 1, dummy, 0, 1, 64 * 65536, 4, 4, 1, 0, None) == -1

    # the fused-epilogue product, the GELU backward and Adam reject bad
    # arguments before any launch
    g = lib.cgfd_fno_gemm_bf16_ex
    g.argtypes = [P, L, L, L, P, L, L, L, P, L, L, I, I, I, I, I, P, P, L, L, P, I, P, L, L, P]
    assert g(dummy, 0, 1, 1, dummy, 0, 1, 1, dummy, 0, 1, 4, 4, 4, 1, 0, None, None, 0, 0, None, 7, None, 0, 0,
             None) == -1
    assert g(dummy, 0, 1, 1, dummy, 0, 1, 1, dummy, 0, 1, 4, 4, 0, 1, 0, None, None, 0, 0, None, 1, None, 0, 0,
             None) == -1
    b = lib.cgfd_fno_gelu_bwd
    b.argtypes = [P, P, P, P, L, P]
    assert b(None, dummy, dummy, None, 4, None) == -1
    assert b(dummy, dummy, dummy, None, -1, None) == -1
    assert b(dummy, dummy, dummy, None, 0, None) == 0          # nothing to do
    a = lib.cgfd_fno_adam
    D, LL = ctypes.c_double, ctypes.c_longlong
    a.argtypes = [P, P, I, D, D, D, D, LL, I, P]
    assert a(dummy, dummy, 1, 1e-3, 0.9, 0.999, 1e-8, 0, 1, None) == -1     # steps count from 1
    assert a(dummy, dummy, 1, 1e-3, 1.0, 0.999, 1e-8, 1, 1, None) == -1     # beta1 must be below 1
    assert a(None, dummy, 1, 1e-3, 0.9, 0.999, 1e-8, 1, 1, None) == -1
    assert a(dummy, dummy, 0, 1e-3, 0.9, 0.999, 1e-8, 1, 1, None) == 0      # no chunks


NC4_HEADER = os.path.join(ROOT, "include", "cgfd_nc4.h")
NC4_LIB = os.path.join(ROOT, "copara-cgfd3d_amd", "libcgfd_nc4.so")


@pytest.mark.skipif(not os.path.exists(NC4_LIB), reason="libcgfd_nc4.so not built (no libhdf5)")
def test_nc4_library_exports_every_declared_symbol(tmp_path):
    names = declared_symbols(NC4_HEADER)
    assert sorted(names) == ["cgfd_nc4_get_att_float", "cgfd_nc4_get_var_double", "cgfd_nc4_get_var_float",
                             "cgfd_nc4_probe", "cgfd_nc4_var_dims"]
    lib = ctypes.CDLL(NC4_LIB)
    assert all(hasattr(lib, n) for n in names)
    lib.cgfd_nc4_probe.argtypes = [ctypes.c_char_p]
    bad = tmp_path / "x.nc"
    bad.write_bytes(b"CDF\x01" + b"\0" * 60)
    assert lib.cgfd_nc4_probe(None) == -1
    assert lib.cgfd_nc4_probe(str(bad).encode()) == 0              # classic, not HDF5
    assert lib.cgfd_nc4_probe(str(tmp_path / "missing.nc").encode()) == 0
