"""NetCDF output files (seisnc.cpp:1229-1490, 1671-1712) written by the
library's own classic-format writer, read back with scipy's independent
netCDF reader: dimensions, variable names and shapes, values, record counts
and the global attributes the reference writes.  No GPU needed."""
import struct

import numpy as np
import pytest
from scipy.io import netcdf_file

cgfd3d = pytest.importorskip("cgfd3d")

NAMES = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz"]


def _assert_text_att_with_nul(path, text):
    """nc_put_att_text(..., sizeof(str), str) stores the terminating NUL too
    (seisnc.cpp:1263, 1322-1327); scipy strips it, so look at the header:
    name, NC_CHAR, nelems = len+1, text, NUL."""
    raw = open(path, "rb").read()
    name = b"Content"
    want = (struct.pack(">I", len(name)) + name + b"\x00" + struct.pack(">II", 2, len(text) + 1) + text + b"\x00")
    assert want in raw


def _lib_or_skip():
    try:
        cgfd3d.load_library()
    except cgfd3d.CgfdError as e:  # library not built
        pytest.skip(str(e))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_seismo_file_round_trip(tmp_path, dtype):
    _lib_or_skip()
    rng = np.random.default_rng(0)
    nrec, npnt = 7, 5
    wave = rng.standard_normal((9, nrec, npnt)).astype(dtype)
    t = ((np.arange(nrec) + 1) * dtype(0.005)).astype(dtype)
    pnum = np.arange(npnt, dtype=np.int32)
    lnum = np.array([1, 1, 2, 2, 2], np.int32)
    pos = [rng.uniform(0, 1e4, npnt).astype(dtype) for _ in range(3)]
    path = tmp_path / "seismo.nc"
    cgfd3d.nc_seismo(path, wave, t, pnum, lnum, *pos)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions["TimeNum"] is None and f.dimensions["PointNum"] == npnt
        assert f._recs == nrec
        for c, n in enumerate(NAMES):
            v = f.variables[n]
            assert v.dimensions == ("TimeNum", "PointNum")
            assert v.data.dtype == np.dtype(dtype).newbyteorder(">")
            np.testing.assert_array_equal(v.data, wave[c])
        np.testing.assert_array_equal(f.variables["currTime"].data, t)
        np.testing.assert_array_equal(f.variables["Pnum"].data, pnum)
        np.testing.assert_array_equal(f.variables["Lnum"].data, lnum)
        for n, p in zip(("posx", "posy", "posz"), pos):
            np.testing.assert_array_equal(f.variables[n].data, p)
        assert f.Content == b"Velocity and Stress seismogram"
    _assert_text_att_with_nul(path, b"Velocity and Stress seismogram")


@pytest.mark.parametrize("cmp", [1, 2, 3])
def test_snapshot_file_round_trip(tmp_path, cmp):
    _lib_or_skip()
    rng = np.random.default_rng(cmp)
    spec = np.array([2, 3, 4, 3, 2, 5, 1, 2, 3, 10, cmp], np.int32)
    nrec = 4
    comps = {1: range(3), 2: range(3, 9), 3: range(9)}[cmp]
    wave = {c: rng.standard_normal((nrec, 3, 2, 5)).astype(np.float32) for c in comps}
    t = ((np.arange(nrec) * 10 + 1) * np.float32(0.004)).astype(np.float32)
    path = tmp_path / "snap_001.nc"
    cgfd3d.nc_snap(path, spec, wave, t)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.dimensions["TimeNum"] is None
        assert (f.dimensions["i"], f.dimensions["j"], f.dimensions["k"]) == (3, 2, 5)
        assert f._recs == nrec
        assert set(f.variables) == {NAMES[c] for c in comps} | {"currTime"}
        for c in comps:
            assert f.variables[NAMES[c]].dimensions == ("TimeNum", "i", "j", "k")
            np.testing.assert_array_equal(f.variables[NAMES[c]].data, wave[c])
        np.testing.assert_array_equal(f.variables["currTime"].data, t)
        np.testing.assert_array_equal(f.IdxStart, [2, 3, 4])
        np.testing.assert_array_equal(f.IdxNumber, [3, 2, 5])
        np.testing.assert_array_equal(f.IdxInterval, [1, 2, 3])
        assert int(np.asarray(f.TimeInterval).reshape(-1)[0]) == 10
        want = {1: b"Velocity field", 2: b"Stress field", 3: b"Velocity and Stress field"}[cmp]
        assert f.Content == want
    _assert_text_att_with_nul(path, want)


def test_peakvel_file_round_trip(tmp_path):
    _lib_or_skip()
    rng = np.random.default_rng(5)
    v = [rng.uniform(0, 1, (6, 4)).astype(np.float32) for _ in range(3)]
    path = tmp_path / "peakvel.nc"
    cgfd3d.nc_peakvel(path, *v)
    with netcdf_file(path, "r", mmap=False) as f:
        assert (f.dimensions["i"], f.dimensions["j"]) == (6, 4)
        for n, a in zip(("Vx", "Vy", "Vz"), v):
            assert f.variables[n].dimensions == ("i", "j")
            np.testing.assert_array_equal(f.variables[n].data, a)


def test_odd_sized_record_variables_keep_alignment(tmp_path):
    """Records whose slabs are not multiples of 8 bytes, and a single
    point: offsets of every variable stay consistent for the reader."""
    _lib_or_skip()
    wave = np.arange(9 * 3 * 1, dtype=np.float32).reshape(9, 3, 1)
    t = np.array([1, 2, 3], np.float32)
    path = tmp_path / "one.nc"
    cgfd3d.nc_seismo(path, wave, t, [0], [1], [1.0], [2.0], [3.0])
    with netcdf_file(path, "r", mmap=False) as f:
        for c, n in enumerate(NAMES):
            np.testing.assert_array_equal(f.variables[n].data, wave[c])
        np.testing.assert_array_equal(f.variables["posz"].data, [3.0])
