"""The C++ parameter-file front end (copara-cgfd3d_amd/host/cgfd_conf.cpp,
run as `cgfd_child --conf ... --dump`) against the Python front end
(seisconf.load): every array the child hands to the engine -- coordinates,
metrics, media, PML tables, sources, focal series, receivers with their
metadata, snapshot specs, ConIndex -- equal value for value.

Cases: the cut-down reference example (hybrid grid, netCDF focal file,
volume media, receivers and lines), the same with the example's own
netCDF-4 focal file stf-GPU.nc (when the reference is present), and
generated sets with force and angle-mode moment sources, a focal file, a
hybrid grid, volume media with smoothing, and interface3D media with and
without smoothing; plus the refusals.  No GPU is used."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import seisconf as S
import seisconf_fixture as FX

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHILD = os.path.join(ROOT, "copara-cgfd3d_amd", "cgfd_child")
GOLDEN = os.path.join(ROOT, "tests", "golden", "hillmodel_cut")
EXAMPLE_STF = "/root/reference/EXAMPLE/hillmodel/stf-GPU.nc"
F32 = np.float32


def _dump(conf, out, seispath=None):
    if not os.path.exists(CHILD):
        pytest.fail("cgfd_child is not built (copara-cgfd3d_amd/build.sh)")
    os.makedirs(out, exist_ok=True)
    cmd = [CHILD, "--conf", str(conf), "--dump", str(out)] + (["--seispath", str(seispath)] if seispath else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode:
        raise RuntimeError(r.stderr.strip())
    with open(os.path.join(out, "dims.txt")) as f:
        a = f.readline().split()
        b = f.readline().split()
    ni, nj, nk, nt = int(a[0]), int(a[1]), int(a[2]), int(a[5])
    nf, nm, nr, ns = int(a[7]), int(a[8]), int(a[9]), int(a[10])
    nfoc, fnt = int(b[0]), int(b[1])
    N = (ni + 6) * (nj + 6) * (nk + 6)
    nstf = 2 * (nt + 1)

    def rd(name, dt, shape):
        v = np.fromfile(os.path.join(out, name), dt)
        return v.reshape(shape)

    d = {"ni": ni, "nj": nj, "nk": nk, "nt": nt, "dh": float(a[3]), "dt": float(a[4]), "pgv": int(a[11]),
         "focal_dt": float(b[2]), "conindex": int(b[3]), "seismo_tinv": int(b[4])}
    shp = (ni + 6, nj + 6, nk + 6)
    d["coords"] = rd("coords.bin", F32, (3,) + shp)
    d["metrics"] = rd("metrics.bin", F32, (10,) + shp)
    d["media"] = rd("media.bin", F32, (3,) + shp)
    d["nabs"] = rd("nabs.bin", np.int32, (6,))
    d["pml"] = np.fromfile(os.path.join(out, "pml.bin"), F32)
    d["force_loc"], d["force_vec"] = rd("force_loc.bin", np.int32, (nf, 3)), rd("force_vec.bin", F32, (nf, 3))
    d["force_stf"] = rd("force_stf.bin", F32, (nf, nstf))
    d["moment_loc"], d["moment_vec"] = rd("moment_loc.bin", np.int32, (nm, 3)), rd("moment_vec.bin", F32, (nm, 6))
    d["moment_stf"] = rd("moment_stf.bin", F32, (nm, nstf))
    d["focal_loc"] = rd("focal_loc.bin", np.int32, (nfoc, 3))
    d["focal_series"] = rd("focal_series.bin", F32, (6, nfoc, fnt))
    d["recv"] = rd("recv.bin", np.int32, (nr, 3))
    d["recv_pos"] = rd("recv_pos.bin", F32, (nr, 3))
    d["recv_pnum"], d["recv_lnum"] = rd("recv_pnum.bin", np.int32, (nr,)), rd("recv_lnum.bin", np.int32, (nr,))
    d["snap"] = rd("snap.bin", np.int32, (ns, 11))
    return d


METRICS = ("xix", "xiy", "xiz", "etax", "etay", "etaz", "zetax", "zetay", "zetaz", "jac")
PML = ("APDx", "APDy", "APDz", "Bx", "By", "Bz", "DBx", "DBy", "DBz")


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    assert a.dtype == b.dtype or a.dtype.kind == b.dtype.kind, f"{what}: dtype {a.dtype} vs {b.dtype}"
    same = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else (a == b)
    assert same.all(), f"{what}: {np.count_nonzero(~same)} of {a.size} values differ"


def _compare(sc, d):
    assert (d["ni"], d["nj"], d["nk"], d["nt"]) == (sc["ni"], sc["nj"], sc["nk"], sc["nt"])
    assert F32(d["dh"]) == F32(sc["dh"]) and F32(d["dt"]) == F32(sc["dt"]) and d["pgv"] == int(sc["pgv"])
    assert d["conindex"] == sc.get("conindex", 0)
    for q, c in enumerate(sc["coords"]):
        _same(d["coords"][q], np.broadcast_to(c, d["coords"][q].shape), f"coords {q}")
    for q, nm in enumerate(METRICS):
        _same(d["metrics"][q], sc["metrics"][nm], nm)
    for q, nm in enumerate(("alpha", "beta", "rho")):
        _same(d["media"][q], sc[nm], nm)
    _same(d["nabs"], np.asarray(sc["nabs"], np.int32), "nabs")
    _same(d["pml"], np.concatenate([sc["pml"][nm] for nm in PML]), "pml")
    for kind, m in (("force", 3), ("moment", 6)):
        if kind + "_locs" in sc:
            _same(d[kind + "_loc"], sc[kind + "_locs"], kind + " locations")
            _same(d[kind + "_vec"], sc[kind + "_vec"], kind + " vectors")
            _same(d[kind + "_stf"], sc[kind + "_stf"], kind + " STF")
        else:
            assert len(d[kind + "_loc"]) == 0
    if "focal_locs" in sc:
        _same(d["focal_loc"], sc["focal_locs"], "focal locations")
        _same(d["focal_series"], sc["focal_series"], "focal series")
        assert d["focal_dt"] == sc["focal_dt"]
    else:
        assert len(d["focal_loc"]) == 0
    _same(d["recv"], sc["recv_locs"], "receiver locations")
    _same(d["recv_pos"], sc["recv_pos"], "receiver positions")
    _same(d["recv_pnum"], sc["recv_pnum"], "Pnum")
    _same(d["recv_lnum"], sc["recv_lnum"], "Lnum")
    assert d["seismo_tinv"] == sc["seismo_tinv"]
    _same(d["snap"], sc["snapshots"], "snapshot specs")


def test_cut_reference_example(tmp_path):
    sc = S.load(os.path.join(GOLDEN, "SeisFD3D.conf"), seispath=GOLDEN)
    d = _dump(os.path.join(GOLDEN, "SeisFD3D.conf"), tmp_path / "out", seispath=GOLDEN)
    _compare(sc, d)
    assert d["conindex"] == 127 and len(d["focal_loc"]) == 1 and len(d["recv"]) == 35


@pytest.mark.skipif(not os.path.exists(EXAMPLE_STF), reason="reference example not present")
@pytest.mark.skipif(S._nc4_library() is None, reason="libcgfd_nc4.so (libhdf5) not built")
def test_cut_reference_example_with_its_netcdf4_focal_file(tmp_path):
    d0 = tmp_path / "set"
    shutil.copytree(GOLDEN, d0)
    shutil.copy(EXAMPLE_STF, d0 / "stf-GPU.nc")
    src = (d0 / "source.dat").read_text().replace("stf-classic.nc", "stf-GPU.nc")
    (d0 / "source.dat").write_text(src)
    sc = S.load(str(d0 / "SeisFD3D.conf"), seispath=str(d0))
    d = _dump(d0 / "SeisFD3D.conf", tmp_path / "out", seispath=d0)
    _compare(sc, d)
    classic = _dump(os.path.join(GOLDEN, "SeisFD3D.conf"), tmp_path / "classic", seispath=GOLDEN)
    _same(d["focal_series"], classic["focal_series"], "netCDF-4 vs classic focal series")
    assert d["focal_series"].shape == (6, 1, 1500) and np.abs(d["focal_series"]).max() > 0


@pytest.mark.parametrize("case", ["force_and_angle_moments", "focal", "hybrid", "volume_smoothing"])
def test_generated_sets(tmp_path, case):
    kw = {"force_and_angle_moments": {}, "focal": {"source": "focal"},
          "hybrid": {"hybrid": 1, "conv": "-1200.0"}, "volume_smoothing": {}}[case]
    conf = FX.write_set(tmp_path / "set", **kw)
    if case == "volume_smoothing":
        txt = conf.read_text().replace("sampling_point_per_cell = 0 0 0", "sampling_point_per_cell = 2 2 3")
        txt = txt.replace("threshold_velocity_jump_of_interface = 100", "threshold_velocity_jump_of_interface = 5")
        conf.write_text(txt)
    sc = S.load(str(conf))
    d = _dump(conf, tmp_path / "out")
    _compare(sc, d)


@pytest.mark.parametrize("sampling", [None, (2, 2, 2)])
def test_interface3d_sets(tmp_path, sampling):
    FX.write_set(tmp_path / "set", media=False)
    conf = FX.write_interface3d(tmp_path / "set", z2=-980.0, sampling=sampling)
    sc = S.load(str(conf))
    d = _dump(conf, tmp_path / "out")
    _compare(sc, d)


def test_refusals_match_the_python_front_end(tmp_path):
    conf = FX.write_set(tmp_path / "p", media=False)
    conf.write_text(conf.read_text().replace("grid_type = vmap", "grid_type = point"))
    with pytest.raises(RuntimeError, match="crd.x"):
        _dump(conf, tmp_path / "o1")
    conf.write_text(conf.read_text().replace("grid_type = point", "grid_type = vmap")
                    .replace("media_type = volume", "media_type = interface"))
    with pytest.raises(RuntimeError, match="NaN"):
        _dump(conf, tmp_path / "o2")
    conf.write_text(conf.read_text().replace("media_type = interface", "media_type = volume"))
    grid = tmp_path / "p" / "grid.dat"
    grid.write_text(grid.read_text().replace("vmap_dims", "vmap_dimz"))
    with pytest.raises(RuntimeError, match="vmap_dims"):
        _dump(conf, tmp_path / "o3")
