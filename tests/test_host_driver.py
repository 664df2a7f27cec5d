"""The C++ host driver (copara-cgfd3d_amd/host/cgfd_child.cpp) runs one
child-rank simulation through the C-ABI from compiled code; its traces,
peak velocities and NetCDF files must equal those of the same run driven
through the ctypes binding (and therefore the oracle's)."""
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

from conftest import assert_same

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "copara-cgfd3d_amd", "cgfd_child")


def dump_inputs(sc, d, dtype, nsnap_spec=None):
    """The master's arrays for one child, as cgfd_child reads them."""
    d.mkdir(parents=True, exist_ok=True)
    nf = 0 if sc.get("force_locs") is None else len(sc["force_locs"])
    nm = 0 if sc.get("moment_locs") is None else len(sc["moment_locs"])
    snaps = np.zeros((0, 11), np.int32) if nsnap_spec is None else np.asarray(nsnap_spec, np.int32)
    prec = 1 if dtype == np.float64 else 0
    (d / "dims.txt").write_text(f"{sc['ni']} {sc['nj']} {sc['nk']} {sc['dh']!r} {sc['dt']!r} {sc['nt']} {prec} "
                                f"{nf} {nm} {len(sc['recv_locs'])} {len(snaps)} 1\n")
    names = ("xix", "xiy", "xiz", "etax", "etay", "etaz", "zetax", "zetay", "zetaz", "jac")
    np.concatenate([sc["metrics"][k].astype(dtype).ravel() for k in names]).tofile(d / "metrics.bin")
    np.concatenate([sc[k].astype(dtype).ravel() for k in ("alpha", "beta", "rho")]).tofile(d / "media.bin")
    np.asarray(sc["nabs"], np.int32).tofile(d / "nabs.bin")
    order = ("APDx", "APDy", "APDz", "Bx", "By", "Bz", "DBx", "DBy", "DBz")
    np.concatenate([sc["pml"][k].astype(dtype) for k in order]).tofile(d / "pml.bin")
    for kind, n in (("force", nf), ("moment", nm)):
        if n:
            np.asarray(sc[f"{kind}_locs"], np.int32).tofile(d / f"{kind}_loc.bin")
            np.asarray(sc[f"{kind}_vec"], dtype).tofile(d / f"{kind}_vec.bin")
            np.asarray(sc[f"{kind}_stf"], dtype).tofile(d / f"{kind}_stf.bin")
    np.asarray(sc["recv_locs"], np.int32).tofile(d / "recv.bin")
    snaps.tofile(d / "snap.bin")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="host driver not built")
def test_driver_reports_missing_inputs(tmp_path):
    # fails on the missing input file before any GPU call
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "dims.txt" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("double", [False, True])
def test_driver_matches_ctypes_run(tmp_path, double):
    cgfd3d = pytest.importorskip("cgfd3d")
    import scenario

    if not os.path.exists(DRIVER):
        pytest.fail("cgfd_child is not built (copara-cgfd3d_amd/build.sh)")
    dtype = np.float64 if double else np.float32
    nt = 14
    sc = scenario.config_random(30, 26, 18, nt, [4, 4, 3, 3, 4, 0], seed=71)
    if double:
        for k in ("alpha", "beta", "rho"):
            sc[k] = sc[k].astype(np.float64)
    snaps = [[1, 1, sc["nk"], sc["ni"], sc["nj"], 1, 1, 1, 1, 3, 1]]
    dump_inputs(sc, tmp_path, dtype, snaps)
    print(f"driver run: double={double}", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"],
                          precision=cgfd3d.FP64 if double else cgfd3d.FP32)
    sc = dict(sc, snapshots=np.array(snaps, np.int32))
    g.setup(sc)
    g.run(nt)
    seis = g.get_seismo()
    traces = np.fromfile(tmp_path / "traces.bin", dtype).reshape(seis.shape)
    assert_same(traces, seis, "traces")
    pv = np.fromfile(tmp_path / "pgv.bin", dtype).reshape(3, sc["ni"] + 6, sc["nj"] + 6)
    assert_same(pv[:, 3:-3, 3:-3], g.get_pgv()[:, 3:-3, 3:-3], "PGV")
    with netcdf_file(tmp_path / "snap_001.nc", "r", mmap=False) as f:
        want = g.get_snap(0, 2).T.reshape(-1, sc["ni"], sc["nj"], 1)
        assert_same(f.variables["Vz"].data, want, "snap Vz")
    with netcdf_file(tmp_path / "seismo.nc", "r", mmap=False) as f:
        assert_same(f.variables["Txy"].data, seis[6].T, "seismo Txy")
    assert np.abs(seis).max() > 0


@pytest.mark.gpu
def test_driver_runs_the_reference_example_from_its_parameter_files(tmp_path):
    """cgfd_child --conf: the cut-down reference example (hybrid grid, focal
    source, its receivers and lines) prepared by the C++ front end and run
    from compiled code gives the traces, PGV and seismo.nc of the same run
    prepared by seisconf.load and driven through ctypes."""
    cgfd3d = pytest.importorskip("cgfd3d")
    import seisconf

    if not os.path.exists(DRIVER):
        pytest.fail("cgfd_child is not built (copara-cgfd3d_amd/build.sh)")
    golden = os.path.join(ROOT, "tests", "golden", "hillmodel_cut")
    steps = 30
    print("driver run from SeisFD3D.conf", flush=True)
    r = subprocess.run([DRIVER, "--conf", os.path.join(golden, "SeisFD3D.conf"), "--seispath", golden,
                        "--out", str(tmp_path), "--steps", str(steps)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    sc = seisconf.load(os.path.join(golden, "SeisFD3D.conf"), seispath=golden)
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g.setup(sc)
    g.run(steps)
    seis = g.get_seismo()
    traces = np.fromfile(tmp_path / "traces.bin", np.float32).reshape(seis.shape)
    assert_same(traces, seis, "traces")
    pv = np.fromfile(tmp_path / "pgv.bin", np.float32).reshape(3, sc["ni"] + 6, sc["nj"] + 6)
    assert_same(pv[:, 3:-3, 3:-3], g.get_pgv()[:, 3:-3, 3:-3], "PGV")
    assert np.abs(seis[:3, :, :steps]).max() > 0
    ref = tmp_path / "py"
    ref.mkdir()
    pos = sc["recv_pos"]
    g.write_seismo(str(ref / "seismo.nc"), sc["seismo_tinv"], sc["recv_pnum"], sc["recv_lnum"], pos[:, 0], pos[:, 1],
                   pos[:, 2])
    with netcdf_file(tmp_path / "seismo.nc", "r", mmap=False) as f, netcdf_file(ref / "seismo.nc", "r", mmap=False) as h:
        for name in h.variables:
            assert_same(f.variables[name].data, h.variables[name].data, f"seismo.nc {name}")
    g.close()


@pytest.mark.gpu
def test_driver_simplecar_option_matches_the_python_path(tmp_path):
    """cgfd_child --conf ... --simplecar runs the reference's SimpleCar build
    (xi/eta metrics taken as the identity) and gives the traces of the same
    build prepared by seisconf.load(simplecar=True)."""
    cgfd3d = pytest.importorskip("cgfd3d")
    import seisconf
    import seisconf_fixture as FX

    if not os.path.exists(DRIVER):
        pytest.fail("cgfd_child is not built (copara-cgfd3d_amd/build.sh)")
    conf = FX.write_set(tmp_path / "set", nt=16, hybrid=1, conv="-1200.0")
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([DRIVER, "--conf", str(conf), "--simplecar", "--out", str(out)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    sc = seisconf.load(str(conf), simplecar=True)
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g.setup(sc)
    g.run(sc["nt"])
    seis = g.get_seismo()
    assert_same(np.fromfile(out / "traces.bin", np.float32).reshape(seis.shape), seis, "traces")
    assert np.abs(seis).max() > 0
    g.close()
