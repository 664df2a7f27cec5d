"""A run assembled by the parameter-file front end (seisconf.load) from a
set of the reference's parameter files: the engine against the oracle, bit
for bit, and the seismogram file written with the station metadata the
front end read (seisplot.cpp:330-440, seisnc.cpp:1229-1365)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from conftest import assert_same

pytestmark = pytest.mark.gpu

cgfd3d = pytest.importorskip("cgfd3d")
import oracle  # noqa: E402
import seisconf  # noqa: E402
import seisconf_fixture as FX  # noqa: E402


def _run(sc, nsteps):
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g.setup(sc)
    o.setup(sc)
    g.run(nsteps)
    o.run(nsteps)
    return g, o


def test_conf_run_matches_oracle(tmp_path):
    nt = 12
    sc = seisconf.load(str(FX.write_set(tmp_path, nt=nt)))
    print(f"conf run: {sc['ni']}x{sc['nj']}x{sc['nk']}, {len(sc['recv_locs'])} receivers", flush=True)
    g, o = _run(sc, nt)
    assert_same(g.get_seismo(), o.get_seismo(), "seismograms")
    assert_same(g.get_pgv()[:, 3:-3, 3:-3], o.get_pgv()[:, 3:-3, 3:-3], "PGV")
    for c in range(9):
        assert_same(g.get_field(c), o.get_field(c), f"field {c}")
    assert np.abs(g.get_seismo()).max() > 0
    # snapshot boxes of the station file
    for i, spec in enumerate(sc["snapshots"]):
        n, t = g.snap_shape(i)
        assert n == spec[3] * spec[4] * spec[5] and t == -(-nt // spec[9])
    path = tmp_path / "seismo.nc"
    g.write_seismo(path, sc["seismo_tinv"], sc["recv_pnum"], sc["recv_lnum"], *sc["recv_pos"].T)
    with netcdf_file(path, "r", mmap=False) as f:
        assert_same(f.variables["posz"].data, sc["recv_pos"][:, 2], "posz")
        assert_same(f.variables["Lnum"].data, sc["recv_lnum"], "Lnum")
        assert_same(f.variables["Vz"].data, g.get_seismo()[2].T, "Vz")


def test_conf_focal_run_matches_oracle(tmp_path):
    nt = 10
    sc = seisconf.load(str(FX.write_set(tmp_path, nt=nt, source="focal")))
    # the point beyond the grid's x range is put on the nearest valid cell,
    # the last x column, as reposition does
    assert sc["focal_locs"].shape == (3, 3) and sc["focal_locs"][2, 0] == 3 + FX.NI - 1
    g, o = _run(sc, nt)
    assert_same(g.get_seismo(), o.get_seismo(), "seismograms")
    for c in range(9):
        assert_same(g.get_field(c), o.get_field(c), f"field {c}")
    assert np.abs(g.get_seismo()).max() > 0


def test_conf_interface3d_media_run(tmp_path):
    """Layered interface3D media (gradient layer over a tilted interface,
    harmonic averaging on a grid row) from the parameter files."""
    nt = 8
    FX.write_set(tmp_path, nt=nt, media=False)
    sc = seisconf.load(str(FX.write_interface3d(tmp_path)))
    g, o = _run(sc, nt)
    assert_same(g.get_seismo(), o.get_seismo(), "seismograms")
    for c in range(9):
        assert_same(g.get_field(c), o.get_field(c), f"field {c}")


def test_conf_hybrid_grid_run_matches_oracle(tmp_path):
    """hybrid_grid = 1 from the parameter files: the hybrid layering puts
    ConIndex at 3 + 4, and the run (Cartesian rates below it) equals the
    oracle's."""
    nt = 10
    sc = seisconf.load(str(FX.write_set(tmp_path, nt=nt, hybrid=1, conv="-1200.0")))
    assert sc["conindex"] == 7
    g, o = _run(sc, nt)
    assert_same(g.get_seismo(), o.get_seismo(), "seismograms")
    assert_same(g.get_pgv()[:, 3:-3, 3:-3], o.get_pgv()[:, 3:-3, 3:-3], "PGV")
    for c in range(9):
        assert_same(g.get_field(c), o.get_field(c), f"field {c}")
    assert np.abs(g.get_seismo()).max() > 0


def test_reference_example_cut_run_matches_oracle():
    """The reference's own example (EXAMPLE/hillmodel: hybrid grid at -3 km,
    focal point at 5 km depth, 12-cell PML, its receivers and lines), cut to
    the 48 x 48 columns around the focal point with stand-in media
    (tests/golden/hillmodel_cut, make_hillmodel_cut.py): 30 steps equal the
    oracle value for value."""
    import os

    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hillmodel_cut")
    sc = seisconf.load(os.path.join(d, "SeisFD3D.conf"), seispath=d)
    assert sc["conindex"] == 127
    g, o = _run(sc, 30)
    assert_same(g.get_seismo(), o.get_seismo(), "seismograms")
    assert_same(g.get_pgv()[:, 3:-3, 3:-3], o.get_pgv()[:, 3:-3, 3:-3], "PGV")
    for c in range(9):
        assert_same(g.get_field(c), o.get_field(c), f"field {c}")
    assert np.abs(g.get_seismo()).max() > 0
