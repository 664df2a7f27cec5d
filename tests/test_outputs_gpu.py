"""End-of-run NetCDF outputs of a context (seismo.nc, snap_NNN.nc,
peakvel.nc) against the engine's own traces, snapshot slabs and PGV maps,
with the record rules of the reference's master (seisplot.cpp:672-760,
seisnc.cpp:1336-1365, 1448-1490, 1671-1712)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from conftest import assert_same

pytestmark = pytest.mark.gpu

cgfd3d = pytest.importorskip("cgfd3d")
import scenario  # noqa: E402

NAMES = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz"]


def _run(nt=11):
    sc = scenario.config_random(28, 24, 16, nt, [4, 4, 3, 3, 4, 0], seed=61)
    nk = sc["nk"]
    sc["snapshots"] = np.array([[1, 1, nk, 28, 24, 1, 1, 1, 1, 4, 1],
                                [2, 2, 2, 4, 3, 5, 5, 6, 3, 3, 3]], np.int32)
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g.setup(sc)
    g.run(nt)
    return sc, g


@pytest.mark.parametrize("tinv", [1, 3])
def test_seismo_nc(tmp_path, tinv):
    nt = 11
    sc, g = _run(nt)
    seis = g.get_seismo()                                      # [9][nrecv][nt]
    n = seis.shape[1]
    pnum = np.arange(n, dtype=np.int32) + 100
    lnum = np.ones(n, np.int32)
    pos = [np.arange(n, dtype=np.float32) * k for k in (1, 2, 3)]
    path = tmp_path / "seismo.nc"
    g.write_seismo(path, tinv, pnum, lnum, *pos)
    nrec = (nt - 1) // tinv + 1
    last = [min(r * tinv + tinv - 1, nt - 1) for r in range(nrec)]  # the last step written into each record
    with netcdf_file(path, "r", mmap=False) as f:
        assert f._recs == nrec and f.dimensions["PointNum"] == n
        for c, name in enumerate(NAMES):
            assert_same(f.variables[name].data, seis[c][:, last].T, f"seismo.nc {name}")
        want_t = np.array([np.float32(it + 1) * np.float32(sc["dt"]) for it in last], np.float32)
        assert_same(f.variables["currTime"].data, want_t, "currTime")
        assert_same(f.variables["Pnum"].data, pnum, "Pnum")
        assert_same(f.variables["posy"].data, pos[1], "posy")


def test_snap_nc(tmp_path):
    nt = 11
    sc, g = _run(nt)
    for i, spec in enumerate(sc["snapshots"]):
        xn, yn, zn, tinv, cmp = int(spec[3]), int(spec[4]), int(spec[5]), int(spec[9]), int(spec[10])
        path = tmp_path / f"snap_{i + 1:03d}.nc"
        g.write_snap(i, path)
        nrec = -(-nt // tinv)
        comps = {1: range(3), 2: range(3, 9), 3: range(9)}[cmp]
        with netcdf_file(path, "r", mmap=False) as f:
            assert f._recs == nrec
            for c in comps:
                want = g.get_snap(i, c).T.reshape(-1, xn, yn, zn)[:nrec]   # [time][i][j][k]
                assert_same(f.variables[NAMES[c]].data, want, f"snap {i} {NAMES[c]}")
            want_t = np.array([np.float32(s * tinv + 1) * np.float32(sc["dt"]) for s in range(nrec)], np.float32)
            assert_same(f.variables["currTime"].data, want_t, "snap currTime")
            np.testing.assert_array_equal(f.IdxStart, spec[:3])


@pytest.mark.parametrize("compat", [True, False])
def test_peakvel_nc(tmp_path, compat):
    sc, g = _run()
    pv = g.get_pgv()[:, 3:-3, 3:-3]
    path = tmp_path / "peakvel.nc"
    g.write_peakvel(path, vz_from_vy=compat)
    with netcdf_file(path, "r", mmap=False) as f:
        assert_same(f.variables["Vx"].data, pv[0], "Vx")
        assert_same(f.variables["Vy"].data, pv[1], "Vy")
        assert_same(f.variables["Vz"].data, pv[1] if compat else pv[2], "Vz")
    assert np.abs(pv[2] - pv[1]).max() > 0  # the switch is observable
