"""Parameter-file front end (seisconf.py) on the CPU.

Pins, each chosen so that a plausible slip fails it:
- get_conf semantics (common.cpp:23-66, 286-322): comment pieces, first
  match, column search past short lines, atoi/atof prefixes, 127-byte fgets
  pieces, anchor blocks.
- vmap grid (gridmesh.cpp:439-513): equal-spacing layers are exactly linear
  between their anchors; a flat grid spaced by steph gives identity metrics.
- smooth_zdepth (gridmesh.cpp:1092-1182): starts at zb and ends at za; the
  geometric branch runs monotonically from hb to ha, the ramp branch is two
  arithmetic ramps around a plateau.
- reposition (common.cpp:225-268): a plain triple loop with a strict `<`.
- source time functions (mathfunc.cpp:481-610): closed forms, integrals and
  derivatives.
- angle2moment in float (source.cpp:134-155): the double-precision tensor
  and the pure strike-slip case.
- volume media (mediaparas.cpp): a volume linear in x, y and z is
  reproduced at every cell by the trilinear interpolation.
- the reference's example files parse, and the paths the reference cannot
  run are refused with the reason.
"""
import math
import os

import numpy as np
import pytest

import scenario
import seisconf as S
import seisconf_fixture as FX

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hillmodel_cut")
EXAMPLE = "/root/reference/EXAMPLE/hillmodel"
F32 = np.float32


# ---------------------------------------------------------------------------
# get_conf
# ---------------------------------------------------------------------------
def test_get_conf_semantics(tmp_path):
    p = tmp_path / "a.conf"
    long_line = "pad = " + "x" * 120 + " tail_key = 7\n"  # 'tail_key' starts the second 127-byte piece
    p.write_text("# nx = 99\n"
                 "nx = 12 # comment\n"
                 "nx = 13\n"
                 "short = 1\n"
                 "short = 1 2 3\n"
                 "f = 1.9e3abc\n"
                 "n = 1.9e3\n"
                 "b = 0\n"
                 "c = 00\n"
                 + long_line +
                 "<anchor>\n"
                 "1.5 -2.25\n3e2\n")
    c = S.ConfFile(str(p))
    assert c.int("nx") == 12                      # first match; the '#' line is skipped
    assert c.int("short", 5) == 3                 # a line without enough columns is passed over
    assert c.real("f") == F32(1900.0)             # atof reads the numeric prefix
    assert c.int("n") == 1                        # atoi stops at '.'
    assert c.bool("b") is False and c.bool("c") is True   # only the string "0" is false
    assert c.int("tail_key") == 7
    s = c.chunk("<anchor>")
    np.testing.assert_array_equal(s.reals(3), np.array([1.5, -2.25, 300.0], F32))
    with pytest.raises(S.ConfError):
        s.real()
    with pytest.raises(S.ConfError, match="no value found"):
        c.int("missing")


def test_real_parse_is_float_rounding_of_double(tmp_path):
    # (float)atof: the double nearest the text, then the float nearest it
    p = tmp_path / "b.conf"
    p.write_text("v = 0.1\nw = 16777217\n")
    c = S.ConfFile(str(p))
    assert c.real("v") == F32(0.1) and c.real("w") == F32(16777216.0)


# ---------------------------------------------------------------------------
# vmap grid and smooth_zdepth
# ---------------------------------------------------------------------------
def test_vmap_equal_spacing_is_linear_between_anchors(tmp_path):
    conf = FX.write_set(tmp_path, media=False, gp=(5, 4, 6), eq=(1, 1, 1))
    m = S.read_main(str(conf))
    posx, posy, posz, gp, eq = S.read_vmap(m["grid_filename"], FX.NI, FX.NJ, FX.NK)
    x, y, z, _ = S.vmap_coords(posx, posy, posz, gp, eq, FX.NK)
    L = S.LFD
    zindex = [L + 15, L + 10, L + 6, L]
    for k in range(4):  # every anchor surface lands exactly on its grid row
        np.testing.assert_array_equal(z[L:-L, L:-L, zindex[k]], posz[:, :, k])
    for k in range(3):  # inside a layer the depths are linear in the row index
        seg = z[L:-L, L:-L, zindex[k + 1]:zindex[k] + 1].astype(np.float64)
        lin = np.linspace(seg[..., 0], seg[..., -1], seg.shape[-1], axis=-1)
        np.testing.assert_allclose(seg, lin, rtol=0, atol=1e-3)
    np.testing.assert_array_equal(x[L:-L, L:-L, L + 5], posx)
    # ghosts: linear extrapolation (interpolated_extend)
    np.testing.assert_array_equal(z[L:-L, L:-L, L - 1], (2.0 * z[L:-L, L:-L, L] - z[L:-L, L:-L, L + 1]).astype(F32))


def test_flat_vmap_gives_identity_metrics(tmp_path):
    ni, nj, nk, dh = 6, 5, 9, 100.0
    lines = [f"vmap_dims = {ni} {nj} 2", f"vmap_gridpoints = {nk - 1}", "vmap_equalspacing = 1", "<vmap_anchor>"]
    for i in range(ni):
        for j in range(nj):
            lines.append(f"{i * dh} {j * dh} 0.0 {-(nk - 1) * dh}")
    (tmp_path / "g.dat").write_text("\n".join(lines) + "\n")
    posx, posy, posz, gp, eq = S.read_vmap(str(tmp_path / "g.dat"), ni, nj, nk)
    x, y, z, _ = S.vmap_coords(posx, posy, posz, gp, eq, nk)
    met = scenario.metrics_from_coords(x, y, z, dh)
    v = (slice(3, -3),) * 3
    for name, want in (("xix", 1), ("etay", 1), ("zetaz", 1), ("jac", 1), ("xiy", 0), ("xiz", 0),
                       ("etax", 0), ("etaz", 0), ("zetax", 0), ("zetay", 0)):
        np.testing.assert_array_equal(met[name][v], np.full((ni, nj, nk), want, F32), err_msg=name)


@pytest.mark.parametrize("ha,hb,n", [(60.0, 140.0, 8), (140.0, 60.0, 8), (95.0, 110.0, 12)])
def test_smooth_zdepth_geometric_branch(ha, hb, n):
    # mean spacing l/n between ha and hb: the spacing grows geometrically
    # from hb at the bottom to ha at the top
    za, zb = 0.0, -800.0 if n == 8 else -1230.0
    zz = S.smooth_zdepth(za, zb, ha, hb, n).astype(np.float64)
    assert zz.size == n + 1 and zz[0] == F32(zb)
    assert abs(zz[-1] - za) < 1e-3
    h = np.diff(zz)
    assert (h > 0).all()
    d = np.diff(h)
    assert (d >= -1e-3).all() if ha > hb else (d <= 1e-3).all()
    assert min(ha, hb) - 1e-3 <= h.min() and h.max() <= max(ha, hb) + 1e-3


@pytest.mark.parametrize("ha,hb,n,l", [(40.0, 50.0, 10, 600.0), (90.0, 80.0, 12, 700.0), (90.0, 80.0, 12, 900.0)])
def test_smooth_zdepth_ramp_branch(ha, hb, n, l):
    # mean spacing outside [ha, hb]: linear ramps of n/4 steps at each end
    # around a plateau, with h = hb + k*hb0 at the bottom and ha + k*ha0 at
    # the top, summing to the layer thickness
    zz = S.smooth_zdepth(0.0, -l, ha, hb, n).astype(np.float64)
    assert zz[0] == F32(-l) and abs(zz[-1]) < 1e-3
    h = np.diff(zz)
    q = n // 4
    hb0, ha0 = h[0] - hb, h[-1] - ha
    np.testing.assert_allclose(h[:q], hb + hb0 * np.arange(1, q + 1), atol=1e-3)
    np.testing.assert_allclose(h[n - q:], ha + ha0 * np.arange(q, 0, -1), atol=1e-3)
    np.testing.assert_allclose(h[q:n - q], h[q - 1], atol=1e-3)
    assert (hb0 > 0) == (l / n > hb)


def test_vmap_transition_layer(tmp_path):
    conf = FX.write_set(tmp_path, media=False)  # middle layer not equally spaced
    m = S.read_main(str(conf))
    posx, posy, posz, gp, eq = S.read_vmap(m["grid_filename"], FX.NI, FX.NJ, FX.NK)
    x, y, z, _ = S.vmap_coords(posx, posy, posz, gp, eq, FX.NK)
    for i, j in ((5, 7), (12, 9), (0, 0)):
        col = z[3 + i, 3 + j, 3:-3].astype(np.float64)
        h = np.diff(col)
        assert (h > 0).all()
        assert col[6] == F32(-900.0) and col[10] == F32(-500.0)   # anchors exact
        # the transition layer's spacings stay between those of the layers
        # below (600 m / 6) and above ((top + 500 m) / 5), and sum to 400 m
        ha, hb = (posz[i, j, 0] + 500.0) / 5, 100.0
        lo, hi = min(ha, hb) - 0.05, max(ha, hb) + 0.05
        assert ((h[6:10] >= lo) & (h[6:10] <= hi)).all(), (i, j, h[6:10], ha, hb)
        assert abs(h[6:10].sum() - 400.0) < 1e-3


def test_vmap_checks(tmp_path):
    conf = FX.write_set(tmp_path, media=False, gp=(5, 4, 5))
    m = S.read_main(str(conf))
    with pytest.raises(S.ConfError, match="sum"):
        S.read_vmap(m["grid_filename"], FX.NI, FX.NJ, FX.NK)
    conf = FX.write_set(tmp_path / "b", media=False, eq=(0, 1, 1))
    m = S.read_main(str(conf))
    with pytest.raises(S.ConfError, match="first layer and last layer"):
        S.read_vmap(m["grid_filename"], FX.NI, FX.NJ, FX.NK)


# ---------------------------------------------------------------------------
# point location
# ---------------------------------------------------------------------------
def _reposition_loops(px, py, pz, x, y, z, hyp):
    """common::reposition as the plain loops of the source."""
    L = 3
    ni, nj, nk = (s - 6 for s in x.shape)
    dist, loc = F32(1e20), (0, 0, 0)
    px, py, pz = F32(px), F32(py), F32(pz)
    if pz > hyp:
        kz = L + nk - 1
        for i in range(L, L + ni):
            for j in range(L, L + nj):
                p = F32(F32((x[i, j, kz] - px) * (x[i, j, kz] - px)) + F32((y[i, j, kz] - py) * (y[i, j, kz] - py)))
                if p < dist:
                    dist, loc = p, (i, j, kz)
        return loc + (z[loc],)
    for i in range(L, L + ni):
        for j in range(L, L + nj):
            for k in range(L, L + nk):
                a = F32((x[i, j, k] - px) * (x[i, j, k] - px))
                b = F32((y[i, j, k] - py) * (y[i, j, k] - py))
                c = F32((z[i, j, k] - pz) * (z[i, j, k] - pz))
                p = F32(F32(a + b) + c)
                if p < dist:
                    dist, loc = p, (i, j, k)
    return loc + (pz,)


def test_reposition_matches_loops(tmp_path):
    conf = FX.write_set(tmp_path, media=False)
    m = S.read_main(str(conf))
    x, y, z, _ = S.vmap_coords(*S.read_vmap(m["grid_filename"], FX.NI, FX.NJ, FX.NK), FX.NK)
    rng = np.random.default_rng(3)
    pts = [(rng.uniform(0, 2300), rng.uniform(0, 1900), rng.uniform(-1500, 100)) for _ in range(6)]
    pts += [(1200.0, 900.0, 9e6), (50.0, 50.0, 9e6),       # surface points
            (150.0, 250.0, -900.0)]                        # equidistant in x and y: ties go to the first index
    for p in pts:
        got = S.reposition(*p, x, y, z, F32(8e6))
        want = _reposition_loops(*p, x, y, z, F32(8e6))
        assert got[:3] == want[:3] and F32(got[3]) == F32(want[3]), p


# ---------------------------------------------------------------------------
# source time functions and mechanisms
# ---------------------------------------------------------------------------
def test_stf_closed_forms():
    t0, fc, a = 0.3, 2.5, 0.05
    # Ricker at its centre: (0 - 1/2) * 1 * sqrt(pi)/2
    assert S._stf_value("Ricker", 0.3, t0, fc) == F32(-math.sqrt(math.pi) / 4)
    # every function is zero before t = 0
    for nm in ("Gauss", "gauss", "Ricker", "ricker", "Step"):
        assert S._stf_value(nm, -0.01, t0, fc) == 0.0
    t = np.linspace(0.0, 0.6, 60001)
    dt = t[1] - t[0]
    G = np.array([S._stf_value("Gauss", v, t0, a) for v in t], np.float64)
    assert abs(G.sum() * dt - 1.0) < 1e-4                          # unit area
    g = np.array([S._stf_value("gauss", v, t0, a) for v in t[::50]], np.float64)
    dG = np.gradient(G, dt)[::50]
    assert np.abs(g - dG)[1:-1].max() < 5e-3 * np.abs(dG).max()    # gauss is d/dt Gauss (away from the cut-offs)
    R = np.array([S._stf_value("Ricker", v, t0, fc) for v in t], np.float64)
    r = np.array([S._stf_value("ricker", v, t0, fc) for v in t[::50]], np.float64)
    dR = np.gradient(R, dt)[::50]
    assert np.abs(r - dR)[1:-1].max() < 5e-3 * np.abs(dR)[1:-1].max()   # ricker is d/dt Ricker (after its jump at t = 0)
    w = 0.2
    B = np.array([S._stf_value("Bell", v, 0.1, w) for v in t], np.float64)
    assert abs(B.sum() * dt - 1.0) < 1e-4
    BB = np.array([S._stf_value("BELL", v, 0.1, w) for v in t], np.float64)
    assert abs(BB[-1] - 1.0) < 1e-7 and np.abs(np.cumsum(B) * dt - BB).max() < 2e-4   # BELL = integral of Bell
    assert S._stf_value("Step", 1e-3, 0, 0) == 1.0


def test_stf_samples_grid_and_flush():
    nt, stept = 20, 0.01
    s = S.stf_samples("gauss", 0.02, 0.05, 0.01, nt, stept, moment=True)
    assert s.shape == (2 * (nt + 1),)
    # sample i sits at t = i*stept/2 - start in float
    i = 11
    tt = F32(F32(i) * F32(stept / 2.0)) - F32(0.02)
    assert s[i] == S._stf_value("gauss", tt, 0.05, 0.01)
    tiny = S.stf_samples("Gauss", 0.0, 0.5, 0.01, nt, stept, moment=True)
    assert (tiny[np.abs(tiny) > 0] >= 1e-20).all()
    assert (S.stf_samples("Gauss", 0.0, 0.5, 0.01, nt, stept, moment=False) == 0).all()  # zero off (0, 2 t0) anyway


def test_angle2moment_float():
    m = S.angle2moment_f(0.0, 90.0, 0.0)     # vertical strike-slip along x: pure Mxy
    np.testing.assert_allclose(m, [0, 0, 0, 1, 0, 0], atol=1e-7)
    rng = np.random.default_rng(5)
    for _ in range(20):
        a = rng.uniform([0, 0, -180], [360, 90, 180])
        af = [float(F32(v)) for v in a]
        np.testing.assert_allclose(S.angle2moment_f(*a), scenario.angle2moment(*af), rtol=0, atol=2e-6)


# ---------------------------------------------------------------------------
# media, sources, stations, the whole set
# ---------------------------------------------------------------------------
def test_volume_media_reproduces_linear_fields(tmp_path):
    conf = FX.write_set(tmp_path)
    m = S.read_main(str(conf))
    x, y, z, _ = S.vmap_coords(*S.read_vmap(m["grid_filename"], FX.NI, FX.NJ, FX.NK), FX.NK)
    alpha, beta, rho = S.volume_media(m["media_filename"], x, y, z)
    v = (slice(3, -3),) * 3
    vp, vs, den = FX.vol_fields(x[v].astype(np.float64), y[v].astype(np.float64), z[v].astype(np.float64))
    np.testing.assert_allclose(alpha[v], vp, rtol=2e-5)
    np.testing.assert_allclose(beta[v], vs, rtol=2e-5)
    np.testing.assert_allclose(rho[v], den, rtol=2e-5)
    assert (alpha[:3, 3:-3, 3:-3] > 0).all()   # ghosts filled by equivalence_extend


def test_load_full_set(tmp_path):
    conf = FX.write_set(tmp_path)
    sc = S.load(str(conf))
    assert sc["nabs"] == [4, 4, 4, 4, 4, 0] and sc["pgv"] is True
    assert sc["force_locs"].shape == (1, 3) and sc["moment_locs"].shape == (2, 3)
    assert sc["moment_stf"].shape == (2, 2 * (12 + 1))
    # angle-mode moments carry mu = rho*beta*beta of their cell and the m0 factor
    ix, iy, iz = sc["moment_locs"][0]
    mu = F32(F32(sc["rho"][ix, iy, iz] * sc["beta"][ix, iy, iz]) * sc["beta"][ix, iy, iz])
    want = (S.angle2moment_f(30, 60, 90) * F32(1e10)).astype(F32) * mu
    np.testing.assert_array_equal(sc["moment_vec"][0], want.astype(F32))
    # receivers: 3 points + a 3-point line; surface ones on the top row
    assert sc["recv_locs"].shape == (6, 3)
    assert (sc["recv_locs"][[0, 1, 3, 4, 5], 2] == 3 + FX.NK - 1).all()
    np.testing.assert_array_equal(sc["recv_lnum"], [0, 0, 0, 1, 1, 1])
    np.testing.assert_array_equal(sc["recv_pnum"], [0, 1, 2, 0, 1, 2])
    np.testing.assert_array_equal(sc["snapshots"][1], [2, 3, 4, 5, 4, 6, 4, 4, 2, 2, 1])
    # a surface receiver's recorded z is the grid's z there
    x, y, z = sc["coords"]
    assert sc["recv_pos"][0, 2] == z[tuple(sc["recv_locs"][0])]


def test_hybrid_grid_layers(tmp_path):
    """hybrid_grid = 1 (gridmesh.cpp:243-431): the layer holding the
    conversion depth (-1200 m, inside the layer between -1500 and -900 m)
    ends there, a flat layer runs from it to the shallowest point of the
    bottom (-1500 m), and the nk planes are shared by thickness (650, 400,
    300 and 300 m of 1650): int(16*650/1650) = 6, int(16*400/1650) = 3,
    int(16*300/1650) = 2 and the 4 left for the flat layer, so ConIndex =
    3 + 4 and the planes below it lie 75 m apart.  The transition layer
    (equalspacing 0) now spans 3 planes between spacings 650/6 and 150 m."""
    sc = S.load(str(FX.write_set(tmp_path, hybrid=1, conv="-1200.0")))
    assert sc["conindex"] == 3 + 4
    z = sc["coords"][2]
    for k in range(3, 3 + 5):
        np.testing.assert_array_equal(z[3:-3, 3:-3, k], F32(-1500.0 + (k - 3) * 75.0))
    for k in range(3 + 4, 3 + 7):
        np.testing.assert_array_equal(z[3:-3, 3:-3, k], F32(-1200.0 + (k - 7) * 150.0))
    np.testing.assert_array_equal(z[3:-3, 3:-3, 3 + 9], F32(-500.0))
    top = np.vectorize(lambda v: F32(float(f"{v:.6f}")))(FX.surfaces()[2][0])   # as the grid file holds it
    np.testing.assert_array_equal(z[3:-3, 3:-3, 3 + 15], top)
    d = np.diff(z[3:-3, 3:-3, 3 + 6:3 + 10].astype(np.float64), axis=2)
    assert (d > 0).all() and np.isfinite(d).all()
    # the all-curvilinear grid of the same files has no ConIndex entry
    assert "conindex" not in S.load(str(FX.write_set(tmp_path / "c", hybrid=0)))
    # a conversion depth outside every layer
    with pytest.raises(S.ConfError, match="conversion"):
        S.load(str(FX.write_set(tmp_path / "o", hybrid=1, conv="-3000.0")))


def test_refusals(tmp_path):
    conf = FX.write_set(tmp_path / "p", media=False)
    conf.write_text(conf.read_text().replace("media_type = volume", "media_type = interface"))
    with pytest.raises(S.ConfError, match="NaN"):
        S.load(str(conf))
    conf.write_text(conf.read_text().replace("grid_type = vmap", "grid_type = point"))
    with pytest.raises(S.ConfError, match="crd.x"):
        S.load(str(conf))
    (tmp_path / "p" / "vol.nc").write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    conf.write_text(conf.read_text().replace("grid_type = point", "grid_type = vmap")
                    .replace("media_type = interface", "media_type = volume"))
    with pytest.raises(S.ConfError, match="classic" if S._nc4_library() is None else "does not open"):
        S.load(str(conf))


@pytest.mark.skipif(not os.path.isdir(EXAMPLE), reason="reference example not present")
def test_reference_example_files():
    m = S.read_main(f"{EXAMPLE}/SeisFD3D.conf", seispath=EXAMPLE)
    assert (m["ni"], m["nj"], m["nk"], m["nt"]) == (201, 201, 200, 2001)
    assert m["stept"] == F32(0.003) and m["steph"] == F32(50.0) and m["hybrid_grid"] == 1
    posx, posy, posz, gp, eq = S.read_vmap(m["grid_filename"], 201, 201, 200)
    assert gp == [199] and eq == [True]
    x, y, z, _ = S.vmap_coords(posx, posy, posz, gp, eq, 200)
    assert (z[3:-3, 3:-3, 3] == F32(-10000.0)).all()
    np.testing.assert_array_equal(z[3:-3, 3:-3, 202], posz[:, :, 0])
    assert S.read_absorb(m["absorb_filename"])[0] == [12, 12, 12, 12, 12, 0]
    st = S.read_stations(m["station_filename"])
    assert len(st["pos"]) == 26 + 9 and st["snapshots"].shape == (1, 11)
    np.testing.assert_array_equal(st["snapshots"][0], [1, 1, 100, 201, 201, 1, 1, 1, 1, 20, 3])
    # its focal file is netCDF-4 (HDF5): read through libcgfd_nc4.so
    src = S.read_sources(m["source_filename"], EXAMPLE, 10, m["stept"])
    f = src["focal"]
    assert f["nt"] == 1500 and f["dt"] == F32(0.006) and f["series"].shape == (6, 1, 1500)
    np.testing.assert_array_equal(f["pos"], [[5000.0, 5000.0, -5000.0]])
    # the whole example (hybrid grid, conversion depth -3 km) with stand-in
    # media, since its media file is not in the reference repository
    shp = (201 + 6, 201 + 6, 200 + 6)
    media = (np.full(shp, 6000.0, F32), np.full(shp, 3464.0, F32), np.full(shp, 2700.0, F32))
    sc = S.load(f"{EXAMPLE}/SeisFD3D.conf", seispath=EXAMPLE, media=media)
    assert sc["conindex"] > 3 and sc["focal_locs"].shape == (1, 3)
    z = sc["coords"][2]
    k = sc["conindex"] - 1
    assert (z[3:-3, 3:-3, k] == z[3, 3, k]).all()       # flat below the conversion depth


@pytest.mark.skipif(not os.path.isdir(EXAMPLE), reason="reference example not present")
def test_netcdf4_focal_file_equals_its_classic_copy():
    """The example's netCDF-4 focal file read through libhdf5 equals the
    classic copy of it in tests/golden/hillmodel_cut (written by
    make_hillmodel_cut.py), variable by variable."""
    from scipy.io import netcdf_file

    a = S._read_netcdf(f"{EXAMPLE}/stf-GPU.nc", "focal process file")
    with netcdf_file(os.path.join(GOLDEN, "stf-classic.nc"), "r", mmap=False) as b:
        assert a.dimensions["NP"] == b.dimensions["NP"] and a.dimensions["NT"] == b.dimensions["NT"]
        assert a.DT == F32(b.DT)
        for nm in ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz", "posx", "posy", "posz"):
            np.testing.assert_array_equal(a.variables[nm][:], b.variables[nm][:])


def test_hillmodel_cut_fixture_loads():
    """The cut-down copy of the reference example (48 x 48 columns around its
    focal point, its own anchors, hybrid grid at -3 km) loads: the hybrid
    layering of its single layer from the Gaussian hill to -10 km puts 124
    planes below the conversion depth (ConIndex 127), every one flat, and
    the focal point at 5 km depth under the hill's flank."""
    sc = S.load(os.path.join(GOLDEN, "SeisFD3D.conf"), seispath=GOLDEN)
    assert (sc["ni"], sc["nj"], sc["nk"]) == (48, 48, 200) and sc["conindex"] == 127
    z = sc["coords"][2]
    for k in range(3, 127):
        assert (z[3:-3, 3:-3, k] == z[3, 3, k]).all()
    assert z[3, 3, 126] <= F32(-3000.0) and z[3, 3, 127] == F32(-3000.0)
    np.testing.assert_array_equal(sc["focal_locs"], [[3 + 24, 3 + 24, 3 + 89]])
    assert sc["nabs"] == [12, 12, 12, 12, 12, 0] and len(sc["recv_locs"]) == 35


def _moduli(v):
    vp, vs, dc = (np.float64(a) for a in v)
    mu = dc * vs * vs
    return mu, dc * vp * vp - 2 * mu, dc


def test_interface3d_media_layers(tmp_path):
    FX.write_set(tmp_path, media=False)
    conf = FX.write_interface3d(tmp_path)
    sc = S.load(str(conf))
    x, y, z = (c.astype(np.float64) for c in sc["coords"])
    v = (slice(3, -3),) * 3
    X, Y, Z = x[v], y[v], z[v]
    alpha, beta, rho = (sc[k][v].astype(np.float64) for k in ("alpha", "beta", "rho"))
    zp = FX.i3d_plane(X, Y)
    # layer 0: linear in z between the sides of interfaces 0 (z = 400) and 1 (the plane)
    a = Z > zp + 1e-3
    L1 = (Z - zp) / (400.0 - zp)
    for q, got in ((0, alpha), (1, beta), (2, rho)):
        want = FX.ROCK_A_TOP[q] * L1 + FX.ROCK_A_BOT[q] * (1 - L1)
        np.testing.assert_allclose(got[a], want[a], rtol=3e-5)
    # layer 1 and below interface 2
    b = (Z < zp - 1e-3) & (Z > -900.0)
    c = Z < -900.0
    for mask, rock in ((b, FX.ROCK_B), (c, FX.ROCK_C)):
        for q, got in ((0, alpha), (1, beta), (2, rho)):
            np.testing.assert_allclose(got[mask], rock[q], rtol=2e-6)
    # cells on interface 2: mean density, harmonic-mean moduli of the two sides
    on = Z == -900.0
    assert on.sum() == FX.NI * FX.NJ
    mu0, la0, d0 = _moduli(FX.ROCK_B)
    mu1, la1, d1 = _moduli(FX.ROCK_C)
    mu = 1 / (0.5 / mu0 + 0.5 / mu1)
    la = 1 / (0.5 / la0 + 0.5 / la1)
    dc = 0.5 * (d0 + d1)
    np.testing.assert_allclose(rho[on], dc, rtol=1e-6)
    np.testing.assert_allclose(beta[on], np.sqrt(mu / dc), rtol=2e-6)
    np.testing.assert_allclose(alpha[on], np.sqrt((la + 2 * mu) / dc), rtol=2e-6)
    assert a.sum() and b.sum() and c.sum()


def test_interface3d_reads_integer_geometry(tmp_path):
    # interface3dpar's origin and spacing are int fields read with atoi:
    # "-5.0e2" is -5, as in the reference
    FX.write_set(tmp_path, media=False)
    conf = FX.write_interface3d(tmp_path)
    t = S.read_interface3d(str(tmp_path / "media.int3d"), 7, 6, S.ConfFile(str(conf)).int("interface3dpar", 5),
                           -500, 500, 500, 3)
    assert t["X"][0] == -500 and t["X"][1] == 0
    conf.write_text(conf.read_text().replace("interface3dpar = 7 6 -500", "interface3dpar = 7 6 -5.0e2"))
    assert S.ConfFile(str(conf)).int("interface3dpar", 5) == -5
    assert t["flat"] == [True, False, True] and t["layer_const"] == [False, True, True]
    assert (t["indexZ"][:, :, :, 1] == np.arange(1, 4)[:, None, None]).all()
    assert (t["indexZ"][0, :, :, 0] == -1).all()       # air above the top interface


def test_smoothing_mixes_the_row_that_straddles_an_interface(tmp_path):
    # interface 2 at -980 m, between grid rows -900 and -1000 (100 m apart):
    # with 2 samples in z, a flagged cell samples z -/+ 25 m, so only the
    # -1000 m row mixes rock B and rock C; the samples of the -900 m row
    # both lie in rock B
    FX.write_set(tmp_path / "s", media=False)
    sm = S.load(str(FX.write_interface3d(tmp_path / "s", z2=-980.0, sampling=(2, 2, 2))))
    FX.write_set(tmp_path / "n", media=False)
    un = S.load(str(FX.write_interface3d(tmp_path / "n", z2=-980.0)))
    z = sm["coords"][2].astype(np.float64)
    v = (slice(3, -3),) * 3
    Z = z[v]
    row = np.isclose(Z, -1000.0, atol=1e-3)
    assert row.sum() == FX.NI * FX.NJ
    muB, laB, dB = _moduli(FX.ROCK_B)
    muC, laC, dC = _moduli(FX.ROCK_C)
    mu, la, dc = 2 / (1 / muB + 1 / muC), 2 / (1 / laB + 1 / laC), (dB + dC) / 2
    a, b, r = (sm[k][v].astype(np.float64) for k in ("alpha", "beta", "rho"))
    np.testing.assert_allclose(r[row], dc, rtol=1e-6)
    np.testing.assert_allclose(b[row], np.sqrt(mu / dc), rtol=2e-6)
    np.testing.assert_allclose(a[row], np.sqrt((la + 2 * mu) / dc), rtol=2e-6)
    # unsmoothed, that row is rock C; rows far from both interfaces are unchanged
    np.testing.assert_allclose(un["alpha"][v][row], FX.ROCK_C[0], rtol=2e-6)
    far = (Z < -1050.0) | np.isclose(Z, -900.0, atol=1e-3)
    for k in ("alpha", "beta", "rho"):
        np.testing.assert_array_equal(sm[k][v][far], un[k][v][far])


def test_smoothing_below_thresholds_changes_nothing(tmp_path):
    FX.write_set(tmp_path / "a")
    base = S.load(str(tmp_path / "a" / "SeisFD3D.conf"))
    FX.write_set(tmp_path / "b")
    conf = tmp_path / "b" / "SeisFD3D.conf"
    conf.write_text(conf.read_text().replace("sampling_point_per_cell = 0 0 0", "sampling_point_per_cell = 3 3 3")
                    .replace("threshold_velocity_jump_of_interface = 100", "threshold_velocity_jump_of_interface = 1e9"))
    sm = S.load(str(conf))
    for k in ("alpha", "beta", "rho"):
        np.testing.assert_array_equal(sm[k], base[k])
    # with every cell flagged, a linear volume stays linear to second order;
    # the bottom row is left out: read_volume_new cuts the volume at the
    # deepest grid z (a volume node here), so samples below it take that
    # node's value
    conf.write_text(conf.read_text().replace("threshold_velocity_jump_of_interface = 1e9",
                                             "threshold_velocity_jump_of_interface = 0"))
    sm = S.load(str(conf))
    v = (slice(3, -3), slice(3, -3), slice(4, -3))
    np.testing.assert_allclose(sm["alpha"][v], base["alpha"][v], rtol=5e-4)
    assert not np.array_equal(sm["alpha"][v], base["alpha"][v])
    bottom = (slice(3, -3), slice(3, -3), 3)
    assert (sm["alpha"][bottom] < base["alpha"][bottom]).all()
