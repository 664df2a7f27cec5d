"""Parity of the gfx950 engine against the CPU oracle.

Every comparison is bit-for-bit (value equality): the engine reproduces the
rounding of each reference operation (FP64 stencil accumulation, FP32
storage, no contraction), so a tolerance of FP32 relative L2 <= 1e-5 is
met with zero difference.  rel_l2 is reported as well.
"""
import numpy as np
import pytest

from conftest import assert_same, rel_l2

pytestmark = pytest.mark.gpu

cgfd3d = pytest.importorskip("cgfd3d")
import oracle  # noqa: E402
import scenario  # noqa: E402


# LDS-staged planes (at the tile width the grid's rows pick, and at 64 x 4 and
# 32 x 8 tiles), direct in-plane loads, split velocity/stress halves
VARIANTS = [0, "0w64", "0w32", 1, 2]


def _pair(sc, double=False, variant=0):
    prec = cgfd3d.FP64 if double else cgfd3d.FP32
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"], precision=prec)
    if isinstance(variant, str):
        g.set_variant(0)
        g.set_tile_width(int(variant[2:]))
    else:
        g.set_variant(variant)
    o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"], double=double)
    if double:
        sc = dict(sc)
        for k in ("alpha", "beta", "rho"):
            sc[k] = sc[k].astype(np.float64)
    g.setup(sc)
    o.setup(sc)
    return g, o


def _compare_state(g, o, what, aux=True):
    for which, name in ((0, "FW"), (1, "W"), (2, "tW")):
        for c in range(9):
            assert_same(g.get_field(c, which), o.get_field(c, which), f"{what} {name}[{c}]")
    if aux:
        for axis in range(3):
            if o.aux_size(axis) <= 0:
                continue
            for which, name in ((0, "FA"), (1, "A"), (2, "tA")):
                for c in range(9):
                    assert_same(g.get_aux(axis, c, which), o.get_aux(axis, c, which), f"{what} aux{axis} {name}[{c}]")


def test_closure_matrices_match():
    sc = scenario.config_random(40, 36, 20, 4, [5, 4, 4, 5, 5, 0], seed=11)
    g, o = _pair(sc)
    gx, gy = g.get_cov()
    ox, oy = o.get_cov()
    s = (slice(3, -3), slice(3, -3))
    assert_same(gx[s], ox[s], "CoVx")
    assert_same(gy[s], oy[s], "CoVy")


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("phase", [0, 1, 2, 3])
def test_random_state_each_stage(phase, variant):
    """Random wavefield and auxiliary state, then one step stage by stage in
    each of the four F/B phases: every intermediate set matches."""
    sc = scenario.config_random(70, 37, 21, 8, [5, 4, 4, 5, 5, 0], seed=100 + phase)
    g, o = _pair(sc, variant=variant)
    if phase:
        g.run(phase)
        o.run(phase)
    rng = np.random.default_rng(phase)
    st = scenario.random_state(g.shape, rng)
    for c in range(9):
        g.set_field(c, st[c])
        o.set_field(c, st[c])
    for axis in range(3):
        n = o.aux_size(axis)
        if n <= 0:
            continue
        for c in range(9):
            a = (rng.standard_normal(n) * (1e-3 if c < 3 else 1e4)).astype(np.float32)
            g.set_aux(axis, c, a)
            # the engine keeps only entries at valid (non-ghost) positions;
            # hand the oracle the same state
            o.set_aux(axis, c, g.get_aux(axis, c, 0))
    for k in range(4):
        g.run_stages(1)
        o.run_stages(1)
        _compare_state(g, o, f"phase {phase} stage {k}")


@pytest.mark.parametrize("variant", VARIANTS)
def test_random_scenario_run(variant):
    """Sources (incl. a doubled surface force and coincident sources),
    receivers at the surface and at depth, PML on five faces, 60 steps."""
    sc = scenario.config_random(50, 44, 30, 60, [6, 5, 4, 6, 6, 0], seed=7)
    g, o = _pair(sc, variant=variant)
    g.run(60)
    o.run(60)
    assert_same(g.get_seismo(), o.get_seismo(), "seismograms")
    assert_same(g.get_pgv()[:, 3:-3, 3:-3], o.get_pgv()[:, 3:-3, 3:-3], "PGV")
    _compare_state(g, o, "after 60 steps")


@pytest.mark.parametrize("variant", VARIANTS)
def test_focal_points_run(variant):
    """Focal points (moment-rate series interpolated at every stage time,
    LoadRmom) next to moment and force sources, one focal point on the cell
    of a moment source; the series ends inside the run."""
    sc = scenario.config_random(44, 38, 26, 48, [5, 4, 4, 5, 5, 0], seed=31, nfoc=4)
    g, o = _pair(sc, variant=variant)
    g.run(48)
    o.run(48)
    assert_same(g.get_seismo(), o.get_seismo(), "seismograms")
    _compare_state(g, o, "after 48 steps")


@pytest.mark.parametrize("seed", [51, 52])
def test_constant_horizontal_metrics_kernel(seed):
    """Topography by vertical stretching only (fixed x/y spacing): the
    default kernel detects the constant xi/eta metrics, loads only the zeta
    metrics, and still matches the oracle bit for bit; forcing the general
    kernel gives the same values."""
    sc = scenario.config_random(41, 35, 22, 30, [5, 4, 4, 5, 5, 0], seed=seed, nfoc=2, shear=False)
    g, o = _pair(sc)
    assert g.metrics_specialized()
    g2 = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g2.set_specialize(False)
    g2.setup(sc)
    assert not g2.metrics_specialized()
    # 6 of the 9 metric loads (24 B per cell-stage) are gone from the byte model
    assert g2.stage_bytes() - g.stage_bytes() == 24 * sc["ni"] * sc["nj"] * sc["nk"]
    g.run(30)
    g2.run(30)
    o.run(30)
    assert_same(g.get_seismo(), o.get_seismo(), "seismograms")
    _compare_state(g, o, "specialised kernel")
    for c in range(9):
        assert_same(g2.get_field(c), g.get_field(c), f"general vs specialised FW[{c}]")


def test_snapshots_match_step_end_fields():
    """Snapshot boxes (SWpick): a full surface plane of velocities every
    step, a strided interior box of all nine components every 3rd step
    (ragged: nt is not a multiple of 3), a corner box reaching into the
    ghost cells (which read zero) with stresses only every 4th step.  The
    expected values are the oracle's step-end field FW at the same steps, in
    the reference's point order k*yn*zn + m*zn + n (seisplot.cpp:513-518)."""
    nt = 13
    sc = scenario.config_random(30, 26, 18, nt, [4, 4, 3, 3, 4, 0], seed=41, nfoc=1)
    ni, nj, nk = sc["ni"], sc["nj"], sc["nk"]
    specs = np.array([[1, 1, nk, ni, nj, 1, 1, 1, 1, 1, 1],
                      [3, 2, 4, 5, 4, 3, 4, 5, 6, 3, 3],
                      [-1, -2, 1, 4, 3, 2, 3, 4, 5, 4, 2]], np.int32)
    sc["snapshots"] = specs
    g, o = _pair(sc)
    fields = {}
    for t in range(nt):
        o.run(1)
        fields[t] = [o.get_field(c, 0) for c in range(9)]
    g.run(nt)
    for i, q in enumerate(specs):
        xs, ys, zs, xn, yn, zn, xi, yi, zi, tinv, cmp = (int(v) for v in q)
        npts, ntime = g.snap_shape(i)
        assert npts == xn * yn * zn and ntime == -(-nt // tinv)
        comps = {1: range(0, 3), 2: range(3, 9), 3: range(9)}[cmp]
        X = xs + 2 + np.arange(xn) * xi
        Y = ys + 2 + np.arange(yn) * yi
        Z = zs + 2 + np.arange(zn) * zi
        for c in comps:
            want = np.zeros((npts, ntime), np.float32)
            for slot in range(ntime):
                f = fields[slot * tinv][c]
                want[:, slot] = f[np.ix_(X, Y, Z)].reshape(-1)
            assert_same(g.get_snap(i, c), want, f"snapshot {i} comp {c}")
        for c in set(range(9)) - set(comps):
            with pytest.raises(cgfd3d.CgfdError):
                g.get_snap(i, c)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("nabs", [[0, 0, 0, 0, 0, 0], [7, 0, 0, 3, 5, 0], [3, 3, 3, 3, 3, 2]])
def test_pml_layouts(nabs, variant):
    sc = scenario.config_random(33, 29, 18, 24, nabs, seed=sum(nabs) + 3)
    g, o = _pair(sc, variant=variant)
    g.run(24)
    o.run(24)
    assert_same(g.get_seismo(), o.get_seismo(), f"seismograms nabs={nabs}")
    _compare_state(g, o, f"nabs={nabs}")


@pytest.mark.parametrize("variant", VARIANTS)
def test_minimum_depth_grid(variant):
    """nk = 4: every row is one of the special top rows or the first below."""
    sc = scenario.config_random(17, 9, 4, 12, [2, 2, 1, 1, 0, 0], seed=5, nmom=1, nfrc=1, duplicate=False,
                                topo_amp=20.0)
    assert np.isfinite(sc["metrics"]["jac"]).all() and (sc["metrics"]["jac"][3:-3, 3:-3, 3:-3] > 0).all()
    g, o = _pair(sc, variant=variant)
    g.run(12)
    o.run(12)
    _compare_state(g, o, "nk=4")


def test_tuning_does_not_change_results():
    sc = scenario.config_random(66, 30, 40, 16, [5, 5, 4, 4, 6, 0], seed=9)
    ref = None
    for variant, zc, rows, tw in [(1, 1, 4, 0), (1, 3, 2, 0), (1, 7, 1, 0), (1, 40, 4, 0), (0, 1, 4, 0),
                                  (0, 5, 4, 0), (0, 40, 4, 0), (0, 1, 4, 64), (0, 5, 4, 64), (0, 40, 4, 64),
                                  (0, 3, 4, 32), (0, 40, 4, 32), (2, 1, 4, 0), (2, 6, 4, 0), (2, 40, 4, 0)]:
        g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
        g.set_variant(variant)
        g.set_tuning(zc, rows)
        g.set_tile_width(tw)
        g.setup(sc)
        g.run(16)
        out = np.stack([g.get_field(c) for c in range(9)])
        if ref is None:
            ref = out
        else:
            assert_same(out, ref, f"variant={variant} zchunks={zc} rows={rows} tile width={tw}")
        g.close()


def test_tile_width_choice_and_refusals():
    """Rows that need over 8 % more lanes in 64-column tiles than in
    32-column ones take 32; a width can be set (0 returns to the automatic
    choice); FP64 runs 64."""
    for ni, want in [(64, 64), (80, 32), (100, 64), (192, 64), (200, 32), (600, 64), (33, 64), (17, 32),
                     (280, 32), (300, 64)]:
        g = cgfd3d.Simulation(ni, 8, 6, 100.0, 0.005, 4)
        assert g.tile_width() == want, ni
        g.set_tile_width(64)
        assert g.tile_width() == 64
        g.set_tile_width(0)
        assert g.tile_width() == want
        with pytest.raises(cgfd3d.CgfdError, match="tile width"):
            g.set_tile_width(48)
        g.close()
    d = cgfd3d.Simulation(80, 8, 6, 100.0, 0.005, 4, precision=cgfd3d.FP64)
    assert d.tile_width() == 64
    with pytest.raises(cgfd3d.CgfdError, match="64-column"):
        d.set_tile_width(32)
    d.close()


@pytest.mark.parametrize("variant", [0, "0w32", 1, 2])
def test_flat_config_full_run(variant):
    """The small flat homogeneous configuration at full length (64^3, 600
    steps): receivers, snapshots every 100 steps and PGV."""
    sc = scenario.config_flat(nt=600)
    g, o = _pair(sc, variant=variant)
    for step in range(100, 601, 100):
        g.run(100)
        o.run(100)
        for c in range(9):
            gf, of = g.get_field(c), o.get_field(c)
            assert rel_l2(gf, of) <= 1e-5
            assert_same(gf, of, f"snapshot step {step} comp {c}")
    gs, os_ = g.get_seismo(), o.get_seismo()
    assert rel_l2(gs, os_) <= 1e-5
    assert_same(gs, os_, "seismograms")
    assert_same(g.get_pgv()[:, 3:-3, 3:-3], o.get_pgv()[:, 3:-3, 3:-3], "PGV")


@pytest.mark.parametrize("variant", [v for v in VARIANTS if v != "0w32"])  # FP64 runs 64-column tiles
def test_fp64_build_matches_fp64_oracle(variant):
    sc = scenario.config_random(40, 30, 24, 30, [5, 4, 4, 5, 5, 0], seed=21, nfoc=2)
    g, o = _pair(sc, double=True, variant=variant)
    g.run(30)
    o.run(30)
    assert_same(g.get_seismo(), o.get_seismo(), "FP64 seismograms")
    for c in range(9):
        assert_same(g.get_field(c), o.get_field(c), f"FP64 field {c}")


def test_fp32_error_against_fp64_reported():
    """FP32 engine vs FP64 engine on the flat configuration (the stated-error
    oracle): the FP32 rounding error stays small."""
    sc = scenario.config_flat(ni=48, nj=48, nk=48, nt=300, npml=8, src_depth=12)
    g32, _ = _pair(sc)
    g64, _ = _pair(sc, double=True)
    g32.run(300)
    g64.run(300)
    err = rel_l2(g32.get_seismo(), g64.get_seismo())
    print(f"FP32 vs FP64 seismogram rel L2 = {err:.3e}")
    assert err < 1e-4
