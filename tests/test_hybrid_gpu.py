"""hybrid_grid = 1 on the GPU: below ConIndex the rates take CalWave's
Cartesian contraction (calculate.cu:3910-3956), above it the curvilinear one.
Every stage variant, FP32 and FP64, the fused batch with a different ConIndex
per grid, and a grid whose metrics are not the horizontal constants, each
against the oracle (or single runs) value for value."""
import numpy as np
import pytest

from conftest import assert_same

pytestmark = pytest.mark.gpu

cgfd3d = pytest.importorskip("cgfd3d")
import oracle  # noqa: E402
import scenario  # noqa: E402


def _scenario(seed, conindex, shear=False, nt=30):
    sc = scenario.config_random(40, 34, 22, nt, [4, 4, 3, 3, 4, 0], seed=seed, shear=shear)
    sc["conindex"] = conindex
    return sc


def _pair(sc, double=False, variant=0):
    prec = cgfd3d.FP64 if double else cgfd3d.FP32
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"], precision=prec)
    g.set_variant(variant)
    o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"], double=double)
    if double:
        sc = dict(sc, **{k: sc[k].astype(np.float64) for k in ("alpha", "beta", "rho")})
    g.setup(sc)
    o.setup(sc)
    return g, o


def _compare(g, o, what):
    assert_same(g.get_seismo(), o.get_seismo(), f"{what} traces")
    assert_same(g.get_pgv(), o.get_pgv(), f"{what} PGV")
    for c in range(9):
        assert_same(g.get_field(c), o.get_field(c), f"{what} field {c}")
    for axis in range(3):
        for c in range(9):
            assert_same(g.get_aux(axis, c, 0), o.get_aux(axis, c, 0), f"{what} aux {axis}.{c}")


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_hybrid_run_matches_oracle(variant):
    sc = _scenario(61, 3 + 12)
    g, o = _pair(sc, variant=variant)
    g.run(30)
    o.run(30)
    _compare(g, o, f"hybrid variant {variant}")
    g.close()


def test_hybrid_general_metrics_and_fp64():
    # sheared grid: the general-metrics kernel, whose Cartesian branch uses
    # the loaded xi_x and eta_y
    for double in (False, True):
        sc = _scenario(62, 3 + 9, shear=True, nt=20)
        g, o = _pair(sc, double=double)
        assert not g.metrics_specialized()
        g.run(20)
        o.run(20)
        _compare(g, o, f"hybrid sheared {'FP64' if double else 'FP32'}")
        g.close()


def test_hybrid_changes_results_below_the_conversion_depth():
    sc = _scenario(63, 3 + 12, nt=20)
    g, _ = _pair(sc)
    h, _ = _pair(dict(sc, conindex=3))
    g.run(20)
    h.run(20)
    assert not np.array_equal(g.get_field(5), h.get_field(5))
    g.close()
    h.close()


def test_fused_batch_with_different_conversion_depths():
    scs = [_scenario(70 + i, ci, nt=24) for i, ci in enumerate((3, 3 + 8, 3 + 14))]
    sims = []
    for sc in scs:
        g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
        g.setup(sc)
        sims.append(g)
    assert cgfd3d.batch_fusable(sims)
    cgfd3d.run_batch(sims, 24)
    for g, sc in zip(sims, scs):
        o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
        o.setup(sc)
        o.run(24)
        _compare(g, o, f"batched hybrid conindex {sc['conindex']}")
        g.close()


def test_set_hybrid_rejects_surface_rows():
    sc = _scenario(64, 3, nt=2)
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    for bad in (2, sc["nk"] + 1, sc["nk"] + 3):
        with pytest.raises(cgfd3d.CgfdError):
            g.set_hybrid(bad)
    g.set_hybrid(sc["nk"])
    g.close()


@pytest.mark.parametrize("conindex", [None, 3 + 10])
def test_simplecar_build_matches_oracle(conindex):
    """The reference's SimpleCar build (xi/eta metrics taken as the identity
    in CalWavemix, CalTIMGmix and VelPDcoeff) on a sheared grid, where it
    differs from the curvilinear run: engine (its constant-horizontal-metrics
    kernel) and oracle value for value, in FP32 and FP64."""
    for double in (False, True):
        sc = _scenario(65, 3 if conindex is None else conindex, shear=True, nt=20)
        if conindex is None:
            del sc["conindex"]
        sc["simplecar"] = True
        g, o = _pair(sc, double=double)
        assert g.metrics_specialized()  # the identity xi/eta metrics select the constant-metrics kernel
        g.run(20)
        o.run(20)
        _compare(g, o, f"SimpleCar conindex {conindex} {'FP64' if double else 'FP32'}")
        g.close()


def test_simplecar_must_precede_the_metrics():
    sc = _scenario(66, 3, nt=2)
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g.set_metrics(sc["metrics"])
    with pytest.raises(cgfd3d.CgfdError):
        g.set_simplecar(True)
    g.close()
