"""Graph replay of 4-step cycles (cgfd_set_graph) against eager stepping:
traces, PGV, snapshots and the final state value for value, with runs that
start at every phase of the forward/backward cycle, leftover eager steps,
reads of the state between replays, and a tuning change in mid-run (which
recaptures).  The step number reaches the kernels through the device
counter, so a captured cycle is valid at any step of its phase."""
import numpy as np
import pytest

from conftest import assert_same

pytestmark = pytest.mark.gpu

cgfd3d = pytest.importorskip("cgfd3d")
import scenario  # noqa: E402


def _scenario(nt):
    sc = scenario.config_random(36, 30, 20, nt, [4, 4, 3, 3, 4, 0], seed=81, nfoc=2)
    sc["snapshots"] = np.array([[1, 1, sc["nk"], 36, 30, 1, 1, 1, 1, 3, 1],
                                [2, 2, 2, 4, 3, 5, 5, 6, 3, 1, 3]], np.int32)
    return sc


def _sim(sc, double=False):
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"],
                          precision=cgfd3d.FP64 if double else cgfd3d.FP32)
    if double:
        sc = dict(sc, **{k: sc[k].astype(np.float64) for k in ("alpha", "beta", "rho")})
    g.setup(sc)
    return g


def _outputs(g, sc):
    out = {"seismo": g.get_seismo(), "pgv": g.get_pgv()[:, 3:-3, 3:-3]}
    for i in range(len(sc["snapshots"])):
        for c in range(9):
            try:
                out[f"snap{i}.{c}"] = g.get_snap(i, c)
            except cgfd3d.CgfdError:
                pass
    for c in range(9):
        out[f"field{c}"] = g.get_field(c)
    return out


@pytest.mark.parametrize("chunks,double", [([23], False), ([6, 9, 8], False), ([1, 2, 3, 4, 5, 8], False),
                                           ([6, 9, 8], True)])
def test_graph_replay_equals_eager(chunks, double):
    nt = sum(chunks)
    sc = _scenario(nt)
    e = _sim(sc, double)
    e.run(nt)
    want = _outputs(e, sc)
    g = _sim(sc, double)
    g.set_graph(True)
    for n in chunks:
        g.run(n)
        g.get_field(0)  # eager work between replays
    assert g.step == nt
    got = _outputs(g, sc)
    for k in want:
        assert_same(got[k], want[k], k)


def test_graph_recaptures_after_tuning_change():
    sc = _scenario(16)
    e = _sim(sc)
    e.run(16)
    g = _sim(sc)
    g.set_graph(True)
    g.run(8)
    g.set_tuning(3, 0)   # new argument blocks: the captured cycles are dropped
    g.run(8)
    for c in range(9):
        assert_same(g.get_field(c), e.get_field(c), f"field {c}")
    assert_same(g.get_seismo(), e.get_seismo(), "seismo")


def test_graph_recaptures_after_variant_change():
    # a variant switch in mid-run drops the captured cycles, so the replays
    # after it launch the new variant's kernels (all variants give the same
    # values, so the state must still equal the eager run's)
    sc = _scenario(24)
    e = _sim(sc)
    e.run(24)
    g = _sim(sc)
    g.set_graph(True)
    g.run(8)
    g.set_variant(2)
    g.run(8)
    g.set_variant(1)
    g.run(4)
    g.set_variant(0)
    g.run(4)
    for c in range(9):
        assert_same(g.get_field(c), e.get_field(c), f"field {c}")
    assert_same(g.get_seismo(), e.get_seismo(), "seismo")


def test_batch_with_graphs_equals_single_eager_runs():
    # contexts of different sizes and metric modes (so one stream each)
    # advanced by cgfd_run_batch with graph replay: 2 whole cycles, then 2
    # eager steps
    scs = [scenario.config_random(n, n - 4, 16, 10, [3, 3, 3, 3, 3, 0], seed=90 + n, shear=n != 30)
           for n in (24, 30, 36)]
    want = []
    for sc in scs:
        e = _sim(sc)
        e.run(10)
        want.append((e.get_seismo(), [e.get_field(c) for c in range(9)]))
    sims = [_sim(sc) for sc in scs]
    for g in sims:
        g.set_graph(True)
    assert not cgfd3d.batch_fusable(sims)
    cgfd3d.run_batch(sims, 10)
    for g, (seis, fields) in zip(sims, want):
        assert g.step == 10
        assert_same(g.get_seismo(), seis, "seismo")
        for c in range(9):
            assert_same(g.get_field(c), fields[c], f"field {c}")
