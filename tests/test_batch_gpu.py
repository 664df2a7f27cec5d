"""Batched engine and autotuner: contexts advanced together on their own
streams (cgfd_run_batch) give exactly the results of running each alone, and
so does a context after the autotuner picked its z-chunking."""
import numpy as np
import pytest

from conftest import assert_same

pytestmark = pytest.mark.gpu

cgfd3d = pytest.importorskip("cgfd3d")
import oracle  # noqa: E402
import scenario  # noqa: E402


def _sim(sc):
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g.setup(sc)
    return g


def test_batch_equals_single_runs():
    scs = [scenario.config_random(37, 29, 18, 26, [4, 4, 3, 3, 4, 0], seed=71),
           scenario.config_random(52, 41, 23, 30, [5, 4, 4, 5, 5, 0], seed=72, nfoc=2),
           scenario.config_random(20, 17, 9, 22, [2, 2, 2, 2, 2, 0], seed=73, shear=False),
           scenario.config_random(64, 12, 30, 24, [0, 0, 0, 0, 0, 0], seed=74)]
    batch = [_sim(sc) for sc in scs]
    assert not cgfd3d.batch_fusable(batch)   # metric modes differ: one stream per grid
    steps = 20
    cgfd3d.run_batch(batch, steps)
    for sc, b in zip(scs, batch):
        s = _sim(sc)
        s.run(steps)
        assert_same(b.get_seismo(), s.get_seismo(), f"{sc['name']} seismograms")
        for c in range(9):
            assert_same(b.get_field(c), s.get_field(c), f"{sc['name']} FW[{c}]")
    # and one of them against the oracle
    o = oracle.Oracle(scs[1]["ni"], scs[1]["nj"], scs[1]["nk"], scs[1]["dh"], scs[1]["dt"], scs[1]["nt"])
    o.setup(scs[1])
    o.run(steps)
    assert_same(batch[1].get_seismo(), o.get_seismo(), "batched run vs oracle")


def test_batch_rejects_running_past_nt():
    sc = scenario.config_random(20, 16, 10, 5, [2, 2, 2, 2, 2, 0], seed=75)
    g = _sim(sc)
    with pytest.raises(cgfd3d.CgfdError):
        cgfd3d.run_batch([g], 6)


@pytest.mark.parametrize("trial", [2, 4])   # 4: the graph-replay choice is timed too
def test_autotune_keeps_results(trial):
    sc = scenario.config_random(48, 40, 36, 24, [5, 4, 4, 5, 5, 0], seed=76)
    g = _sim(sc)
    z = g.autotune(trial_steps=trial)
    assert 1 <= z <= sc["nk"] // 2
    assert g.step == 0
    g.run(24)
    o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    o.setup(sc)
    o.run(24)
    assert_same(g.get_seismo(), o.get_seismo(), "autotuned seismograms")
    for c in range(9):
        assert_same(g.get_field(c), o.get_field(c), f"autotuned FW[{c}]")
