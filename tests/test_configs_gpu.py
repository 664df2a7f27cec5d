"""Every benchmark configuration through the HIP path at its own size.

- 200x200x150 Gaussian hill (3 layers, double couple, 400 surface
  receivers): 100 steps, every trace sample, the PGV maps and the 9
  wavefield components equal the oracle's.
- 600x600x300 von Karman model (20-cell PML, 200 subfaults, 400 receivers):
  3 steps from a random wavefield and PML state, every wavefield and
  auxiliary value, trace sample and PGV value, at both tile widths.  This is the size whose 32-bit plane byte
  offsets and 150-plane z chunks the small cases never reach.
- A fused batch of 4 uneven hill scenarios drawn from the 96-192 x 96-192 x
  64-128 distribution: 150 steps in one cgfd_run_batch call equal each
  scenario run alone, and the smallest of them equals the oracle.
- The 192x192x128 Gaussian-hill medium of the surrogate, 1024 steps: FD
  data (128x128 PGV window and 32 stations' traces) from a fused batch,
  one scenario of which equals its run alone; the 4-layer, width-64,
  16-mode network with its seismogram head trained for a few epochs; its
  forward pass and the gate, whose fallback answers are the FD answers
  (reduced counts: the full 4096-run training set is a benchmark, not a
  test).

The 64^3 flat configuration runs in full (600 steps) in test_parity_gpu.py.
Value equality is the criterion throughout (stricter than relative L2 1e-5).
"""
import numpy as np
import pytest

from conftest import assert_same

pytestmark = pytest.mark.gpu

cgfd3d = pytest.importorskip("cgfd3d")
import oracle  # noqa: E402
import scenario  # noqa: E402


def _run_pair(sc, steps):
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g.setup(sc)
    g.run(steps)
    g.sync()
    o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    o.setup(sc)
    o.run(steps)
    return g, o


def _compare_outputs(g, o, what, nrecv):
    gs, os_ = g.get_seismo(), o.get_seismo()
    assert gs.shape[1] == nrecv
    assert np.abs(os_).max() > 0, f"{what}: no signal reached the receivers"
    assert_same(gs, os_, f"{what} traces")
    assert_same(g.get_pgv(), o.get_pgv(), f"{what} PGV")
    for c in range(9):
        assert_same(g.get_field(c), o.get_field(c), f"{what} field {cgfd3d.COMPONENTS[c]}")


def test_hill_200x200x150_full_size():
    sc = scenario.config_hill(nt=100)
    assert (sc["ni"], sc["nj"], sc["nk"]) == (200, 200, 150) and len(sc["recv_locs"]) == 400
    g, o = _run_pair(sc, 100)
    _compare_outputs(g, o, "hill 200x200x150", 400)
    g.close()
    o.close()


def test_vonkarman_600x600x300_full_size():
    # The subfaults start rupturing later than 3 steps, so both sides start
    # from the same random wavefield and PML state: every cell, PML layer and
    # surface row is active from the first stage.
    sc = scenario.config_vonkarman(nt=3)
    assert (sc["ni"], sc["nj"], sc["nk"]) == (600, 600, 300) and len(sc["moment_locs"]) == 200
    # the engine at its own tile width for 600-cell rows and at the other one
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g.setup(sc)
    g2 = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g2.set_tile_width(64 if g.tile_width() == 32 else 32)
    g2.setup(sc)
    o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    o.setup(sc)
    rng = np.random.default_rng(600)
    for c, a in enumerate(scenario.random_state(g.shape, rng)):
        g.set_field(c, a)
        g2.set_field(c, a)
        o.set_field(c, a)
        del a
    for axis in range(3):
        n = o.aux_size(axis)
        for c in range(9):
            a = (rng.standard_normal(n) * (1e-3 if c < 3 else 1e4)).astype(np.float32)
            g.set_aux(axis, c, a)
            g2.set_aux(axis, c, a)
            o.set_aux(axis, c, g.get_aux(axis, c, 0))  # the engine keeps the valid positions only
    g.run(3)
    g2.run(3)
    g.sync()
    o.run(3)
    for e, tw in ((g, g.tile_width()), (g2, g2.tile_width())):
        for c in range(9):
            a, b = e.get_field(c), o.get_field(c)
            assert np.isfinite(b).all() and np.abs(b).max() > 0
            assert_same(a, b, f"von Karman field {cgfd3d.COMPONENTS[c]}, tile width {tw}")
            del a, b
        for axis in range(3):
            for c in range(9):
                assert_same(e.get_aux(axis, c, 0), o.get_aux(axis, c, 0), f"von Karman aux {axis}.{c}, tile width {tw}")
        assert_same(e.get_seismo(), o.get_seismo(), f"von Karman traces, tile width {tw}")
        assert_same(e.get_pgv(), o.get_pgv(), f"von Karman PGV, tile width {tw}")
    g.close()
    g2.close()
    o.close()


def test_uneven_batch_full_size_scenarios():
    steps = 150
    scs = scenario.config_scenario_batch(4, seed=11, nt_cap=steps)
    sizes = [(sc["ni"], sc["nj"], sc["nk"]) for sc in scs]
    assert len(set(sizes)) == 4 and all(96 <= s[0] <= 192 and 64 <= s[2] <= 128 for s in sizes)
    sims = []
    for sc in scs:
        g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
        g.setup(sc)
        sims.append(g)
    assert cgfd3d.batch_fusable(sims)
    cgfd3d.run_batch(sims, steps)
    for g, sc in zip(sims, scs):
        e = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
        e.setup(sc)
        e.run(steps)
        assert_same(g.get_seismo(), e.get_seismo(), f"{sc['name']} traces")
        assert_same(g.get_pgv(), e.get_pgv(), f"{sc['name']} PGV")
        for c in range(9):
            assert_same(g.get_field(c), e.get_field(c), f"{sc['name']} field {c}")
        e.close()
    small = min(range(4), key=lambda i: np.prod(sizes[i]))
    o = oracle.Oracle(scs[small]["ni"], scs[small]["nj"], scs[small]["nk"], scs[small]["dh"], scs[small]["dt"],
                      scs[small]["nt"])
    o.setup(scs[small])
    o.run(steps)
    _compare_outputs(sims[small], o, scs[small]["name"], len(scs[small]["recv_locs"]))
    o.close()
    for g in sims:
        g.close()


def test_surrogate_192x192x128_reduced_counts():
    import surrogate

    base = surrogate.base_medium(ni=192, nj=192, nk=128, nt=1024, dt=0.006, npml=10, window=128, stations=32,
                                 depth=(20, 101))
    ptr = surrogate.draw_sources(12, base, seed=11)
    pte = surrogate.draw_sources(4, base, seed=12)
    print("FD data, 16 scenarios of 192x192x128 x 1024 steps", flush=True)
    y, tr = surrogate.run_fd(base, np.concatenate([ptr, pte]), group=16, traces=True)
    assert y.shape == (16, 3, 128, 128) and tr.shape == (16, 32, 3, 1024)
    assert np.isfinite(y).all() and (y > 0).all() and np.abs(tr).max() > 0
    # the fused batch's answer for one scenario is its run alone
    y1, tr1 = surrogate.run_fd(base, pte[:1], group=1, traces=True, graph=False)
    np.testing.assert_array_equal(y1[0], y[12])
    np.testing.assert_array_equal(tr1[0], tr[12])
    s = surrogate.Surrogate(base, width=64, modes=16, layers=4)
    losses = s.fit(ptr, y[:12], epochs=4, batch=8, lr=3e-3, augment=True, traces=tr[:12])
    print("losses", [round(v, 4) for v in losses], flush=True)
    assert all(np.isfinite(losses))
    pred, ind, ptr_tr = s.predict(pte, traces=True)
    assert pred.shape == (4, 3, 128, 128) and ptr_tr.shape == (4, 32, 3, 1024) and ind.shape == (4,)
    assert np.isfinite(pred).all() and np.isfinite(ptr_tr).all() and (ind >= 0).all()
    s.calibrate(pte, y[12:], target_err=0.05)
    s.threshold = float(np.median(ind))
    got, ok = s.answer(pte, fd_group=4)
    assert ok.any() and not ok.all()
    np.testing.assert_array_equal(got[~ok], y[12:][~ok])
    np.testing.assert_array_equal(got[ok], pred[ok])
