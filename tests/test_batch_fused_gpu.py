"""Batched launches (cgfd_run_batch on a fusable list): grids advanced by
one launch per kernel and stage give, value for value, the traces, PGV,
snapshots and final state of running each grid alone, with different
sizes, media, PML widths, sources and receiver counts per grid, with
single runs interleaved between batches, in FP64, and with grids that
share one device copy of their medium.  Lists that cannot
share a launch (another step phase or metric mode, a repeated context,
another stage variant, kernel timing) are reported as such and still run
correctly on one stream per grid."""
import numpy as np
import pytest

from conftest import assert_same

pytestmark = pytest.mark.gpu

cgfd3d = pytest.importorskip("cgfd3d")
import scenario  # noqa: E402

NI, NJ, NK = 36, 30, 20
PMLS = ([4, 4, 3, 3, 4, 0], [2, 5, 4, 2, 3, 0], [0, 0, 0, 0, 0, 0])
NRECV = (6, 2, 9)


def _scenarios(nt, n=3, ni=NI, sizes=None, shear=True):
    """n random scenarios; nt is one step count for all or one per scenario."""
    scs = []
    for i in range(n):
        gi, gj, gk = sizes[i] if sizes else (ni, NJ, NK)
        nti = nt[i] if isinstance(nt, (list, tuple)) else nt
        sc = scenario.config_random(gi, gj, gk, nti, PMLS[i % 3], seed=90 + i, nfoc=2 if i == 1 else 0,
                                    nrecv=NRECV[i % 3], shear=shear)
        x0 = 2 + i % 3  # a 4 x 3 x 3 box spread over the grid, every step
        sc["snapshots"] = np.array([[1, 1, gk, gi, gj, 1, 1, 1, 1, 3 + i % 3, 1],
                                    [x0, 2, 2, 4, 3, 3, max(1, (gi - x0) // 3), max(1, (gj - 2) // 2),
                                     max(1, (gk - 2) // 2), 1, 3]], np.int32)
        scs.append(sc)
    return scs


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


def _check_against_single(batch, scs, nt, double=False):
    for n, (b, sc) in enumerate(zip(batch, scs)):
        s = _sim(sc, double)
        s.run(nt)
        want, got = _outputs(s, sc), _outputs(b, sc)
        assert b.step == nt
        for k in want:
            assert_same(got[k], want[k], f"grid {n} {k}")


@pytest.mark.parametrize("chunks,double", [([6, 9, 8], False), ([23], False), ([5, 7], True)])
def test_fused_batch_equals_single_runs(chunks, double):
    nt = sum(chunks)
    scs = _scenarios(nt)
    batch = [_sim(sc, double) for sc in scs]
    assert cgfd3d.batch_fusable(batch)
    for n in chunks:
        cgfd3d.run_batch(batch, n)
        batch[1].get_field(0)  # a read between batches
    _check_against_single(batch, scs, nt, double)


def test_fused_batch_interleaved_with_single_runs():
    """Single runs of every grid (on its own stream) between batched
    launches (on the first grid's stream): the streams wait for each other."""
    scs = _scenarios(20)
    batch = [_sim(sc) for sc in scs]
    cgfd3d.run_batch(batch, 4)
    for b in batch:
        b.run(8)
    assert cgfd3d.batch_fusable(batch)
    cgfd3d.run_batch(batch, 8)
    _check_against_single(batch, scs, 20)


def test_fused_batch_with_graphs_on():
    """A fusable list runs batched even when its grids replay graphs alone."""
    scs = _scenarios(14)
    batch = [_sim(sc) for sc in scs]
    for b in batch:
        b.set_graph(True)
    assert cgfd3d.batch_fusable(batch)
    cgfd3d.run_batch(batch, 14)
    _check_against_single(batch, scs, 14)


def test_graph_replay_around_fused_batches():
    """Grids that replay captured cycles alone, then join a fused batch, then
    replay again: the captured launches read the step from the device
    counter that the batch advanced."""
    scs = _scenarios(20)
    batch = [_sim(sc) for sc in scs]
    for b in batch:
        b.set_graph(True)
        b.run(4)          # captures and replays the phase-0 cycle
    cgfd3d.run_batch(batch, 8)
    for b in batch:
        b.run(8)          # replays again, two cycles
    _check_against_single(batch, scs, 20)


SIZES = [(36, 30, 20), (70, 33, 26), (20, 17, 12), (65, 64, 30), (129, 9, 8)]


@pytest.mark.parametrize("double", [False, True])
def test_fused_batch_mixed_sizes(double):
    """Grids of different sizes in one launch: each finds its workgroups by a
    search over the batch's prefix table and keeps its own z-chunk length."""
    scs = _scenarios(11, n=len(SIZES), sizes=SIZES)
    batch = [_sim(sc, double) for sc in scs]
    assert cgfd3d.batch_fusable(batch)
    cgfd3d.run_batch(batch, 6)
    cgfd3d.run_batch(batch[::-1], 5)  # another leading grid and order
    _check_against_single(batch, scs, 11, double)


def test_fused_batch_scenarios_of_different_lengths():
    """Scenarios with their own step counts: each call advances the
    unfinished ones to the next one's end, and finished ones leave the list
    (so later calls fuse a different set of grids)."""
    nts = [7, 15, 11, 15]
    scs = _scenarios(nts, n=4, sizes=SIZES[:4])
    batch = [_sim(sc) for sc in scs]
    done, active = 0, list(range(len(batch)))
    while active:
        n = min(nts[i] - done for i in active)
        cgfd3d.run_batch([batch[i] for i in active], n)
        done += n
        active = [i for i in active if nts[i] > done]
    for b, sc, nt in zip(batch, scs, nts):
        s = _sim(sc)
        s.run(nt)
        assert b.step == nt
        want, got = _outputs(s, sc), _outputs(b, sc)
        for k in want:
            assert_same(got[k], want[k], f"{sc['name']} {k}")


def test_fused_batch_many_grids():
    scs = _scenarios(9, n=12)
    batch = [_sim(sc) for sc in scs]
    assert cgfd3d.batch_fusable(batch)
    cgfd3d.run_batch(batch, 9)
    _check_against_single(batch, scs, 9)


def test_unfusable_lists_are_reported_and_still_run():
    scs = _scenarios(12)
    batch = [_sim(sc) for sc in scs]
    assert not cgfd3d.batch_fusable(batch[:1])
    assert not cgfd3d.batch_fusable([batch[0], batch[1], batch[0]])
    other = _sim(_scenarios(12, n=1, shear=False)[0])   # constant horizontal metrics
    assert other.metrics_specialized() and not batch[0].metrics_specialized()
    assert not cgfd3d.batch_fusable([batch[0], other])
    batch[2].set_variant(1)
    assert not cgfd3d.batch_fusable(batch)
    batch[2].set_variant(0)
    batch[0].profile(True)
    assert not cgfd3d.batch_fusable(batch)
    batch[0].profile(False)
    assert cgfd3d.batch_fusable(batch)
    batch[0].set_tile_width(32)
    batch[1].set_tile_width(64)
    assert cgfd3d.batch_fusable(batch)  # one stage launch per tile width
    # another step phase: one stream per grid
    batch[0].run(1)
    assert not cgfd3d.batch_fusable(batch)
    cgfd3d.run_batch(batch[1:], 1)
    assert cgfd3d.batch_fusable(batch)
    cgfd3d.run_batch(batch, 11)
    _check_against_single(batch, scs, 12)
    other.close()


@pytest.mark.parametrize("width", [32, 64])
def test_fused_batch_at_each_tile_width(width):
    """Every grid of a fused batch runs at its own tile width (one stage
    launch per width, here one or two); single runs give the same values."""
    nt = 9
    scs = _scenarios(nt, sizes=[(80, 30, 20), (64, 22, 16), (20, 26, 18)])
    batch = [_sim(sc) for sc in scs]
    assert [b.tile_width() for b in batch] == [32, 64, 32]
    batch[1].set_tile_width(width)
    assert cgfd3d.batch_fusable(batch)
    cgfd3d.run_batch(batch, nt)
    _check_against_single(batch, scs, nt)


def test_fused_batch_keeps_each_grids_tuned_chunks():
    """Grids whose z-chunk count was set (cgfd_set_tuning) or timed
    (cgfd_autotune) keep it inside a fused batch, next to a grid on the
    batch's own count (including counts of 1 and of nk, which the batch's
    rule never picks): results equal the single runs."""
    nt = 12
    scs = _scenarios(nt, sizes=[(40, 30, 20), (36, 22, 16), (33, 26, 18)])
    batch = [_sim(sc) for sc in scs]
    batch[0].set_tuning(zchunks=1)
    batch[1].set_tuning(zchunks=16)
    assert cgfd3d.batch_fusable(batch)
    cgfd3d.run_batch(batch, 7)
    cgfd3d.run_batch(batch, nt - 7)
    _check_against_single(batch, scs, nt)
    tuned = [_sim(sc) for sc in scs]
    tuned[2].autotune(trial_steps=2)
    assert cgfd3d.batch_fusable(tuned)
    cgfd3d.run_batch(tuned, nt)
    _check_against_single(tuned, scs, nt)


def _on_one_medium(n, size, nt):
    """n scenarios with their own sources, receivers and PML on the medium
    (metrics and media) of the first."""
    scs = _scenarios(nt, n=n, sizes=[size] * n)
    for sc in scs[1:]:
        for k in ("metrics", "alpha", "beta", "rho"):
            sc[k] = scs[0][k]
    return scs


@pytest.mark.parametrize("n,size", [(8, (36, 30, 20)), (3, (70, 26, 18))])
def test_batch_on_one_shared_medium_equals_single_runs(n, size):
    """Grids that share one device copy of the medium (cgfd_share_medium)
    run in a fused batch with the copies of each tile together; tile counts
    that are and are not multiples of 8 (the interleave's tail).  Each equals
    a single run with its own copy of the medium, value for value, also after
    the context that loaded the medium is closed."""
    nt = 10
    scs = _on_one_medium(n, size, nt)
    batch = [_sim(scs[0])]
    for sc in scs[1:]:
        batch.append(_sim(dict(sc, medium_of=batch[0])))
    assert cgfd3d.batch_fusable(batch)
    cgfd3d.run_batch(batch, 6)
    _check_against_single(batch, scs, 6)
    first = batch.pop(0)
    first.close()  # the others still hold the medium
    cgfd3d.run_batch(batch, nt - 6)
    _check_against_single(batch, scs[1:], nt)


def test_shared_medium_refuses_what_would_change_it():
    scs = _on_one_medium(2, (36, 30, 20), 4)
    a = _sim(scs[0])
    b = cgfd3d.Simulation(36, 30, 20, scs[1]["dh"], scs[1]["dt"], 4)
    b.share_medium(a)
    for g in (a, b):
        with pytest.raises(cgfd3d.CgfdError, match="shared"):
            g.set_metrics(scs[0]["metrics"])
        with pytest.raises(cgfd3d.CgfdError, match="shared"):
            g.set_media(scs[0]["alpha"], scs[0]["beta"], scs[0]["rho"])
    with pytest.raises(cgfd3d.CgfdError, match="before"):
        b.share_medium(a)  # b has them now
    other = cgfd3d.Simulation(36, 30, 21, scs[1]["dh"], scs[1]["dt"], 4)
    with pytest.raises(cgfd3d.CgfdError, match="grid size"):
        other.share_medium(a)
    fp64 = cgfd3d.Simulation(36, 30, 20, scs[1]["dh"], scs[1]["dt"], 4, precision=cgfd3d.FP64)
    with pytest.raises(cgfd3d.CgfdError, match="precision"):
        fp64.share_medium(a)
    empty = cgfd3d.Simulation(36, 30, 20, scs[1]["dh"], scs[1]["dt"], 4)
    fresh = cgfd3d.Simulation(36, 30, 20, scs[1]["dh"], scs[1]["dt"], 4)
    with pytest.raises(cgfd3d.CgfdError, match="no metrics"):
        fresh.share_medium(empty)
    b.close()
    # a holds the medium alone again and may replace it
    a.set_metrics(scs[0]["metrics"])
    for g in (a, other, fp64, empty, fresh):
        g.close()


def test_two_width_groups_each_on_one_shared_medium():
    """A list mixing two media whose rows take different tile widths (80
    cells: 32, 64 cells: 64), listed alternately: one stage launch per width,
    each interleaving its grids' copies of a tile; every grid equals its
    single run."""
    nt = 8
    a = _on_one_medium(2, (80, 30, 20), nt)
    b = _on_one_medium(2, (64, 22, 16), nt)
    ga = [_sim(a[0])]
    ga.append(_sim(dict(a[1], medium_of=ga[0])))
    gb = [_sim(b[0])]
    gb.append(_sim(dict(b[1], medium_of=gb[0])))
    assert [g.tile_width() for g in ga + gb] == [32, 32, 64, 64]
    batch = [ga[0], gb[0], ga[1], gb[1]]
    assert cgfd3d.batch_fusable(batch)
    cgfd3d.run_batch(batch, 5)
    cgfd3d.run_batch(batch, nt - 5)
    _check_against_single(batch, [a[0], b[0], a[1], b[1]], nt)
