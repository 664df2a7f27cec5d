"""cgfd_run_batch_static, the allocation-free form of cgfd_run_batch's fused
path, against cgfd_run_batch itself: two calls in a row on the same device,
the second with fewer and smaller grids, every trace, PGV map and wavefield
component equal value for value.  It must pass alone in a fresh process and
after every other test (the fixed array keeps the last call's images)."""
import numpy as np
import pytest

from conftest import assert_same

pytestmark = pytest.mark.gpu

cgfd3d = pytest.importorskip("cgfd3d")
import scenario  # noqa: E402


def _sims(scs):
    out = []
    for sc in scs:
        g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
        g.setup(sc)
        out.append(g)
    return out


def _outputs(g):
    return [g.get_seismo(), g.get_pgv()] + [g.get_field(c) for c in range(9)]


def _case(scs, steps):
    want = _sims(scs)
    assert cgfd3d.batch_fusable(want)
    cgfd3d.run_batch(want, steps)
    got = _sims(scs)
    cgfd3d.run_batch_static(got, steps)
    for g, w, sc in zip(got, want, scs):
        assert g.step == steps
        for k, (a, b) in enumerate(zip(_outputs(g), _outputs(w))):
            assert_same(a, b, f"{sc['name']} output {k}")
        assert np.abs(w.get_seismo()).max() > 0
    for g in got + want:
        g.close()


def test_static_batch_equals_fused_batch_twice():
    nabs = [4, 4, 3, 3, 4, 0]
    big = [scenario.config_random(n, n - 3, 18, 30, nabs, seed=200 + n, shear=False) for n in (40, 52, 33)]
    small = [scenario.config_random(n, n - 2, 14, 24, nabs, seed=300 + n, shear=False) for n in (24, 28)]
    for sc in big + small:
        sc["name"] = f"random-{sc['ni']}x{sc['nj']}x{sc['nk']}"
    _case(big, 30)
    _case(small, 24)


def test_static_batch_rejects_what_it_cannot_run():
    nabs = [3, 3, 3, 3, 3, 0]
    scs = [scenario.config_random(20, 18, 12, 4, nabs, seed=400 + i, shear=False) for i in range(2)]
    sims = _sims(scs)
    with pytest.raises(cgfd3d.CgfdError, match="fusable"):
        cgfd3d.run_batch_static(sims[:1], 2)   # a single grid is not a fused batch
    sims[1].set_variant(1)
    with pytest.raises(cgfd3d.CgfdError, match="fusable"):
        cgfd3d.run_batch_static(sims, 2)
    many = _sims([scenario.config_random(12, 10, 8, 2, [2, 2, 2, 2, 2, 0], seed=500 + i, shear=False)
                  for i in range(65)])
    with pytest.raises(cgfd3d.CgfdError, match="at most 64"):
        cgfd3d.run_batch_static(many, 1)
    for g in sims + many:
        g.close()
