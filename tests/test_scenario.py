"""Host-side input builders (copara-cgfd3d_amd/scenario.py) against closed
forms: CFS-PML profiles (absorb.cpp:346-800), central-difference metrics
(gridmesh.cpp:1476-1627), source time functions (mathfunc.cpp:481-522) and
the strike/dip/rake moment tensor (source.cpp:134-155)."""
import math

import numpy as np
import pytest

import scenario

DH = 100.0


def flat_metrics(ni, nj, nk):
    return scenario.metrics_from_coords(*scenario.stretched_grid(ni, nj, nk, DH, None, (nk - 1) * DH), DH)


@pytest.mark.parametrize("n", [12, 20])
def test_cfs_reflection_coefficient_from_layer_count(n):
    # absorb.cpp:467; N=12 -> 5.46e-4, N=20 -> exactly 1e-4 ((log10 20 - 1)/log10 2 = 1)
    R = 10 ** (-(math.log10(n) - 1) / math.log10(2) - 3)
    assert {12: 5.46e-4, 20: 1.0e-4}[n] == pytest.approx(R, rel=0.01)


def test_cfs_profiles_closed_form():
    ni, nj, nk = 40, 36, 30
    nabs = [8, 6, 5, 7, 9, 0]
    vel, bmax, amax = 6000.0, 2.0, 2 * math.pi
    t = scenario.cfs_pml_tables(ni, nj, nk, nabs, [vel] * 6, [bmax] * 6, [amax] * 6, DH, flat_metrics(ni, nj, nk))
    nv = {"x": ni, "y": nj, "z": nk}
    for ax, (lo, hi) in zip("xyz", ((nabs[0], nabs[1]), (nabs[2], nabs[3]), (nabs[4], nabs[5]))):
        n = nv[ax]
        for side, N in (("lo", lo), ("hi", hi)):
            if N == 0:
                continue
            L0 = N * DH
            R = 10 ** (-(math.log10(N) - 1) / math.log10(2) - 3)
            d0 = -vel / 2 / L0 * math.log(R) * 3
            for k in range(N):
                # distance from the interior edge: the innermost cell has L = h
                L = (N - k) * DH
                idx = 3 + k if side == "lo" else 3 + n - 1 - k
                r = L / L0
                d, b, a = d0 * r * r, 1 + (bmax - 1) * r * r, amax * (1 - r)
                assert t["B" + ax][idx] == pytest.approx(b, rel=1e-5)
                assert t["DB" + ax][idx] == pytest.approx(d / b, rel=1e-5)
                assert t["APD" + ax][idx] == pytest.approx(a + d / b, rel=1e-5, abs=1e-4)
        inner = slice(3 + lo, 3 + n - hi)
        assert np.all(t["B" + ax][inner] == 1) and np.all(t["DB" + ax][inner] == 0)
        assert np.all(t["APD" + ax][inner] == 0)


def _hill_metric_errors(dh):
    """Max errors of the central-difference metrics of a Gaussian hill
    (h = 1.5 km, sigma = 3 km) meshed by vertical stretching, against the
    analytic derivatives of the mapping, over the same physical box."""
    ni, nj, nk = int(6000 / dh), int(5000 / dh), int(4000 / dh)
    depth = (nk - 1) * dh
    cx, cy, hh, sg = 3000.0, 2500.0, 1500.0, 3000.0
    top = lambda X, Y: hh * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / sg ** 2)
    x, y, z = scenario.stretched_grid(ni, nj, nk, dh, top, depth)
    m = scenario.metrics_from_coords(x, y, z, dh)
    X, Y = np.meshgrid((np.arange(ni + 6) - 3) * dh, (np.arange(nj + 6) - 3) * dh, indexing="ij")
    frac = ((np.arange(nk + 6) - 3) / (nk - 1))[None, None, :]
    T = top(X, Y)[:, :, None]
    zx = T * (-2 * (X[:, :, None] - cx) / sg ** 2) * frac
    zy = T * (-2 * (Y[:, :, None] - cy) / sg ** 2) * frac
    zz = np.broadcast_to((T + depth) / ((nk - 1) * dh), zx.shape)
    s = (slice(3, -3),) * 3
    assert np.all(m["xix"][s] == 1) and np.all(m["etay"][s] == 1)
    assert np.all(m["xiz"][s] == 0) and np.all(m["etaz"][s] == 0)
    # the first valid cell of each face differences against a linearly
    # extrapolated ghost coordinate (common.cpp:67-99), i.e. a first-order
    # one-sided difference; the second-order claim is for the inside
    s = (slice(4, -4),) * 3
    return {
        "jac": np.abs(m["jac"][s] / zz[s] - 1).max(),
        "zetaz": np.abs(m["zetaz"][s] * zz[s] - 1).max(),
        "zetax": np.abs(m["zetax"][s] + (zx / zz)[s]).max(),
        "zetay": np.abs(m["zetay"][s] + (zy / zz)[s]).max(),
    }


def test_metrics_of_gaussian_hill_converge_at_second_order():
    coarse, fine = _hill_metric_errors(100.0), _hill_metric_errors(50.0)
    for k in ("zetax", "zetay"):
        assert coarse[k] < 1e-2
        assert coarse[k] / fine[k] > 3.2, (k, coarse[k], fine[k])  # O(h^2): ~4x per halving
    for k in ("jac", "zetaz"):  # exact up to FP32 rounding: linear in the vertical index
        assert coarse[k] < 1e-5 and fine[k] < 1e-5, (k, coarse[k], fine[k])


def test_ricker_closed_form():
    fc, t0, dt, nt = 2.0, 0.6, 0.005, 400
    s = scenario.stf_table("Ricker", nt, dt, t0, fc)
    t = np.arange(2 * (nt + 1)) * dt / 2
    u = (t - t0) * 2 * math.pi * fc
    want = np.where(t <= 0, 0.0, (u * u / 4 - 0.5) * np.exp(-u * u / 4) * math.sqrt(math.pi) / 2)
    want = np.where(np.abs(want) < 1e-20, 0.0, want)
    assert len(s) == 2 * (nt + 1)
    np.testing.assert_allclose(s, want, atol=2e-6)
    assert s[int(round(2 * t0 / dt))] == pytest.approx(-math.sqrt(math.pi) / 4, rel=1e-5)


def test_angle2moment_double_couple_identities():
    m = scenario.angle2moment(0.0, 90.0, 0.0)
    np.testing.assert_allclose(m, [0, 0, 0, 1, 0, 0], atol=1e-15)  # vertical strike-slip on x: pure Mxy
    rng = np.random.default_rng(0)
    for _ in range(20):
        mxx, myy, mzz, mxy, mxz, myz = scenario.angle2moment(*rng.uniform([0, 0, -180], [360, 90, 180]))
        M = np.array([[mxx, mxy, mxz], [mxy, myy, myz], [mxz, myz, mzz]])
        assert np.trace(M) == pytest.approx(0, abs=1e-12)
        assert np.linalg.det(M) == pytest.approx(0, abs=1e-12)
        assert np.sum(M * M) == pytest.approx(2, rel=1e-12)


def test_ghost_extensions():
    v = np.zeros((10, 9, 8), np.float32)
    v[3:-3, 3:-3, 3:-3] = np.arange(4 * 3 * 2, dtype=np.float32).reshape(4, 3, 2)
    lin = scenario.interpolated_extend(v.copy())
    # linear extrapolation: ghost = 2*edge - mirror
    assert np.array_equal(lin[2, 4, 4], 2 * lin[3, 4, 4] - lin[4, 4, 4])
    mir = scenario.mirror_extend(v.copy())
    assert np.array_equal(mir[2, 3:-3, 3:-3], mir[4, 3:-3, 3:-3])
    assert np.array_equal(mir[-1, 3:-3, 3:-3], mir[-7, 3:-3, 3:-3])
