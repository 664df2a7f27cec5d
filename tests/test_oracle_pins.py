"""Pins of the CPU oracle (the parity checker) by closed forms, exact
invariants and physics, independent of the oracle's own formulas.

No golden vectors exist for this path in the reference (no tests, no outputs
are shipped), and the reference kernels cannot be compiled here without
stand-ins for headers the image lacks, so the oracle is pinned by:
  * closed forms of single stages (DRP coefficient sums and moments, Hooke
    and momentum contractions, RK4 amplification),
  * the textbook traction-free closure on a flat surface,
  * exact invariants (x<->y transposition symmetry, translation invariance,
    linearity),
  * physics at tolerance level (P and S speeds, PML absorption).
"""
import math

import numpy as np
import pytest

import oracle
import scenario

DH, DT = 100.0, 0.005
A = {-1: -0.30874, 0: -0.6326, 1: 1.23300, 2: -0.3334, 3: 0.04168}  # typenew.h:52-56
SUM_A = sum(A.values())                      # -6e-5: the scheme is not exactly consistent
MOM_A = sum(k * v for k, v in A.items())     # 0.99998
VP, VS, RHO = 6000.0, 3464.0, 2700.0
MU = RHO * VS * VS
L2M = RHO * VP * VP
LAM = L2M - 2 * MU


def homogeneous(ni, nj, nk, nt, nabs=(0, 0, 0, 0, 0, 0), pml=None):
    x, y, z = scenario.stretched_grid(ni, nj, nk, DH, None, (nk - 1) * DH)
    met = scenario.metrics_from_coords(x, y, z, DH)
    shp = (ni + 6, nj + 6, nk + 6)
    sc = dict(ni=ni, nj=nj, nk=nk, dh=DH, dt=DT, nt=nt, metrics=met,
              alpha=np.full(shp, VP, np.float32), beta=np.full(shp, VS, np.float32),
              rho=np.full(shp, RHO, np.float32), nabs=list(nabs),
              pml=pml if pml is not None else scenario.unit_pml(ni, nj, nk), pgv=False)
    return sc


def new_oracle(sc, double=False, lib=None):
    o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"], double=double, lib=lib)
    o.setup(sc)
    return o


def interior(a, m=8):
    return a[m:-m, m:-m, m:-m]


# ---------------------------------------------------------------------------
# closed forms of one stage
# ---------------------------------------------------------------------------
def test_identity_metrics_of_flat_uniform_grid():
    sc = homogeneous(12, 10, 8, 1)
    s = (slice(3, -3),) * 3
    for k, want in (("xix", 1), ("etay", 1), ("zetaz", 1), ("jac", 1), ("xiy", 0), ("zetax", 0), ("etaz", 0)):
        assert np.all(sc["metrics"][k][s] == want), k


@pytest.mark.parametrize("phase,sign", [(0, 1), (2, -1)])
def test_constant_velocity_gives_coefficient_sum(phase, sign):
    """Vx = c everywhere in the valid box: away from edges every one-sided
    derivative is c*sum(A)/h regardless of the F/B direction, so after
    stage 0 Txx = a*dt*(lam+2mu)*c*sum(A)/h (forward and backward runs of
    the 8-phase schedule), Tyy = Tzz = a*dt*lam*c*sum(A)/h and Txy = Txz =
    a*dt*mu*c*sum(A)/h."""
    ni = nj = nk = 24
    sc = homogeneous(ni, nj, nk, 4)
    o = new_oracle(sc, double=True)
    if phase:
        o.run(phase)
    c = 1e-3
    v = np.zeros(o.shape)
    v[3:-3, 3:-3, 3:-3] = c
    o.set_field(0, v)
    o.run_stages(1)
    alpha = 0.5 * DT
    d = sign * c * SUM_A / DH  # a backward stencil divides by -h
    W = lambda k: interior(o.get_field(k, 1))
    np.testing.assert_allclose(W(3), alpha * L2M * d, rtol=1e-9)
    np.testing.assert_allclose(W(4), alpha * LAM * d, rtol=1e-9)
    np.testing.assert_allclose(W(5), alpha * LAM * d, rtol=1e-9)
    np.testing.assert_allclose(W(6), alpha * MU * d, rtol=1e-9)
    np.testing.assert_allclose(W(7), alpha * MU * d, rtol=1e-9)
    np.testing.assert_allclose(W(8), 0.0, atol=0)
    # the velocity itself is carried over unchanged (no stress gradient)
    np.testing.assert_allclose(W(0), c, rtol=0, atol=0)


@pytest.mark.parametrize("phase,forward", [(0, True), (0, False)])
def test_linear_stress_gives_first_moment(phase, forward):
    pin_linear_stress_first_moment(forward)


def pin_linear_stress_first_moment(forward, lib=None):
    """Txx = s*i (i the x index): the forward x-derivative is
    s*(i*sum(A) + M1)/h and the backward one s*(M1 - i*sum(A))/h with
    M1 = sum(k*A_k); hVx = dTxx/dx / rho.  Stage 0 of step 0 is forward in
    x, stage 1 backward."""
    ni, nj, nk = 30, 20, 20
    sc = homogeneous(ni, nj, nk, 4)
    o = new_oracle(sc, double=True, lib=lib)
    s = 1e4
    t = np.zeros(o.shape)
    i = np.arange(o.shape[0], dtype=np.float64)[:, None, None] - 3
    t[3:-3, 3:-3, 3:-3] = (s * i * np.ones((1, nj, nk)))[3:-3]
    o.set_field(3, t)
    ii = np.arange(8, ni - 8)
    o.run_stages(1)  # stage 0 of step 0: forward in x
    if forward:
        want = 0.5 * DT * s * (ii * SUM_A + MOM_A) / DH / RHO
    else:
        # stage 1 is backward in x; its input stresses equal the initial ones
        # (no velocity gradient existed in stage 0) and FW.Vx = 0
        o.run_stages(1)
        want = 0.5 * DT * s * (MOM_A - ii * SUM_A) / DH / RHO
    vx = o.get_field(0, 1)[ii + 3, 10, 10]
    np.testing.assert_allclose(vx, want, rtol=1e-6)


def test_free_surface_closure_flat():
    """Flat surface, identity metrics: the traction-free condition gives
    dVx/dz = -dVz/dx, dVy/dz = -dVz/dy, dVz/dz = -lam/(lam+2mu)(dVx/dx+dVy/dy)."""
    sc = homogeneous(10, 9, 8, 1)
    o = new_oracle(sc, double=True)
    cx, cy = o.get_cov()
    r = LAM / L2M
    wx = np.array([0, 0, -1, 0, 0, 0, -r, 0, 0])
    wy = np.array([0, 0, 0, 0, 0, -1, 0, -r, 0])
    np.testing.assert_allclose(cx[3:-3, 3:-3], np.broadcast_to(wx, cx[3:-3, 3:-3].shape), atol=1e-15)
    np.testing.assert_allclose(cy[3:-3, 3:-3], np.broadcast_to(wy, cy[3:-3, 3:-3].shape), atol=1e-15)


def test_rk4_amplification_on_auxiliary_ode():
    pin_rk4_amplification()


def pin_rk4_amplification(lib=None):
    """With d = 0 (DB = 0) and a constant APD = g, each auxiliary field obeys
    dA/dt = -g A independently of the wavefield, so one step multiplies it by
    the classical RK4 polynomial 1 + z + z^2/2 + z^3/6 + z^4/24, z = -g dt."""
    ni, nj, nk = 16, 14, 12
    nabs = [3, 3, 3, 3, 3, 0]
    g = 40.0
    pml = scenario.unit_pml(ni, nj, nk)
    for nm in "xyz":
        pml["APD" + nm][:] = g
    sc = homogeneous(ni, nj, nk, 2, nabs, pml)
    o = new_oracle(sc, double=True, lib=lib)
    rng = np.random.default_rng(0)
    start = {}
    for axis in range(3):
        for c in range(9):
            a0 = rng.standard_normal(o.aux_size(axis))
            o.set_aux(axis, c, a0)
            start[axis, c] = a0
    o.run(1)
    z = -g * DT
    amp = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    NX, NY, NZ = ni + 6, nj + 6, nk + 6
    # packed layouts: axis 0 [p][NY][NZ], axis 1 [p][NX][NZ], axis 2 [p][NX][NY];
    # only entries at valid (non-ghost) positions belong to cells
    dims = {0: (NY, NZ), 1: (NX, NZ), 2: (NX, NY)}
    for (axis, c), a0 in start.items():
        u, v = dims[axis]
        got = o.get_aux(axis, c).reshape(-1, u, v)[:, 3:-3, 3:-3]
        want = (a0 * amp).reshape(-1, u, v)[:, 3:-3, 3:-3]
        np.testing.assert_allclose(got, want, rtol=1e-12)


# ---------------------------------------------------------------------------
# exact invariants
# ---------------------------------------------------------------------------
def test_transpose_symmetry_is_exact():
    """Square box, identity metrics, homogeneous, PML only at the bottom,
    explosive source on the diagonal: the solution is symmetric under x<->y
    with Vx<->Vy, Txx<->Tyy, Txz<->Tyz (bit-for-bit, since x and y always
    share their F/B sign and every contraction sums x and y terms with a
    commutative +)."""
    n, nk, nt = 28, 24, 30
    pml = scenario.cfs_pml_tables(n, n, nk, [0, 0, 0, 0, 6, 0], [6000.0] * 6, [2.0] * 6, [6.28] * 6, DH,
                                  scenario.metrics_from_coords(*scenario.stretched_grid(n, n, nk, DH, None,
                                                                                        (nk - 1) * DH), DH))
    sc = homogeneous(n, n, nk, nt, [0, 0, 0, 0, 6, 0], pml)
    sc["moment_locs"] = np.array([[n // 2 + 3, n // 2 + 3, nk - 8 + 3]], np.int32)
    sc["moment_vec"] = np.array([[1e14, 1e14, 1e14, 0, 0, 0]], np.float32)
    sc["moment_stf"] = scenario.stf_table("Ricker", nt, DT, 0.3, 3.0)[None, :]
    o = new_oracle(sc)
    o.run(nt)
    f = [o.get_field(c) for c in range(9)]
    T = lambda a: np.transpose(a, (1, 0, 2))
    pairs = [(0, 1), (3, 4), (7, 8)]
    for a, b in pairs:
        assert np.array_equal(f[a], T(f[b])), (a, b)
    for a in (2, 5, 6):
        assert np.array_equal(f[a], T(f[a])), a
    assert np.abs(f[0]).max() > 0


def test_translation_invariance_is_exact():
    """Away from boundaries the update is the same at every cell: a local
    random state shifted by one cell in x, y or z evolves into the shifted
    result, bit-for-bit."""
    ni, nj, nk = 48, 46, 50
    sc = homogeneous(ni, nj, nk, 2)
    rng = np.random.default_rng(1)
    blob = rng.standard_normal((9, 5, 5, 5))
    base = (24, 22, 20)

    def run_with(offset):
        o = new_oracle(sc)
        for c in range(9):
            a = np.zeros(o.shape, np.float32)
            x0, y0, z0 = (base[i] + offset[i] + 3 for i in range(3))
            a[x0 - 2:x0 + 3, y0 - 2:y0 + 3, z0 - 2:z0 + 3] = blob[c] * (1e-3 if c < 3 else 1e5)
            o.set_field(c, a)
        o.run(2)
        return np.stack([o.get_field(c) for c in range(9)])

    ref = run_with((0, 0, 0))
    assert np.abs(ref).max() > 0
    for ax in range(3):
        off = [0, 0, 0]
        off[ax] = 1
        got = run_with(tuple(off))
        assert np.array_equal(np.roll(ref, 1, axis=1 + ax), got), f"axis {ax}"


def test_source_linearity():
    """Doubling every source amplitude doubles every output (power-of-two
    scaling commutes with rounding away from the subnormal range)."""
    sc = scenario.config_random(24, 20, 16, 30, [4, 3, 3, 4, 4, 0], seed=2)
    o1 = new_oracle(sc)
    sc2 = dict(sc, moment_vec=sc["moment_vec"] * 2, force_vec=sc["force_vec"] * 2)
    o2 = new_oracle(sc2)
    o1.run(30)
    o2.run(30)
    s1, s2 = o1.get_seismo(), o2.get_seismo()
    big = np.abs(s1) > 1e-30 * np.abs(s1).max()
    assert big.any()
    assert np.array_equal(2 * s1[big], s2[big])


# ---------------------------------------------------------------------------
# physics at tolerance level
# ---------------------------------------------------------------------------
def test_focal_rate_is_linear_interpolation_of_the_series():
    """LoadRmom with InterpFocus: with a zero stage input every stencil term
    vanishes, so the stress rate at a focal cell is exactly -m(t)/(h^3 J) with
    m(t) the linear interpolation of the series at the stage time
    (currT + tinc/2)*dt, and zero once t passes the last sample.  numpy's
    interp states the interpolation independently."""
    ni, nj, nk, nt = 12, 10, 8, 10
    sc = homogeneous(ni, nj, nk, nt)
    # vertical stretching by 1.3 makes J = 1.3, so the 1/J factor is pinned too
    sc["metrics"] = scenario.metrics_from_coords(
        *scenario.stretched_grid(ni, nj, nk, DH, None, 1.3 * (nk - 1) * DH), DH)
    loc = np.array([[6 + 3, 5 + 3, 4 + 3]], np.int32)
    fdt, fnt = np.float32(0.0075), 6        # 1.5 dt: stage times hit samples every third half step
    rng = np.random.default_rng(4)
    ser = (rng.standard_normal((6, 1, fnt)) * 1e12).astype(np.float32)
    o = oracle.Oracle(ni, nj, nk, DH, DT, nt)
    o.set_metrics(sc["metrics"])
    o.set_media(sc["alpha"], sc["beta"], sc["rho"])
    o.set_pml(sc["nabs"], sc["pml"])
    o.set_focal(loc, ser, float(fdt))
    o.init()
    zero = np.zeros(o.shape, np.float32)
    x, y, z = loc[0]
    jac = sc["metrics"]["jac"][x, y, z]
    assert abs(jac - 1.3) < 1e-5
    A = np.float32(1.0) / (np.float32(DH) ** 3 * jac)
    tgrid = np.arange(fnt) * float(fdt)
    seen_zero = False
    for step in range(nt):
        for k, tinc in enumerate((0, 1, 1, 2)):
            for which in (0, 1, 2):
                for c in range(9):
                    o.set_field(c, zero, which)
            o.run_stages(1)
            t = float(np.float32((step + tinc * 0.5) * float(np.float32(DT))))
            for m in range(6):
                got = o.get_field(3 + m, 3)[x, y, z]
                if t > tgrid[-1] + 1e-9:
                    assert got == 0, (step, k, m)
                    seen_zero = True
                else:
                    want = -np.interp(t, tgrid, ser[m, 0].astype(np.float64)) * float(A)
                    # float32 interpolation: a few ulps of the larger neighbouring sample
                    tol = 4e-7 * np.abs(ser[m, 0]).max() * float(A)
                    assert abs(got - want) <= tol, (step, k, m, got, want)
    assert seen_zero


def _peak_time(trace, dt):
    i = int(np.argmax(np.abs(trace)))
    if 0 < i < len(trace) - 1:  # parabolic refinement
        a, b, c = np.abs(trace[i - 1:i + 2])
        den = a - 2 * b + c
        return (i + (0.5 * (a - c) / den if den else 0.0)) * dt
    return i * dt


def test_p_and_s_wave_speeds():
    """Peak-arrival moveout between two interior receivers on the x axis:
    explosion -> P speed sqrt((lam+2mu)/rho); Mxy double couple -> S speed
    sqrt(mu/rho) in Vy.  Checks the media enter the Hooke and momentum
    terms correctly (a swapped lambda / lambda+2mu fails)."""
    ni, nj, nk, nt = 80, 40, 50, 380
    dt = 0.004
    nabs = [10, 10, 10, 10, 10, 0]
    met = scenario.metrics_from_coords(*scenario.stretched_grid(ni, nj, nk, DH, None, (nk - 1) * DH), DH)
    pml = scenario.cfs_pml_tables(ni, nj, nk, nabs, [VP] * 6, [2.0] * 6, [math.pi * 4] * 6, DH, met)
    sc = homogeneous(ni, nj, nk, nt, nabs, pml)
    sc["dt"] = dt
    cz = 22
    src = [[15 + 3, nj // 2 + 3, cz + 3]]
    recv = np.array([[15 + d + 3, nj // 2 + 3, cz + 3] for d in (20, 40)], np.int32)
    stf = scenario.stf_table("Ricker", nt, dt, 0.25, 4.0)[None, :]
    for kind, comp, speed in (("P", 0, VP), ("S", 1, VS)):
        mvec = [[1e14, 1e14, 1e14, 0, 0, 0]] if kind == "P" else [[0, 0, 0, 1e14, 0, 0]]
        s = dict(sc, moment_locs=np.array(src, np.int32), moment_vec=np.array(mvec, np.float32),
                 moment_stf=stf, recv_locs=recv)
        o = new_oracle(s)
        o.run(nt)
        tr = o.get_seismo()[comp]
        t1, t2 = _peak_time(tr[0], dt), _peak_time(tr[1], dt)
        v = 20 * DH / (t2 - t1)
        assert abs(v / speed - 1) < 0.05, (kind, v)


def test_pml_absorbs_outgoing_waves():
    """The same explosion with and without the 5-face CFS-PML: with absorbing
    layers the velocity energy left in the box after the waves reach the
    edges is a small fraction of the zero-ghost (reflecting) box's."""
    def energy(nabs):
        ni = nj = nk = 40
        nt = 400
        met = scenario.metrics_from_coords(*scenario.stretched_grid(ni, nj, nk, DH, None, (nk - 1) * DH), DH)
        pml = scenario.cfs_pml_tables(ni, nj, nk, nabs, [VP] * 6, [2.0] * 6, [math.pi * 3] * 6, DH, met) \
            if sum(nabs) else scenario.unit_pml(ni, nj, nk)
        sc = homogeneous(ni, nj, nk, nt, nabs, pml)
        sc["moment_locs"] = np.array([[23, 23, 23]], np.int32)
        sc["moment_vec"] = np.array([[1e14, 1e14, 1e14, 0, 0, 0]], np.float32)
        sc["moment_stf"] = scenario.stf_table("Ricker", nt, DT, 0.2, 5.0)[None, :]
        o = new_oracle(sc)
        o.run(nt)
        return sum(float(np.sum(o.get_field(c).astype(np.float64) ** 2)) for c in range(3))

    # the scheme itself is mildly dissipative, so compare once the direct and
    # converted waves have crossed the box several times (2 s)
    assert energy([8, 8, 8, 8, 8, 0]) < 0.05 * energy([0, 0, 0, 0, 0, 0])
