"""Pins of the CPU oracle on curvilinear grids, free surfaces, PML splits,
sources and the forward/backward schedule.

The oracle restates the reference kernels expression by expression, so
checking it against another restatement of the same formulas would pin
nothing.  Every check here states what the mathematics fixes instead:

  * the chain rule on an affine (sheared and tilted) grid: a stage input
    linear in the physical coordinates has exact one-sided differences
    (first moment times the gradient plus the stencil's coefficient sum
    times the value), so every rate is Hooke's law / the momentum equation
    of that gradient, with every off-diagonal metric a distinct nonzero
    constant;
  * the traction-free condition on a tilted planar surface: the closure
    matrices must make the stress rate's traction on the surface vanish,
    solved here from the elastic tensor rather than from the reference's
    3x3 blocks;
  * the ADE CFS-PML split: an axis' auxiliary rate is DB * B times the part
    of the total rate that comes from that axis' derivatives, on every row
    including the free surface;
  * the reference's forward/backward table, transcribed literally from
    simulation.cc:264-641;
  * source magnitudes on a grid with J != 1;
  * the free-surface physics of a plane P or S pulse at normal incidence
    (velocity doubling, stress sign reversal) for the traction image.

Each check also runs against copies of the oracle with one planted mistake
each (test_oracle_plants.py) and must fail there.
"""
import math

import numpy as np
import pytest

import oracle
import scenario

DH = 100.0
A = {-1: -0.30874, 0: -0.6326, 1: 1.23300, 2: -0.3334, 3: 0.04168}  # typenew.h:52-56
SUM_A = sum(A.values())
MOM_A = sum(k * v for k, v in A.items())
VP, VS, RHO = 6000.0, 3464.0, 2700.0
MU = RHO * VS * VS
L2M = RHO * VP * VP
LAM = L2M - 2 * MU

# physical position = M @ (i, j, k) * h for valid-relative indices (i, j, k):
# a sheared and tilted grid, every entry nonzero, so every metric
# d(xi, eta, zeta)/d(x, y, z) is a distinct nonzero constant
M_AFFINE = np.array([[1.00, 0.12, -0.08],
                     [0.07, 0.95, 0.11],
                     [0.10, -0.13, 1.20]])

# (Xvec, Yvec, Zvec) of RKite per step % 4 and stage, as simulation.cc
# spells them at lines 264-300, 377-413, 491-527 and 605-641
FB_TABLE = {
    0: [(1, 1, 1), (-1, -1, -1), (1, 1, 1), (-1, -1, -1)],
    1: [(1, 1, -1), (-1, -1, 1), (1, 1, -1), (-1, -1, 1)],
    2: [(-1, -1, -1), (1, 1, 1), (-1, -1, -1), (1, 1, 1)],
    3: [(-1, -1, 1), (1, 1, -1), (-1, -1, 1), (1, 1, -1)],
}

COMPS = ("Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")
SIG = ((3, 6, 7), (6, 4, 8), (7, 8, 5))  # stress component of sigma_ij


# ---------------------------------------------------------------------------
# scenario builders
# ---------------------------------------------------------------------------
def affine_metrics(shape, M):
    Minv = np.linalg.inv(M)
    names = ("xix", "xiy", "xiz", "etax", "etay", "etaz", "zetax", "zetay", "zetaz")
    met = {n: np.full(shape, Minv[i // 3, i % 3]) for i, n in enumerate(names)}
    met["jac"] = np.full(shape, np.linalg.det(M))
    return met


def affine_scenario(ni, nj, nk, nt, nabs=(0, 0, 0, 0, 0, 0), M=M_AFFINE, dt=0.005):
    shp = (ni + 6, nj + 6, nk + 6)
    met = affine_metrics(shp, M)
    if any(nabs):
        pml = scenario.cfs_pml_tables(ni, nj, nk, list(nabs), [VP] * 6, [2.0] * 6, [math.pi * 2] * 6, DH, met)
    else:
        pml = scenario.unit_pml(ni, nj, nk)
    return dict(ni=ni, nj=nj, nk=nk, dh=DH, dt=dt, nt=nt, metrics=met,
                alpha=np.full(shp, VP), beta=np.full(shp, VS), rho=np.full(shp, RHO),
                nabs=list(nabs), pml=pml, pgv=False)


def make_oracle(sc, lib=None, double=True):
    o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"], double=double, lib=lib)
    o.setup(sc)
    return o


def linear_state(ni, nj, nk, M, grads, consts):
    """Nine fields f_c = grads[c] . x + consts[c] on the valid cells (x the
    physical position), zero on the ghosts."""
    i, j, k = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    X = np.stack([i, j, k], -1).astype(np.float64) @ M.T * DH
    out = np.zeros((9, ni + 6, nj + 6, nk + 6))
    for c in range(9):
        out[c, 3:-3, 3:-3, 3:-3] = X @ grads[c] + consts[c]
    return out


def random_linear(seed):
    rng = np.random.default_rng(seed)
    grads = rng.uniform(0.5, 1.5, (9, 3)) * rng.choice([-1, 1], (9, 3))
    grads[:3] *= 1e-4   # velocities: 1/s
    # values large enough that the stencil's coefficient-sum term is about a
    # tenth of the gradient term, so a wrong F/B sign changes every rate
    consts = np.concatenate([rng.uniform(1.0, 2.0, 3) * 2e-4, rng.uniform(1.0, 2.0, 6) * 2.0]) * 0.1 * DH / 6e-5
    consts *= rng.choice([-1, 1], 9)
    return grads, consts


def pml_profile(sc, axis, idx):
    """(APD, B, DB) of valid-relative index idx on axis 0/1/2."""
    nm = "xyz"[axis]
    p = sc["pml"]
    return p["APD" + nm][idx + 3], p["B" + nm][idx + 3], p["DB" + nm][idx + 3]


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
def traction_closure(Minv, lam, mu, B=(1.0, 1.0, 1.0)):
    """Matrices Cx, Cy with d_zeta v = Cx d_xi v + Cy d_eta v making the
    traction of the stress rate on the zeta surface vanish, for the
    stretched gradient G = sum_a (1/B_a) (d_a v) (grad a)^T.  Solved from the
    isotropic elastic tensor column by column."""
    n = Minv[2]

    def traction(G):
        S = lam * np.trace(G) * np.eye(3) + mu * (G + G.T)
        return S @ n

    def op(vec, axis):
        return traction(np.outer(vec, Minv[axis]) / B[axis])

    K = np.stack([op(e, 2) for e in np.eye(3)], 1)       # traction per unit d_zeta v
    Kx = np.stack([op(e, 0) for e in np.eye(3)], 1)
    Ky = np.stack([op(e, 1) for e in np.eye(3)], 1)
    return -np.linalg.solve(K, Kx), -np.linalg.solve(K, Ky)


def rates_of(D, Minv, B, lam, mu, rho):
    """Rates of the nine fields from the one-sided derivatives D[a] (n, 9)
    along xi, eta, zeta: the physical derivative d/dx_j is
    sum_a (d a/d x_j) D[a] / B_a; velocity rates are div(sigma)/rho and stress
    rates Hooke's law of grad v.  Returns (total (n, 9), axis parts
    [3] (n, 9) without the 1/B_a)."""
    def assemble(Da):
        P = [sum(Minv[a, jj] * Da[a] for a in range(3)) for jj in range(3)]  # P[j][:, c]
        out = np.zeros_like(Da[0])
        for i in range(3):
            out[:, i] = sum(P[jj][:, SIG[i][jj]] for jj in range(3)) / rho
        div = sum(P[kk][:, kk] for kk in range(3))
        for i in range(3):
            for jj in range(i, 3):
                out[:, SIG[i][jj]] = lam * div * (i == jj) + mu * (P[jj][:, i] + P[i][:, jj])
        return out

    zero = np.zeros_like(D[0])
    parts = [assemble([D[a] if b == a else zero for b in range(3)]) for a in range(3)]
    total = assemble([D[a] / B[a][:, None] for a in range(3)])
    return total, parts


def one_sided(grads, consts, M, cells, s):
    """D[a] (n, 9): one-sided DRP differences of the linear fields at the
    cells for F/B signs s = (sx, sy, sz): first moment times the exact
    index-space derivative plus sign * coefficient sum * value / h."""
    X = cells.astype(np.float64) @ M.T * DH
    F = X @ grads.T + consts
    dI = grads @ M  # (9, 3): derivative per unit of the computational coordinate
    return [MOM_A * dI[:, a][None, :] + s[a] * SUM_A * F / DH for a in range(3)], F, dI


def cells_of(ilist, jlist, klist):
    i, j, k = np.meshgrid(ilist, jlist, klist, indexing="ij")
    return np.stack([i.ravel(), j.ravel(), k.ravel()], -1)


def close(got, want, scale, tol, what):
    err = np.abs(got - want)
    bad = err > tol * scale
    assert not bad.any(), f"{what}: {bad.sum()} cells off, worst {err.max():.3e} (scale {np.max(scale):.3e})"


# ---------------------------------------------------------------------------
# pins (each a plain function so that planted copies of the oracle can run it)
# ---------------------------------------------------------------------------
def check_linear_stage(o, sc, grads, consts, s, M=M_AFFINE, aux=True):
    """One stage from the linear state: rates of interior cells, of the
    stress on the 2-4 and two-point closure rows, the traction-free surface,
    and the x/y/z auxiliary rates of PML cells."""
    ni, nj, nk = sc["ni"], sc["nj"], sc["nk"]
    Minv = np.linalg.inv(M)
    st = linear_state(ni, nj, nk, M, grads, consts)
    for which in (0, 1):
        for c in range(9):
            o.set_field(c, st[c], which)
        for ax in range(3):
            if o.aux_size(ax) > 0:
                for c in range(9):
                    o.set_aux(ax, c, np.zeros(o.aux_size(ax)), which)
    o.run_stages(1)
    hW = np.stack([o.get_field(c, 3) for c in range(9)])

    def Bof(cells):
        return [np.array([pml_profile(sc, a, int(v))[1] for v in cells[:, a]]) for a in range(3)]

    def at(arr, cells):
        return arr[:, cells[:, 0] + 3, cells[:, 1] + 3, cells[:, 2] + 3].T

    # interior rows: one-sided DRP differences along all three axes
    cells = cells_of(range(3, ni - 3), range(3, nj - 3), range(3, nk - 3))
    D, F, dI = one_sided(grads, consts, M, cells, s)
    B = Bof(cells)
    want, parts = rates_of(D, Minv, B, LAM, MU, RHO)
    scale = rates_of([np.abs(d) for d in D], np.abs(Minv), B, abs(LAM), MU, RHO)[0]
    close(at(hW, cells), want, scale, 1e-9, "interior rates")

    # the two rows under the surface: the velocity's zeta difference is the
    # 2-4 MacCormack (row nk-3) or two-point (row nk-2) closure, both exact
    # for a linear field, so only the exact derivative remains
    cells = cells_of(range(3, ni - 3), range(3, nj - 3), [nk - 3, nk - 2])
    D, F, dI = one_sided(grads, consts, M, cells, s)
    D[2] = D[2].copy()
    D[2][:, :3] = dI[:3, 2][None, :]
    B = Bof(cells)
    want, _ = rates_of(D, Minv, B, LAM, MU, RHO)
    scale = rates_of([np.abs(d) for d in D], np.abs(Minv), B, abs(LAM), MU, RHO)[0]
    close(at(hW, cells)[:, 3:], want[:, 3:], scale[:, 3:], 1e-9, "stress rates of the closure rows")

    # the surface row: the stress rate exerts no traction on the surface
    cells = cells_of(range(3, ni - 3), range(3, nj - 3), [nk - 1])
    got = at(hW, cells)
    n = Minv[2]
    S = np.zeros((len(cells), 3, 3))
    for i in range(3):
        for jj in range(3):
            S[:, i, jj] = got[:, SIG[i][jj]]
    tr = S @ n
    smax = np.abs(S).max(axis=(1, 2)) * np.abs(n).sum()
    assert smax.min() > 0
    close(tr, 0.0, smax[:, None], 1e-12, "surface traction of the stress rate")

    if not aux:
        return
    # ADE split: each axis' auxiliary rate is DB_a * B_a times the part of
    # the total rate that comes from that axis' derivatives (aux inputs are
    # zero).  On the three top rows the velocity rates come from the
    # traction image, whose xi/eta parts are those of the linear field; on
    # the surface the stress parts include the closure's share.
    nabs = sc["nabs"]
    for ax in range(3):
        if o.aux_size(ax) == 0:
            continue
        n_ax = (ni, nj, nk)[ax]
        layer = list(range(3, nabs[2 * ax])) + list(range(n_ax - nabs[2 * ax + 1], n_ax - 3))
        rng_ax = [range(3, ni - 3), range(3, nj - 3), range(3, nk)]
        rng_ax[ax] = layer
        if ax == 2:
            rng_ax[2] = [v for v in layer if v < nk - 3]
        cells = cells_of(*rng_ax)
        if not len(cells):
            continue
        D, F, dI = one_sided(grads, consts, M, cells, s)
        top = cells[:, 2]
        # zeta differences of the velocity on the closure rows are exact
        for row in (nk - 3, nk - 2):
            m = top == row
            D[2][m, :3] = dI[:3, 2][None, :]
        # surface: zeta difference of the velocity from the traction-free closure
        Bc = Bof(cells)
        surf = top == nk - 1
        want_aux = np.zeros((len(cells), 9))
        for idx in np.flatnonzero(surf):
            b = [Bc[a][idx] for a in range(3)]
            Cx, Cy = traction_closure(Minv, LAM, MU, b)
            Dc = [D[a][idx:idx + 1].copy() for a in range(3)]
            Dv = [Dc[0][0, :3], Dc[1][0, :3]]
            zpart = (Cx @ Dv[0]) if ax == 0 else (Cy @ Dv[1]) if ax == 1 else np.zeros(3)
            only = [Dc[a] if a == ax else np.zeros_like(Dc[a]) for a in range(3)]
            only[2] = only[2].copy()
            only[2][0, :3] = zpart if ax < 2 else 0.0
            tot, _ = rates_of(only, Minv, [np.array([b[a]]) for a in range(3)], LAM, MU, RHO)
            want_aux[idx] = tot[0] * b[ax]
        _, parts = rates_of(D, Minv, Bc, LAM, MU, RHO)
        keep = ~surf
        want_aux[keep] = parts[ax][keep]
        DB = np.array([pml_profile(sc, ax, int(v))[2] for v in cells[:, ax]])
        want_aux = want_aux * DB[:, None]
        scale = np.abs(want_aux).max(axis=0)[None, :] + 1e-300
        # packed auxiliary layouts (calculate.cu:325-327)
        NX, NY, NZ = ni + 6, nj + 6, nk + 6
        lo = nabs[2 * ax]
        p = np.where(cells[:, ax] < lo, cells[:, ax], cells[:, ax] - (n_ax - nabs[2 * ax + 1]) + lo)
        if ax == 0:
            lin = (p * NY + cells[:, 1] + 3) * NZ + cells[:, 2] + 3
        elif ax == 1:
            lin = (p * NX + cells[:, 0] + 3) * NZ + cells[:, 2] + 3
        else:
            lin = (p * NX + cells[:, 0] + 3) * NY + cells[:, 1] + 3
        got = np.stack([o.get_aux(ax, c, 3)[lin] for c in range(9)], -1)
        # velocity auxiliaries of the top rows come from the traction image
        # and see only the xi/eta differences: same as the linear parts
        close(got, want_aux, scale, 1e-9, f"auxiliary rates of axis {'xyz'[ax]}")


def pin_affine_linear_stage(lib=None):
    """(a) Chain rule on a sheared, tilted grid, stage 0 of step 0 (all
    forward), with 6-cell CFS-PML layers on five faces."""
    sc = affine_scenario(22, 21, 20, 2, nabs=(6, 6, 6, 6, 6, 0))
    o = make_oracle(sc, lib)
    grads, consts = random_linear(11)
    check_linear_stage(o, sc, grads, consts, FB_TABLE[0][0])


def pin_fb_schedule(lib=None):
    """(e) The forward/backward signs of all 16 (step % 4, stage) cases,
    read off the coefficient-sum term of each rate."""
    sc = affine_scenario(18, 17, 16, 6)
    o = make_oracle(sc, lib)
    grads, consts = random_linear(12)
    for p in range(4):
        for k in range(4):
            assert (o.step % 4, o.stage) == (p, k)
            check_linear_stage(o, sc, grads, consts, FB_TABLE[p][k], aux=False)


def pin_closure_matrices(lib=None):
    """(b) VelPDcoeff on a tilted planar surface with PML stretching: the
    closure matrices equal the traction-free solution."""
    ni, nj, nk = 16, 15, 8
    sc = affine_scenario(ni, nj, nk, 1, nabs=(5, 5, 5, 5, 3, 0))
    o = make_oracle(sc, lib)
    cx, cy = o.get_cov()
    Minv = np.linalg.inv(M_AFFINE)
    seen_stretch = False
    for i in range(ni):
        for j in range(nj):
            b = [pml_profile(sc, 0, i)[1], pml_profile(sc, 1, j)[1], pml_profile(sc, 2, nk - 1)[1]]
            seen_stretch |= b[0] != 1 and b[1] != 1
            Cx, Cy = traction_closure(Minv, LAM, MU, b)
            gx = cx[i + 3, j + 3].reshape(3, 3)
            gy = cy[i + 3, j + 3].reshape(3, 3)
            s = max(np.abs(Cx).max(), np.abs(Cy).max())
            assert np.abs(gx - Cx).max() <= 1e-12 * s, (i, j, gx, Cx)
            assert np.abs(gy - Cy).max() <= 1e-12 * s, (i, j, gy, Cy)
    assert seen_stretch


def pin_source_magnitudes(lib=None):
    """(d) With a zero state, a vertically stretched grid (J = 1.3) and
    sources of known tables: the stress rate at a moment source is
    -M stf / (h^3 J), the velocity rate at a force is f stf / (h^3 J rho),
    doubled on the surface row, for every stage's table index."""
    ni, nj, nk, nt = 12, 11, 10, 3
    s_ = 1.3
    M = np.diag([1.0, 1.0, s_])
    sc = affine_scenario(ni, nj, nk, nt, M=M)
    rng = np.random.default_rng(5)
    nstf = 2 * nt + 1
    sc["moment_locs"] = np.array([[5 + 3, 4 + 3, 3 + 3]], np.int32)
    sc["moment_vec"] = rng.standard_normal((1, 6)) * 1e15
    sc["moment_stf"] = rng.uniform(0.5, 2.0, (1, nstf))
    sc["force_locs"] = np.array([[4 + 3, 6 + 3, 2 + 3], [7 + 3, 5 + 3, nk - 1 + 3]], np.int32)
    sc["force_vec"] = rng.standard_normal((2, 3)) * 1e12
    sc["force_stf"] = rng.uniform(0.5, 2.0, (2, nstf))
    o = make_oracle(sc, lib)
    zero = np.zeros(o.shape)
    J = s_
    for step in range(nt):
        for k, tinc in enumerate((0, 1, 1, 2)):
            for which in (0, 1, 2):
                for c in range(9):
                    o.set_field(c, zero, which)
            o.run_stages(1)
            t = 2 * step + tinc
            x, y, z = sc["moment_locs"][0]
            got = np.array([o.get_field(3 + m, 3)[x, y, z] for m in range(6)])
            want = -sc["moment_vec"][0] * sc["moment_stf"][0, t] / (DH ** 3 * J)
            np.testing.assert_allclose(got, want, rtol=1e-13, err_msg=f"moment, step {step} stage {k}")
            for f in range(2):
                x, y, z = sc["force_locs"][f]
                got = np.array([o.get_field(c, 3)[x, y, z] for c in range(3)])
                want = sc["force_vec"][f] * sc["force_stf"][f, t] / (DH ** 3 * J * RHO)
                if z == nk - 1 + 3:
                    want = want * 2
                np.testing.assert_allclose(got, want, rtol=1e-13, err_msg=f"force {f}, step {step} stage {k}")


def pin_aux_enters_rate_over_b(lib=None):
    """With a zero wavefield and random auxiliary inputs, every rate is
    -sum_a A_a / B_a over the axes whose layer holds the cell."""
    ni, nj, nk = 14, 13, 12
    nabs = (4, 4, 4, 4, 4, 0)
    sc = affine_scenario(ni, nj, nk, 2, nabs=nabs)
    o = make_oracle(sc, lib)
    rng = np.random.default_rng(6)
    NX, NY, NZ = ni + 6, nj + 6, nk + 6
    auxv = {}
    for ax in range(3):
        for c in range(9):
            v = rng.standard_normal(o.aux_size(ax))
            o.set_aux(ax, c, v, 0)
            auxv[ax, c] = v
    o.run_stages(1)
    # corner column: x and y layers; rows nk-4 and below are clear of the
    # traction image, row 1 is in the bottom (z) layer too
    for (i, j, k) in [(1, 2, 1), (1, 2, 6), (ni - 2, 5, 2), (6, nj - 1, 7), (6, 6, 2)]:
        want = np.zeros(9)
        for ax, (idx, n_ax) in enumerate(((i, ni), (j, nj), (k, nk))):
            lo, hi = nabs[2 * ax], nabs[2 * ax + 1]
            if not (idx < lo or idx >= n_ax - hi):
                continue
            p = idx if idx < lo else idx - (n_ax - hi) + lo
            lin = ((p * NY + j + 3) * NZ + k + 3 if ax == 0 else
                   (p * NX + i + 3) * NZ + k + 3 if ax == 1 else (p * NX + i + 3) * NY + j + 3)
            B = pml_profile(sc, ax, idx)[1]
            want -= np.array([auxv[ax, c][lin] for c in range(9)]) / B
        got = np.array([o.get_field(c, 3)[i + 3, j + 3, k + 3] for c in range(9)])
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=0, err_msg=str((i, j, k)))


def _plane_pulse(lib, kind):
    """Laterally uniform plane P (kind 'P') or S (kind 'S') pulse travelling
    up a homogeneous half-space meshed with vertical stretching 1.25.
    Returns the centre column's surface velocity and its stress 5 cells
    below the surface over time, and the incident amplitudes.  The pulse
    starts 10 cells below the surface; waves from the lateral edges need
    0.8 s to reach the centre column, after the last sample used."""
    ni = nj = 96
    nk, dt = 30, 0.005
    nt = 90 if kind == "P" else 150
    s_ = 1.25
    M = np.diag([1.0, 1.0, s_])
    sc = affine_scenario(ni, nj, nk, nt, M=M, dt=dt)
    c = VP if kind == "P" else VS
    zc, w = (nk - 1 - 10) * DH * s_, 2.5 * DH * s_
    k = np.arange(nk) * DH * s_
    f = np.exp(-((k - zc) / w) ** 2)
    amp = 1e-3
    st = np.zeros((9, ni + 6, nj + 6, nk + 6))
    # upgoing wave in +z: sigma = -rho c v on the pulse (sigma_zz for P,
    # sigma_xz for S), and the lateral P stresses lambda/(lambda+2mu) of it
    if kind == "P":
        st[2, 3:-3, 3:-3, 3:-3] = amp * f
        st[5, 3:-3, 3:-3, 3:-3] = -RHO * c * amp * f
        st[3, 3:-3, 3:-3, 3:-3] = st[4, 3:-3, 3:-3, 3:-3] = -LAM / c * amp * f
        vcomp, scomp = 2, 5
    else:
        st[0, 3:-3, 3:-3, 3:-3] = amp * f
        st[7, 3:-3, 3:-3, 3:-3] = -RHO * c * amp * f
        vcomp, scomp = 0, 7
    ctr = (ni // 2 + 3, nj // 2 + 3)
    sc["recv_locs"] = np.array([[ctr[0], ctr[1], nk - 1 + 3], [ctr[0], ctr[1], nk - 6 + 3]], np.int32)
    o = make_oracle(sc, lib)
    for cc in range(9):
        o.set_field(cc, st[cc], 0)
    o.run(nt)
    seis = o.get_seismo()
    return seis[vcomp, 0], seis[scomp, 1], amp, RHO * c * amp


def pin_free_surface_plane_waves(lib=None):
    """(c) The traction image at normal incidence: the surface velocity of an
    upgoing plane P or S pulse peaks at twice the incident amplitude, and
    the reflected stress pulse comes back with the opposite sign (free
    surface: reflection coefficient +1 for velocity, -1 for traction),
    before any wave from the lateral edges reaches the centre column."""
    for kind in ("P", "S"):
        vs, ts, amp, tamp = _plane_pulse(lib, kind)
        assert abs(vs.max() / (2 * amp) - 1) < 0.03, (kind, vs.max() / amp)
        i_in = int(np.argmin(ts))     # incident: negative stress pulse
        i_out = int(np.argmax(ts))    # reflected: positive, later
        assert i_out > i_in, kind
        assert abs(ts.min() / tamp + 1) < 0.05, (kind, ts.min() / tamp)
        assert abs(ts.max() / tamp - 1) < 0.05, (kind, ts.max() / tamp)


def pin_outputs_are_step_end_fields(lib=None):
    """Receivers record FW after each step and PGV is the running maximum of
    |V| on the surface row over the step-end fields."""
    sc = scenario.config_random(16, 14, 10, 12, [3, 3, 3, 3, 3, 0], seed=7)
    sc["recv_locs"] = np.array([[5, 6, 12], [9, 10, 7], [12, 4, 12]], np.int32)
    o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"], double=True, lib=lib)
    o.setup(sc)
    nk = sc["nk"]
    pgv = np.zeros((3, sc["ni"] + 6, sc["nj"] + 6))
    trace = np.zeros((9, 3, sc["nt"]))
    for t in range(sc["nt"]):
        o.run(1)
        f = [o.get_field(c, 0) for c in range(9)]
        for r, (x, y, z) in enumerate(sc["recv_locs"]):
            trace[:, r, t] = [f[c][x, y, z] for c in range(9)]
        for c in range(3):
            top = np.abs(f[c][:, :, nk + 2])
            pgv[c, 3:-3, 3:-3] = np.maximum(pgv[c, 3:-3, 3:-3], top[3:-3, 3:-3])
    assert np.abs(trace).max() > 0
    assert np.array_equal(o.get_seismo(), trace)
    got = o.get_pgv()
    assert np.array_equal(got[:, 3:-3, 3:-3], pgv[:, 3:-3, 3:-3])


M_DIAG = np.diag([1.15, 0.9, 1.3])


def pin_hybrid_cartesian_branch(lib=None):
    """(g) hybrid_grid = 1: below ConIndex CalWave uses a Cartesian form of
    the contraction (calculate.cu:3910-3956).  On a grid stretched along the
    three axes (diagonal metrics, as the hybrid mesh builds below the
    conversion depth) it is exact, so every rate of a linear state, stress
    and velocity, and every PML auxiliary rate equals the chain-rule closed
    form; ConIndex puts all rows below the top six in that branch."""
    sc = affine_scenario(22, 21, 20, 2, nabs=(6, 6, 6, 6, 6, 0), M=M_DIAG)
    sc["conindex"] = 3 + 20 - 6
    o = make_oracle(sc, lib)
    grads, consts = random_linear(13)
    check_linear_stage(o, sc, grads, consts, FB_TABLE[0][0], M=M_DIAG)


def pin_hybrid_equals_curvilinear_on_identity_metrics(lib=None):
    """On identity metrics the Cartesian and the curvilinear contraction
    give the same values term by term (products with 1 and sums with zeros
    are exact), so a hybrid run equals the all-curvilinear run value for
    value, PML, surface and sources included."""
    sc = scenario.config_random(24, 22, 18, 8, [4, 4, 3, 3, 4, 0], seed=31, topo_amp=0.0, shear=False)
    a = make_oracle(sc, lib, double=True)
    sc2 = dict(sc, conindex=3 + 12)
    b = make_oracle(sc2, lib, double=True)
    rng = np.random.default_rng(31)
    for c, st in enumerate(scenario.random_state(a.shape, rng)):
        a.set_field(c, st.astype(np.float64))
        b.set_field(c, st.astype(np.float64))
    a.run(8)
    b.run(8)
    for c in range(9):
        x, y = a.get_field(c), b.get_field(c)
        assert np.abs(x).max() > 0
        assert np.array_equal(x, y), f"{COMPS[c]}: {np.count_nonzero(x != y)} values differ"


XI_ETA_IDENTITY = (("xix", 1.0), ("xiy", 0.0), ("xiz", 0.0), ("etax", 0.0), ("etay", 1.0), ("etaz", 0.0))


def _run_pair(sa, sb, lib, steps=8, seed=41):
    """Two FP64 oracles from the same random state, run `steps` steps."""
    a, b = make_oracle(sa, lib, double=True), make_oracle(sb, lib, double=True)
    rng = np.random.default_rng(seed)
    for c, st in enumerate(scenario.random_state(a.shape, rng)):
        a.set_field(c, st.astype(np.float64))
        b.set_field(c, st.astype(np.float64))
    a.run(steps)
    b.run(steps)
    return a, b


def _all_equal(a, b):
    out = [np.array_equal(a.get_field(c), b.get_field(c)) for c in range(9)]
    out += [np.array_equal(a.get_aux(ax, c, 0), b.get_aux(ax, c, 0)) for ax in range(3) for c in range(9)]
    out += [np.array_equal(a.get_seismo(), b.get_seismo()), np.array_equal(a.get_pgv(), b.get_pgv())]
    return all(out)


def pin_simplecar_is_the_identity_xi_eta_metrics(lib=None):
    """(h) The SimpleCar build restates CalWavemix (calculate.cu:4339-4456),
    CalTIMGmix (5729-5740) and VelPDcoeff's constants (3369-3375).  On a
    sheared grid, whose xi/eta metrics are far from the identity, it equals
    value for value the curvilinear run of the same grid with its xi/eta
    metrics replaced by 1 and 0 (products with 1 and sums with zeros are
    exact), with and without a hybrid part, PML, surface and sources; and
    it differs from the run with the grid's own metrics."""
    for conindex in (None, 3 + 8):
        sc = scenario.config_random(24, 22, 18, 8, [4, 4, 3, 3, 4, 0], seed=41, shear=True)
        sc = {k: (v.astype(np.float64) if k in ("alpha", "beta", "rho") else v) for k, v in sc.items()}
        if conindex:
            sc["conindex"] = conindex
        met = dict(sc["metrics"])
        for nm, v in XI_ETA_IDENTITY:
            met[nm] = np.full(met[nm].shape, v, met[nm].dtype)
        a, b = _run_pair(dict(sc, simplecar=True), dict(sc, metrics=met), lib)
        assert np.abs(a.get_field(3)).max() > 0
        assert _all_equal(a, b), f"SimpleCar vs identity xi/eta metrics, conindex {conindex}"
        _, c = _run_pair(dict(sc, simplecar=True), sc, lib)
        assert not np.array_equal(a.get_field(3), c.get_field(3))


def pin_simplecar_equals_curvilinear_on_stretched_grids(lib=None):
    """On a grid stretched only vertically under a topography (uniform x/y
    spacing, as every vmap grid), the discrete xi/eta metrics are exactly 1
    and 0, so the SimpleCar build and the curvilinear run agree value for
    value, hybrid or not."""
    for conindex in (None, 3 + 8):
        sc = scenario.config_random(24, 22, 18, 8, [4, 4, 3, 3, 4, 0], seed=43, shear=False)
        sc = {k: (v.astype(np.float64) if k in ("alpha", "beta", "rho") else v) for k, v in sc.items()}
        for nm, v in XI_ETA_IDENTITY:
            assert (sc["metrics"][nm][3:-3, 3:-3, 3:-3] == v).all(), nm
        if conindex:
            sc["conindex"] = conindex
        a, b = _run_pair(dict(sc, simplecar=True), sc, lib, seed=43)
        assert np.abs(a.get_field(0)).max() > 0
        assert _all_equal(a, b), f"SimpleCar vs curvilinear, conindex {conindex}"


PINS = {
    "affine_linear_stage": pin_affine_linear_stage,
    "hybrid_cartesian_branch": pin_hybrid_cartesian_branch,
    "hybrid_equals_curvilinear_on_identity_metrics": pin_hybrid_equals_curvilinear_on_identity_metrics,
    "simplecar_is_the_identity_xi_eta_metrics": pin_simplecar_is_the_identity_xi_eta_metrics,
    "simplecar_equals_curvilinear_on_stretched_grids": pin_simplecar_equals_curvilinear_on_stretched_grids,
    "fb_schedule": pin_fb_schedule,
    "closure_matrices": pin_closure_matrices,
    "source_magnitudes": pin_source_magnitudes,
    "aux_enters_rate_over_b": pin_aux_enters_rate_over_b,
    "free_surface_plane_waves": pin_free_surface_plane_waves,
    "outputs_are_step_end_fields": pin_outputs_are_step_end_fields,
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    PINS[name]()


def test_host_metrics_of_affine_grid():
    """scenario.metrics_from_coords on an affine grid gives the inverse of the
    map's matrix and its determinant at every valid cell."""
    ni, nj, nk = 10, 9, 8
    i, j, k = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    X = np.stack([i, j, k], -1).astype(np.float64) @ M_AFFINE.T * DH
    met = scenario.metrics_from_coords(X[..., 0], X[..., 1], X[..., 2], DH)
    Minv = np.linalg.inv(M_AFFINE)
    names = ("xix", "xiy", "xiz", "etax", "etay", "etaz", "zetax", "zetay", "zetaz")
    s = (slice(3, -3),) * 3
    for n_, nm in enumerate(names):
        np.testing.assert_allclose(met[nm][s], Minv[n_ // 3, n_ % 3], rtol=2e-6, atol=2e-7, err_msg=nm)
    np.testing.assert_allclose(met["jac"][s], np.linalg.det(M_AFFINE), rtol=2e-6)
