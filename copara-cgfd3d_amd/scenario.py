"""Host-side input builders for the time-stepping engine.

These restate the reference's one-time host preprocessing (fes478/CoPara-CGFD3D,
src/gridmesh.cpp, src/absorb.cpp, src/source.cpp, src/mathfunc.cpp,
src/common.cpp) in numpy so that a scenario can be assembled without the
reference's parameter files, and provide generators for the benchmark
configurations (flat homogeneous box, Gaussian hill, random von Karman model).

Arrays use the reference layout: [x][y][z], z fastest, 3 ghost cells per side
(shape (ni+6, nj+6, nk+6)); the last valid z row is the free surface.
Arithmetic is FP32 where the reference's `Real` is float, with the reference's
double-literal promotions, so the tables match what the reference would hand
to its kernels.
"""
import ctypes
import math

import numpy as np

LFD = 3
F32 = np.float32
PI_F = F32(3.141592653589793)  # constant::PI is a Real (typenew.h:115)

# float-argument pow/log bind powf/logf in the reference's float build; the
# C library's own routines keep the tables independent of numpy's SIMD
# implementations (and equal to a C++ host's)
_libm = ctypes.CDLL("libm.so.6")
for _fn in ("powf", "logf"):
    getattr(_libm, _fn).restype = ctypes.c_float
_libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
_libm.logf.argtypes = [ctypes.c_float]


def _powf(a, b):
    return F32(_libm.powf(float(a), float(b)))


def _logf(a):
    return F32(_libm.logf(float(a)))


# ---------------------------------------------------------------------------
# ghost extension (common.cpp:67-166)
# ---------------------------------------------------------------------------
def interpolated_extend(v):
    """Linear extrapolation into the ghosts, z then y then x (common.cpp:67-99)."""
    n = v.shape
    for k in range(1, LFD + 1):
        v[:, :, LFD - k] = (2.0 * v[:, :, LFD] - v[:, :, LFD + k]).astype(v.dtype)
        v[:, :, n[2] - LFD + k - 1] = (2.0 * v[:, :, n[2] - LFD - 1] - v[:, :, n[2] - LFD - k - 1]).astype(v.dtype)
    for j in range(1, LFD + 1):
        v[:, LFD - j, :] = (2.0 * v[:, LFD, :] - v[:, LFD + j, :]).astype(v.dtype)
        v[:, n[1] - LFD + j - 1, :] = (2.0 * v[:, n[1] - LFD - 1, :] - v[:, n[1] - LFD - j - 1, :]).astype(v.dtype)
    for i in range(1, LFD + 1):
        v[LFD - i] = (2.0 * v[LFD] - v[LFD + i]).astype(v.dtype)
        v[n[0] - LFD + i - 1] = (2.0 * v[n[0] - LFD - 1] - v[n[0] - LFD - i - 1]).astype(v.dtype)
    return v


def mirror_extend(v):
    """Mirror the valid region into the face ghosts, no corners (common.cpp:100-135)."""
    n = v.shape
    a, bx, by, bz = LFD, n[0] - LFD, n[1] - LFD, n[2] - LFD
    for i in range(1, LFD + 1):
        v[a - i, a:by, a:bz] = v[a + i, a:by, a:bz]
        v[bx - 1 + i, a:by, a:bz] = v[bx - 1 - i, a:by, a:bz]
    for j in range(1, LFD + 1):
        v[a:bx, a - j, a:bz] = v[a:bx, a + j, a:bz]
        v[a:bx, by - 1 + j, a:bz] = v[a:bx, by - 1 - j, a:bz]
    for k in range(1, LFD + 1):
        v[a:bx, a:by, a - k] = v[a:bx, a:by, a + k]
        v[a:bx, a:by, bz - 1 + k] = v[a:bx, a:by, bz - 1 - k]
    return v


def equivalence_extend(v):
    """Ghosts of the media (common.cpp:136-166), including its y mirror."""
    n = v.shape
    a, bx, by, bz = LFD, n[0] - LFD, n[1] - LFD, n[2] - LFD
    for i in range(1, LFD + 1):
        v[a - i] = v[a]
        v[bx + i - 1] = v[bx - 1]
    for j in range(1, LFD + 1):
        v[:, a - j, :] = v[:, a + j, :]
        v[:, by + j - 1, :] = v[:, by - 1, :]
    for k in range(1, LFD + 1):
        v[:, :, a - k] = v[:, :, a + k]
        v[:, :, bz + k - 1] = v[:, :, bz - 1]
    return v


# ---------------------------------------------------------------------------
# metrics (gridmesh.cpp:1476-1627)
# ---------------------------------------------------------------------------
def metrics_from_coords(x, y, z, dh):
    """Second-order central-difference metrics and Jacobian.

    x, y, z: coordinates with ghosts (ghosts already extended).  Returns the
    10 metric arrays (xix ... zetaz, jac) mirror-extended into the ghosts.
    Float32 arithmetic as in calmetric: differences in float, /2.0/steph in
    double, cross products and the Jacobian dot product in float.
    """
    steph = float(F32(dh))
    shape = np.broadcast_shapes(x.shape, y.shape, z.shape)
    out = {k: np.zeros(shape, F32) for k in ("xix", "xiy", "xiz", "etax", "etay", "etaz",
                                                "zetax", "zetay", "zetaz", "jac")}
    s = slice(LFD, -LFD)

    def dcent(c, axis):
        c = np.broadcast_to(c, shape)
        if axis == 0:
            d = c[LFD + 1:shape[0] - LFD + 1, s, s] - c[LFD - 1:shape[0] - LFD - 1, s, s]
        elif axis == 1:
            d = c[s, LFD + 1:shape[1] - LFD + 1, s] - c[s, LFD - 1:shape[1] - LFD - 1, s]
        else:
            d = c[s, s, LFD + 1:shape[2] - LFD + 1] - c[s, s, LFD - 1:shape[2] - LFD - 1]
        return (d.astype(np.float64) / 2.0 / steph).astype(F32)

    x_xi, y_xi, z_xi = dcent(x, 0), dcent(y, 0), dcent(z, 0)
    x_et, y_et, z_et = dcent(x, 1), dcent(y, 1), dcent(z, 1)
    x_zt, y_zt, z_zt = dcent(x, 2), dcent(y, 2), dcent(z, 2)
    # vec1 x vec2 (crossproduct, mathfunc.cpp:153-159)
    c0 = y_xi * z_et - z_xi * y_et
    c1 = z_xi * x_et - x_xi * z_et
    c2 = x_xi * y_et - y_xi * x_et
    jac = ((F32(0.0) + c0 * x_zt) + c1 * y_zt) + c2 * z_zt
    out["jac"][s, s, s] = jac
    out["zetax"][s, s, s] = c0 / jac
    out["zetay"][s, s, s] = c1 / jac
    out["zetaz"][s, s, s] = c2 / jac
    # vec3 x vec1
    c0 = y_zt * z_xi - z_zt * y_xi
    c1 = z_zt * x_xi - x_zt * z_xi
    c2 = x_zt * y_xi - y_zt * x_xi
    out["etax"][s, s, s] = c0 / jac
    out["etay"][s, s, s] = c1 / jac
    out["etaz"][s, s, s] = c2 / jac
    # vec2 x vec3
    c0 = y_et * z_zt - z_et * y_zt
    c1 = z_et * x_zt - x_et * z_zt
    c2 = x_et * y_zt - y_et * x_zt
    out["xix"][s, s, s] = c0 / jac
    out["xiy"][s, s, s] = c1 / jac
    out["xiz"][s, s, s] = c2 / jac
    for v in out.values():
        mirror_extend(v)
    return out


def stretched_grid(ni, nj, nk, dh, surface, depth):
    """Coordinates of a grid with uniform x, y spacing and a vertically
    stretched z between a flat bottom at -depth and the surface z=surface(x,y).

    surface: callable on (X, Y) arrays in metres, or None for a flat z=0 top.
    Returns x (NX,1,1), y (1,NY,1), z (NX,NY,NZ), ghosts extended.
    """
    NX, NY, NZ = ni + 2 * LFD, nj + 2 * LFD, nk + 2 * LFD
    xs = ((np.arange(NX) - LFD) * dh).astype(F32)
    ys = ((np.arange(NY) - LFD) * dh).astype(F32)
    X, Y = np.meshgrid(xs.astype(np.float64), ys.astype(np.float64), indexing="ij")
    top = np.zeros_like(X) if surface is None else surface(X, Y)
    frac = (np.arange(NZ) - LFD) / (nk - 1)
    z = (-depth + (top[:, :, None] + depth) * frac[None, None, :]).astype(F32)
    interpolated_extend(z)
    return xs[:, None, None], ys[None, :, None], z


# ---------------------------------------------------------------------------
# CFS-PML tables (absorb.cpp:346-800)
# ---------------------------------------------------------------------------
def cfs_pml_tables(ni, nj, nk, nabs, velabs, bmax, amax, dh, metrics):
    """ADE CFS-PML 1-D factors APD = alpha + d/beta, B = beta, DB = d/beta.

    Layer thickness is accumulated inward from the interior edge with the
    local grid spacing steph/|grad(xi)| sampled on the (ni1, nj1, nk1) line;
    d = d0 (L/L0)^2, beta = 1 + (bmax-1)(L/L0)^2, alpha = amax (1 - L/L0),
    d0 = -v/2/L0 ln(R) (PD+1) with R = 10^(-(log10 N - 1)/log10 2 - 3).
    """
    n = {0: ni, 1: nj, 2: nk}
    N = {0: ni + 6, 1: nj + 6, 2: nk + 6}
    names = (("xix", "xiy", "xiz"), ("etax", "etay", "etaz"), ("zetax", "zetay", "zetaz"))
    steph = F32(dh)
    PD, PB, PA = F32(2.0), F32(2.0), F32(1.0)
    A, B, D = {}, {}, {}
    for ax in range(3):
        A[ax] = np.zeros(N[ax], F32)
        B[ax] = np.ones(N[ax], F32)
        D[ax] = np.zeros(N[ax], F32)

    def grad_norm(ax, i):
        idx = [LFD, LFD, LFD]
        idx[ax] = i
        vg = [metrics[nm][tuple(idx)] for nm in names[ax]]
        dp = F32(0.0)
        for c in vg:
            dp = F32(dp + F32(c * c))
        return F32(np.sqrt(dp))

    for side in range(6):
        ax, hi = side // 2, side % 2
        nl = int(nabs[side])
        if nl <= 0:
            continue
        Rpp = F32(10.0 ** (-(math.log10(nl * 1.0) - 1) / math.log10(2.0) - 3.0))
        L = np.zeros(N[ax] + 1, F32)
        if hi == 0:
            i2 = LFD + nl - 1
            for i in range(i2, LFD - 1, -1):
                L[i] = F32(L[i + 1] + F32(steph / grad_norm(ax, i)))
            L0 = L[LFD]
            rng = range(LFD, i2 + 1)
        else:
            i1 = LFD + n[ax] - nl
            for i in range(i1, LFD + n[ax]):
                L[i] = F32(L[i - 1] + F32(steph / grad_norm(ax, i)))
            L0 = L[LFD + n[ax] - 1]
            rng = range(i1, LFD + n[ax])
        dmax = F32(-float(velabs[side]) / 2.0 / float(L0) * float(_logf(Rpp)) * (float(PD) + 1.0))
        for i in rng:
            r = F32(L[i] / L0)
            D[ax][i] = F32(dmax * _powf(r, PD))
            B[ax][i] = F32(1.0 + (float(bmax[side]) - 1.0) * float(_powf(r, PB)))
            A[ax][i] = F32(float(amax[side]) * (1.0 - float(_powf(r, PA))))
    tabs = {}
    for ax, nm in enumerate("xyz"):
        tabs["APD" + nm] = (A[ax] + D[ax] / B[ax]).astype(F32)
        tabs["B" + nm] = B[ax].copy()
        tabs["DB" + nm] = (D[ax] / B[ax]).astype(F32)
    return tabs


def unit_pml(ni, nj, nk):
    tabs = {}
    for nm, n in zip("xyz", (ni, nj, nk)):
        tabs["APD" + nm] = np.zeros(n + 6, F32)
        tabs["B" + nm] = np.ones(n + 6, F32)
        tabs["DB" + nm] = np.zeros(n + 6, F32)
    return tabs


# ---------------------------------------------------------------------------
# source time functions (mathfunc.cpp:481-610) and mechanisms (source.cpp:134-155)
# ---------------------------------------------------------------------------
def ricker(t, fc, t0):
    """mathfunc::Ricker: (u^2/4 - 1/2) exp(-u^2/4) sqrt(pi)/2, u = (t-t0) 2 pi fc; 0 for t<=0."""
    t = np.asarray(t, F32)
    u = ((t - F32(t0)).astype(np.float64) * 2.0 * float(PI_F) * float(F32(fc))).astype(F32)
    uu = (u * u).astype(np.float64)
    y = (uu / 4.0 - 0.5) * np.exp(-uu / 4.0) * float(np.sqrt(PI_F)) / 2.0
    return np.where(t <= 0, F32(0.0), y.astype(F32)).astype(F32)


def gauss_deriv(t, a, t0):
    """mathfunc::gauss: derivative of a Gaussian of width a centred at t0, zero outside (0, 2 t0)."""
    t = np.asarray(t, F32)
    d = (t - F32(t0)).astype(np.float64)
    y = np.exp(-(d * d) / (float(a) ** 2)) / (math.sqrt(float(PI_F)) * float(a)) * (-2.0 * d / float(a) ** 2)
    off = (abs(float(t0)) >= 1e-20) & ((t <= 0) | (t >= 2.0 * float(t0)))
    return np.where(off, F32(0.0), y.astype(F32)).astype(F32)


def stf_table(kind, nt, dt, t0, f0, start=0.0):
    """Samples every dt/2 for i in [0, 2(nt+1)) at t = i*dt/2 - start
    (source.cpp:102-132, nstf = 2*(nt+1) at source.cpp:408); moment tables
    zero values below 1e-20 in magnitude (source.cpp:131)."""
    nstf = 2 * (nt + 1)
    t = (np.arange(nstf) * np.float64(F32(dt / 2.0))).astype(F32) - F32(start)
    if kind == "Ricker":
        s = ricker(t, f0, t0)
    elif kind == "gauss":
        s = gauss_deriv(t, f0, t0)
    else:
        raise ValueError(kind)
    return np.where(np.abs(s) < 1e-20, F32(0.0), s).astype(F32)


def angle2moment(strike, dip, rake):
    """Unit double-couple tensor (mxx, myy, mzz, mxy, mxz, myz), source.cpp:134-155."""
    S, D, R = (math.radians(v) for v in (strike, dip, rake))
    m11 = -(math.sin(D) * math.cos(R) * math.sin(2 * S) + math.sin(2 * D) * math.sin(R) * math.sin(S) ** 2)
    m22 = math.sin(D) * math.cos(R) * math.sin(2 * S) - math.sin(2 * D) * math.sin(R) * math.cos(S) ** 2
    m33 = -(m11 + m22)
    m12 = math.sin(D) * math.cos(R) * math.cos(2 * S) + math.sin(2 * D) * math.sin(R) * math.sin(2 * S) * 0.5
    m13 = -(math.cos(D) * math.cos(R) * math.cos(S) + math.cos(2 * D) * math.sin(R) * math.sin(S))
    m23 = -(math.cos(D) * math.cos(R) * math.sin(S) - math.cos(2 * D) * math.sin(R) * math.cos(S))
    return np.array([m11, m22, m33, m12, m13, m23], dtype=np.float64)


# ---------------------------------------------------------------------------
# random fields
# ---------------------------------------------------------------------------
def von_karman_field(shape, dh, a, nu, rng, dtype=F32):
    """Zero-mean, unit-RMS random field with a von Karman spectrum
    P(k) ~ (1 + k^2 a^2)^-(nu + d/2), generated by spectral filtering of white
    noise (d = number of dimensions)."""
    d = len(shape)
    noise = rng.standard_normal(shape).astype(dtype)
    F = np.fft.rfftn(noise)
    del noise
    ks = [np.fft.fftfreq(n, d=dh) * 2 * np.pi for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1], d=dh) * 2 * np.pi]
    k2 = np.zeros(F.shape, np.float32)
    for ax, k in enumerate(ks):
        sh = [1] * d
        sh[ax] = -1
        k2 = k2 + (k.reshape(sh) ** 2).astype(np.float32)
    F *= ((1.0 + k2 * a * a) ** (-(nu + d / 2.0) / 2.0)).astype(np.float32)
    del k2
    f = np.fft.irfftn(F, s=shape, axes=tuple(range(len(shape)))).astype(dtype)
    del F
    f -= f.mean()
    f /= max(float(np.sqrt(np.mean(f.astype(np.float64) ** 2))), 1e-30)
    return f


# ---------------------------------------------------------------------------
# scenario assembly
# ---------------------------------------------------------------------------
def _media_from_vp(vp, rho):
    vp = vp.astype(F32)
    vs = (vp / F32(math.sqrt(3.0))).astype(F32)
    for a in (vp, vs, rho):
        equivalence_extend(a)
    return vp, vs, rho


def surface_receivers(ni, nj, nk, nx, ny, margin):
    xs = np.linspace(margin, ni - 1 - margin, nx).round().astype(int)
    ys = np.linspace(margin, nj - 1 - margin, ny).round().astype(int)
    return np.array([[x + LFD, y + LFD, nk - 1 + LFD] for x in xs for y in ys], np.int32)


def config_flat(ni=64, nj=64, nk=64, nt=600, dh=100.0, dt=0.005, npml=12, src_depth=20):
    """Plain method, small: identity metrics, homogeneous Vp 6000 / Vs 3464 /
    rho 2700, 12-cell CFS-PML on 5 faces, explosive moment source M0=1e15 at
    (ni/2, nj/2, src_depth cells below the surface), Ricker fc=2 Hz with
    t0=0.6 s, 9 surface receivers."""
    x, y, z = stretched_grid(ni, nj, nk, dh, None, (nk - 1) * dh)
    met = metrics_from_coords(x, y, z, dh)
    shp = (ni + 6, nj + 6, nk + 6)
    vp = np.full(shp, 6000.0, F32)
    vs = np.full(shp, 3464.0, F32)
    rho = np.full(shp, 2700.0, F32)
    nabs = [npml, npml, npml, npml, npml, 0]
    fc = 2.0
    pml = cfs_pml_tables(ni, nj, nk, nabs, [6000.0] * 6, [2.0] * 6, [math.pi * fc] * 6, dh, met)
    m0 = 1e15
    src = np.array([[ni // 2 + LFD, nj // 2 + LFD, nk - 1 - src_depth + LFD]], np.int32)
    mvec = np.array([[m0, m0, m0, 0.0, 0.0, 0.0]], F32)
    stf = stf_table("Ricker", nt, dt, 0.6, fc)[None, :]
    recv = surface_receivers(ni, nj, nk, 3, 3, ni // 4)
    return dict(ni=ni, nj=nj, nk=nk, dh=dh, dt=dt, nt=nt, metrics=met, alpha=vp, beta=vs, rho=rho,
                nabs=nabs, pml=pml, moment_locs=src, moment_vec=mvec, moment_stf=stf, recv_locs=recv,
                pgv=True, name=f"flat-homogeneous-{ni}x{nj}x{nk}")


def config_hill(ni=200, nj=200, nk=150, nt=2000, dh=100.0, dt=0.006, npml=12, nrecv_side=20):
    """Plain method on hill topography: Gaussian hill h=1.5 km, sigma=3 km by
    vertical stretching; 3 flat layers Vp 4000/5500/6500 (Vs=Vp/sqrt3, rho
    2400/2600/2800) with interfaces at z=-3 km and -8 km; double couple
    30/60/90, M0=1e16, Gaussian-derivative STF (width 1/(pi*1.5 Hz)), 4 km
    below the surface at the hill centre; 20x20 surface receivers."""
    cx, cy = (ni - 1) * dh / 2, (nj - 1) * dh / 2
    hill = lambda X, Y: 1500.0 * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / 3000.0 ** 2)
    depth = (nk - 1) * dh
    x, y, z = stretched_grid(ni, nj, nk, dh, hill, depth)
    met = metrics_from_coords(x, y, z, dh)
    zz = z
    vp = np.where(zz > -3000.0, F32(4000.0), np.where(zz > -8000.0, F32(5500.0), F32(6500.0))).astype(F32)
    rho = np.where(zz > -3000.0, F32(2400.0), np.where(zz > -8000.0, F32(2600.0), F32(2800.0))).astype(F32)
    vp, vs, rho = _media_from_vp(vp, rho)
    nabs = [npml, npml, npml, npml, npml, 0]
    pml = cfs_pml_tables(ni, nj, nk, nabs, [6500.0] * 6, [2.0] * 6, [math.pi * 1.5] * 6, dh, met)
    m0 = 1e16
    mvec = (angle2moment(30.0, 60.0, 90.0) * m0).astype(F32)[None, :]
    src = np.array([[ni // 2 + LFD, nj // 2 + LFD, nk - 1 - int(round(4000.0 / dh)) + LFD]], np.int32)
    stf = stf_table("gauss", nt, dt, 1.0, 1.0 / (math.pi * 1.5))[None, :]
    recv = surface_receivers(ni, nj, nk, nrecv_side, nrecv_side, npml + 2)
    return dict(ni=ni, nj=nj, nk=nk, dh=dh, dt=dt, nt=nt, metrics=met, alpha=vp, beta=vs, rho=rho,
                nabs=nabs, pml=pml, moment_locs=src, moment_vec=mvec, moment_stf=stf, recv_locs=recv,
                pgv=True, name=f"gaussian-hill-{ni}x{nj}x{nk}")


def config_scenario_batch(count, nt_range=(800, 2500), seed=0, size_range=((96, 192), (64, 128)), dh=100.0,
                          dt=0.006, npml=12, nsta=32, nt_cap=None):
    """Many independent scenarios on the Gaussian-hill medium (batched FD):
    grid sizes drawn from U[96,192]^2 x U[64,128], a double couple with random
    strike/dip/rake at a random interior position 2-10 km deep, log-uniform
    M0 in [1e14, 1e17], Ricker fc in [1, 2.5] Hz, durations drawn from
    nt_range steps (nt_cap trims them for short measurements), 32 surface
    stations per scenario."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        ni, nj = (int(rng.integers(size_range[0][0], size_range[0][1] + 1)) for _ in range(2))
        nk = int(rng.integers(size_range[1][0], size_range[1][1] + 1))
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        if nt_cap:
            nt = min(nt, nt_cap)
        sc = config_hill(ni=ni, nj=nj, nk=nk, nt=nt, dh=dh, dt=dt, npml=npml, nrecv_side=2)
        depth = rng.uniform(2000.0, min(10000.0, (nk - npml - 4) * dh))
        x = int(rng.integers(npml + 4, ni - npml - 4))
        y = int(rng.integers(npml + 4, nj - npml - 4))
        zk = max(npml + 2, nk - 1 - int(round(depth / dh)))
        m0 = 10 ** rng.uniform(14.0, 17.0)
        fc = rng.uniform(1.0, 2.5)
        sc["moment_locs"] = np.array([[x + LFD, y + LFD, zk + LFD]], np.int32)
        sc["moment_vec"] = (angle2moment(*rng.uniform([0, 10, -180], [360, 90, 180])) * m0).astype(F32)[None, :]
        sc["moment_stf"] = stf_table("Ricker", nt, dt, 1.2 / fc, fc)[None, :]
        sc["recv_locs"] = surface_receivers(ni, nj, nk, 4, max(1, nsta // 4), npml + 2)
        sc["name"] = f"scenario{i}-{ni}x{nj}x{nk}-nt{nt}"
        out.append(sc)
    return out


def config_vonkarman(ni=600, nj=600, nk=300, nt=4000, dh=50.0, dt=0.004, npml=20, nsub=200, seed=2024,
                     nrecv_side=20):
    """Large heterogeneous run: random von Karman surface (a=5 km, RMS 300 m),
    1-D gradient Vp 3500->7000 m/s with a 5 % von Karman perturbation (a=2 km,
    nu=0.3) on Vp and Vs, Poisson solid, Gardner density 310 Vp^0.25,
    20-cell CFS-PML, finite fault of `nsub` double-couple subfaults with
    random rupture times."""
    rng = np.random.default_rng(seed)
    topo = von_karman_field((ni + 6, nj + 6), dh, 5000.0, 0.5, rng, np.float64) * 300.0
    depth = (nk - 1) * dh
    NX, NY, NZ = ni + 6, nj + 6, nk + 6
    frac = (np.arange(NZ) - LFD) / (nk - 1)
    z = np.empty((NX, NY, NZ), F32)
    for i in range(NX):  # row by row keeps the float64 temporaries small
        z[i] = (-depth + (topo[i][:, None] + depth) * frac[None, :]).astype(F32)
    interpolated_extend(z)
    xs = ((np.arange(NX) - LFD) * dh).astype(F32)[:, None, None]
    ys = ((np.arange(NY) - LFD) * dh).astype(F32)[None, :, None]
    met = metrics_from_coords(xs, ys, z, dh)
    pert = np.clip(von_karman_field((NX, NY, NZ), dh, 2000.0, 0.3, rng), -3.0, 3.0)
    vp = np.empty((NX, NY, NZ), F32)
    for i in range(NX):
        dep = np.clip((topo[i][:, None] - z[i].astype(np.float64)) / depth, 0.0, 1.0)
        vp[i] = ((3500.0 + 3500.0 * dep) * (1.0 + 0.05 * pert[i])).astype(F32)
    del pert
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(F32)
    vp, vs, rho = _media_from_vp(vp, rho)
    nabs = [npml, npml, npml, npml, npml, 0]
    pml = cfs_pml_tables(ni, nj, nk, nabs, [7000.0] * 6, [2.0] * 6, [math.pi * 2.0] * 6, dh, met)
    # finite fault: subfaults on a vertical plane through the centre, 200 cells
    cols = int(math.ceil(math.sqrt(nsub * 2)))
    locs, vecs, stfs = [], [], []
    fc, vr = 2.0, 2800.0
    for s in range(nsub):
        ix = ni // 2 - cols + 2 * (s % cols)
        kz = nk - 1 - 40 - 2 * (s // cols)
        locs.append([ix + LFD, nj // 2 + LFD, kz + LFD])
        r = math.hypot((ix - ni // 2) * dh, (kz - (nk - 1 - 60)) * dh)
        t_rup = r / vr + 0.05 * rng.random()
        m0 = 10 ** rng.uniform(14.0, 15.0)
        vecs.append(angle2moment(30.0 + 10 * rng.standard_normal(), 70.0, 90.0 + 10 * rng.standard_normal()) * m0)
        stfs.append(stf_table("Ricker", nt, dt, 1.0 / fc, fc, start=t_rup))
    recv = surface_receivers(ni, nj, nk, nrecv_side, nrecv_side, npml + 2)
    return dict(ni=ni, nj=nj, nk=nk, dh=dh, dt=dt, nt=nt, metrics=met, alpha=vp, beta=vs, rho=rho,
                nabs=nabs, pml=pml, moment_locs=np.array(locs, np.int32), moment_vec=np.array(vecs, F32),
                moment_stf=np.array(stfs, F32), recv_locs=recv, pgv=True,
                name=f"vonkarman-{ni}x{nj}x{nk}")


def focal_series(n, fnt, fdt, rng, m0=1e14):
    """Moment-rate series of n focal points, [6][n][fnt] float32: a random
    double couple times a Brune-type rate m0*t/tau^2*exp(-t/tau) starting at
    a random onset, sampled every fdt from t = 0 (the layout of the
    reference's focal file, seisnc.cpp:1148-1228)."""
    t = np.arange(fnt) * fdt
    ser = np.zeros((6, n, fnt), np.float64)
    for i in range(n):
        mech = angle2moment(*rng.uniform([0, 10, -180], [360, 90, 180]))
        tau, t0 = rng.uniform(3, 8) * fdt, rng.uniform(0, 5) * fdt
        u = np.clip(t - t0, 0, None)
        rate = m0 * u / tau ** 2 * np.exp(-u / tau)
        for c in range(6):
            ser[c, i] = mech[c] * rate
    return ser.astype(F32)


def config_random(ni, nj, nk, nt, nabs, seed=0, dh=100.0, dt=0.005, topo_amp=300.0, nmom=3, nfrc=2,
                  nrecv=6, duplicate=True, nfoc=0, fdt=None, fnt=None, shear=True):
    """Small randomised case for parity tests: curvilinear grid with random
    smooth topography, heterogeneous media, PML on the requested faces,
    moment and force sources (one force on the surface row, and optionally a
    second source on the same cell as the first), receivers on the surface
    and at depth.  nfoc > 0 adds focal points (moment-rate series every fdt,
    default 1.5*dt, fnt samples, default covering two thirds of the run so
    the series also ends inside it); with duplicate, the first focal point
    shares the cell of the first moment source.  shear=False keeps x and y
    uniform (the reference's vmap meshes), so the xi/eta metrics are exact
    constants."""
    rng = np.random.default_rng(seed)
    kx, ky = rng.uniform(0.5, 1.5, 2)
    ph = rng.uniform(0, 2 * np.pi, 2)
    Lx, Ly = ni * dh, nj * dh
    surf = lambda X, Y: topo_amp * np.sin(2 * np.pi * kx * X / Lx + ph[0]) * np.cos(2 * np.pi * ky * Y / Ly + ph[1])
    x, y, z = stretched_grid(ni, nj, nk, dh, surf, (nk - 1) * dh)
    # a gentle horizontal shear makes the xi/eta metrics non-trivial too
    NX, NY, NZ = z.shape
    if shear:
        xx = np.broadcast_to(x, (NX, NY, NZ)).astype(np.float64)
        xx = (xx + 0.05 * (z.astype(np.float64) + (nk - 1) * dh)).astype(F32)
        yy = np.broadcast_to(y, (NX, NY, NZ)).astype(np.float64)
        yy = (yy - 0.03 * xx.astype(np.float64)).astype(F32)
        met = metrics_from_coords(xx, yy, z, dh)
    else:
        met = metrics_from_coords(x, y, z, dh)
    shp = (NX, NY, NZ)
    vp = (5000.0 + 800.0 * rng.standard_normal(shp)).clip(3500, 6500).astype(F32)
    rho = (2500.0 + 200.0 * rng.standard_normal(shp)).clip(2000, 3000).astype(F32)
    vp, vs, rho = _media_from_vp(vp, rho)
    pml = cfs_pml_tables(ni, nj, nk, nabs, [6500.0] * 6, [2.0] * 6, [math.pi * 2.0] * 6, dh, met) \
        if sum(nabs) else unit_pml(ni, nj, nk)
    nstf = 2 * (nt + 1)
    t_ = np.arange(nstf) * dt / 2
    mloc = [[rng.integers(0, ni), rng.integers(0, nj), rng.integers(0, nk)] for _ in range(nmom)]
    floc = [[rng.integers(0, ni), rng.integers(0, nj), rng.integers(0, nk)] for _ in range(nfrc)]
    if nfrc:
        floc[0][2] = nk - 1  # a force on the free surface (doubled amplitude, calculate.cu:6177)
    if duplicate and nmom >= 2:
        mloc[1] = list(mloc[0])
    if duplicate and nfrc >= 2:
        floc[1] = list(floc[0])
    mloc = (np.array(mloc, np.int32) + LFD).reshape(-1, 3)
    floc = (np.array(floc, np.int32) + LFD).reshape(-1, 3)
    mvec = (rng.standard_normal((nmom, 6)) * 1e14).astype(F32)
    fvec = (rng.standard_normal((nfrc, 3)) * 1e10).astype(F32)
    mstf = np.stack([ricker(t_ - rng.uniform(0, 0.05), 3.0, 0.3) for _ in range(nmom)]).astype(F32) \
        if nmom else np.zeros((0, nstf), F32)
    fstf = np.stack([ricker(t_ - rng.uniform(0, 0.05), 3.0, 0.3) for _ in range(nfrc)]).astype(F32) \
        if nfrc else np.zeros((0, nstf), F32)
    recv = [[rng.integers(0, ni), rng.integers(0, nj), nk - 1] for _ in range(nrecv // 2)] + \
           [[rng.integers(0, ni), rng.integers(0, nj), rng.integers(0, nk)] for _ in range(nrecv - nrecv // 2)]
    recv = np.array(recv, np.int32) + LFD
    sc = dict(ni=ni, nj=nj, nk=nk, dh=dh, dt=dt, nt=nt, metrics=met, alpha=vp, beta=vs, rho=rho,
              nabs=list(nabs), pml=pml, moment_locs=mloc, moment_vec=mvec, moment_stf=mstf,
              force_locs=floc, force_vec=fvec, force_stf=fstf, recv_locs=recv, pgv=True,
              name=f"random-{ni}x{nj}x{nk}-s{seed}")
    if nfoc:
        fdt = float(np.float32(1.5 * dt if fdt is None else fdt))
        fnt = int(fnt if fnt is not None else max(2, (2 * nt * dt / 3) // fdt))
        cloc = [[rng.integers(0, ni), rng.integers(0, nj), rng.integers(0, nk)] for _ in range(nfoc)]
        cloc = np.array(cloc, np.int32).reshape(-1, 3) + LFD
        if duplicate and nmom:
            cloc[0] = mloc[0]
        sc.update(focal_locs=cloc, focal_series=focal_series(nfoc, fnt, fdt, rng), focal_dt=fdt)
    return sc


def random_state(shape, rng, scale_v=1e-3, scale_t=1e5):
    """Random smooth-ish wavefield (9 components) with zero ghosts, for
    single-stage checks that exercise every code path at once."""
    out = []
    for c in range(9):
        a = rng.standard_normal(shape).astype(F32) * F32(scale_v if c < 3 else scale_t)
        a[:LFD] = 0
        a[-LFD:] = 0
        a[:, :LFD] = 0
        a[:, -LFD:] = 0
        a[:, :, :LFD] = 0
        a[:, :, -LFD:] = 0
        out.append(a)
    return out
