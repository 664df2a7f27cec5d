"""Parameter-file front end: the inputs of one run from the reference's own
parameter files.

`load(conf)` reads a SeisFD3D.conf and the files it names (vmap grid,
absorb, source, station, media) the way the reference's master does
(fes478/CoPara-CGFD3D src/simulation.cc:759-790, gridmesh.cpp, absorb.cpp,
source.cpp, seisplot.cpp, mediaparas.cpp) and returns the scenario dict that
`cgfd3d.Simulation.setup` takes, plus the receiver and snapshot metadata the
NetCDF writers need.

Arithmetic follows the reference's `Real = float` build:
- values parsed by `get_conf` are `(float)atof`;
- anchor blocks are parsed by `fscanf("%f")`;
- glibc's own `atoi`/`atof`/`strtof` do the parsing;
- every expression keeps C's float/double promotions;
- transcendental calls go to the same glibc routines the reference binds:
  `sinf`/`cosf`/`expf` for float arguments, `sin`/`cos`/`exp` for double ones.

Supported inputs and the reference paths they follow:

| input | reference |
|---|---|
| grid_type vmap, hybrid_grid 0 and 1 (conversion_interface_depth) | gridmesh.cpp:72-532, 1092-1210, 1446-1470 |
| absorb file (ADE CFS-PML) | absorb.cpp:20-46 |
| source_type 2/3: force and moment anchors, 8 STF types, moment or angle mechanisms | source.cpp:12-155, 671-849, 985-1289 |
| source_type 1: focal file, netCDF classic or netCDF-4 | source.cpp:43-58, 787-848, 1185-1262 |
| station file: snapshots, receivers, lines | seisplot.cpp:12-55, 330-457 |
| media_type volume, netCDF classic or netCDF-4 | mediaparas.cpp:376-441, 1273-1425, 2140-2219 |
| media_type interface3D | mediaparas.cpp:34-43, 52-374, 442-914, 1837-1905 |
| sampling_point_per_cell > 1 (media smoothing) | mediaparas.cpp:1712-1835 |

Refused, with the reason in the error:
- grid_type point: `read_point` stores x, y and z all into `crd.x`
  (gridmesh.cpp:52-68), so the reference has no valid grid from it.
- media_type interface: `readdata` overwrites the media from `read_interface`
  with zero Vp/Vs/rho via `get2in1` (mediaparas.cpp:2163-2200), and
  `MediaStatistics` then sets every cell to sqrt(0/0).

Callers may always pass their own `media=(alpha, beta, rho)` arrays instead.
"""
import ctypes
import math
import os

import numpy as np

import scenario as _sc

F32 = np.float32
LFD = 3
PI_F = F32(3.141592653589793)        # constant::PI (typenew.h:111)
SEIS_ZERO = F32(1e-20)               # constant::SeisZero
SEIS_EQUAL = F32(1e-5)               # constant::SeisEqual
SEIS_INF = F32(1e20)                 # constant::SeisInf
STRLEN = 128                         # constant::SeisStrLen: the fgets buffer

_libc = ctypes.CDLL(None)
_libc.atoi.restype = ctypes.c_int
_libc.atoi.argtypes = [ctypes.c_char_p]
_libc.atof.restype = ctypes.c_double
_libc.atof.argtypes = [ctypes.c_char_p]
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]
_libm = ctypes.CDLL("libm.so.6")
for _fn in ("sinf", "cosf", "expf"):
    getattr(_libm, _fn).restype = ctypes.c_float
    getattr(_libm, _fn).argtypes = [ctypes.c_float]


def _sinf(x):
    return F32(_libm.sinf(float(x)))


def _cosf(x):
    return F32(_libm.cosf(float(x)))


def _expf(x):
    return F32(_libm.expf(float(x)))


class ConfError(ValueError):
    """A parameter file is missing, malformed, or asks for something this
    front end refuses (the message says which)."""


# ---------------------------------------------------------------------------
# get_conf (common.cpp:23-66, 286-322)
# ---------------------------------------------------------------------------
class ConfFile:
    """One parameter file, searched the way `common::getConfStr` searches it.

    The search reads the file in fgets pieces of at most 127 bytes, skips
    pieces that start with '#', and takes the first piece whose first token
    is the key and that has at least `col` tokens; the value is the col-th
    token (1-based).  `chunk(anchor)` returns the tokens after the anchor
    line, which `fscanf` reads in the reference."""

    def __init__(self, path, what="configure"):
        self.path = path
        try:
            with open(path, "rb") as f:
                raw = f.read()
        except OSError as e:
            raise ConfError(f"fail to open {what} file {path}: {e.strerror}") from None
        self.pieces = []
        lines = raw.split(b"\n")
        for n, line in enumerate(lines):
            if n + 1 < len(lines):
                line += b"\n"
            for s in range(0, len(line), STRLEN - 1):
                self.pieces.append(line[s:s + STRLEN - 1])

    def _find(self, key, col):
        k = key.encode()
        for n, piece in enumerate(self.pieces):
            if piece[:1] == b"#":
                continue
            toks = piece.split()
            if toks and toks[0] == k and len(toks) >= col:
                return toks[col - 1], n
        raise ConfError(f"there is no value found at {col} colomn for {key} in {self.path}")

    def has(self, key, col=3):
        try:
            self._find(key, col)
            return True
        except ConfError:
            return False

    def str(self, key, col=3):
        return self._find(key, col)[0].decode()

    def int(self, key, col=3):
        return int(_libc.atoi(self._find(key, col)[0]))

    def real(self, key, col=3):
        """(float)atof of the token, as get_conf(..., float *) stores it."""
        return F32(_libc.atof(self._find(key, col)[0]))

    def bool(self, key, col=3):
        return self._find(key, col)[0] != b"0"

    def chunk(self, anchor):
        _, n = self._find(anchor, 1)
        toks = []
        for piece in self.pieces[n + 1:]:
            toks.extend(piece.split())
        return _Scanner(toks, f"{self.path} after {anchor}")


class _Scanner:
    """fscanf("%f") over a token stream: strtof of each token."""

    def __init__(self, toks, where):
        self.toks, self.pos, self.where = toks, 0, where

    def real(self):
        if self.pos >= len(self.toks):
            raise ConfError(f"{self.where}: fewer values than expected")
        tok = self.toks[self.pos]
        end = ctypes.c_char_p()
        buf = ctypes.create_string_buffer(tok)
        v = _libc.strtof(buf, ctypes.byref(end))
        if ctypes.cast(end, ctypes.c_void_p).value == ctypes.addressof(buf):
            raise ConfError(f"{self.where}: '{tok.decode()}' is not a number")
        self.pos += 1
        return F32(v)

    def reals(self, n):
        return np.array([self.real() for _ in range(n)], F32)


# ---------------------------------------------------------------------------
# main parameter file (simulation.cc:759-790 and the getConf of each part)
# ---------------------------------------------------------------------------
def read_main(conf_path, seispath=None):
    """Scalars of SeisFD3D.conf and the paths of the files it names.
    `seispath` overrides the file's own `seispath` (the directory that holds
    the other files)."""
    c = ConfFile(conf_path, "main par")
    m = {
        "ni": c.int("ni"), "nj": c.int("nj"), "nk": c.int("nk"), "nt": c.int("nt"),
        "stept": c.real("stept"), "steph": c.real("steph"),
        "seispath": seispath if seispath is not None else c.str("seispath"),
        "grid_type": c.str("grid_type"), "hybrid_grid": c.int("hybrid_grid"),
        "conversion_depth": c.real("conversion_interface_depth"),
        "media_type": c.str("media_type"),
        "sampling": [c.int("sampling_point_per_cell", 3 + i) for i in range(3)],
        "crit_value": c.real("threshold_velocity_jump_of_interface"),
        "crit_perc": c.real("threshold_percent_change_of_interface"),
        "output_peakvel": c.int("output_peakvel"),
    }
    for key in ("grid_filename", "media_filename", "source_filename", "station_filename", "absorb_filename"):
        m[key] = f"{m['seispath']}/{c.str(key)}"  # sprintf("%s/%s", parpath, name)
    if m["ni"] <= 0 or m["nj"] <= 0 or m["nk"] <= 0 or m["nt"] <= 0:
        raise ConfError(f"{conf_path}: ni, nj, nk and nt must be positive")

    return m


# ---------------------------------------------------------------------------
# vmap grid (gridmesh.cpp:72-532 with hybrid_grid = 0; 1092-1210)
# ---------------------------------------------------------------------------
def _set_twoside_space(l, na, nb, ha1, ha2, hb1, hb2):
    """gridmesh::set_twoside_space (gridmesh.cpp:1184-1210), float arithmetic."""
    l0 = F32(F32(l - F32(F32(na) * ha1)) - F32(F32(nb) * hb1))
    la = F32(F32(na) * F32(ha2 - ha1))
    lb = F32(F32(nb) * F32(hb2 - hb1))
    if abs(l0) < SEIS_EQUAL:
        return ha1, hb1
    if abs(la) < SEIS_EQUAL:
        return ha1, F32(hb1 + F32(l0 / F32(nb)))
    if abs(lb) < SEIS_EQUAL:
        return F32(ha1 + F32(l0 / F32(na))), hb1
    return (F32(ha1 + F32(F32(F32(l0 * la) / F32(la + lb)) / F32(na))),
            F32(hb1 + F32(F32(F32(l0 * lb) / F32(la + lb)) / F32(nb))))


def smooth_zdepth(za, zb, ha, hb, n):
    """gridmesh::smooth_zdepth (gridmesh.cpp:1092-1182): n+1 depths from zb
    (bottom) to about za whose spacing changes from hb at the bottom to ha
    at the top; float arithmetic with the source's double promotions."""
    za, zb, ha, hb = F32(za), F32(zb), F32(ha), F32(hb)
    l = F32(za - zb)
    h0 = F32(l / F32(n))
    h = [F32(0.0)] * (n + 2)
    h[0], h[n + 1] = hb, ha
    if (h0 >= ha and h0 <= hb) or (h0 <= ha and h0 >= hb):
        nb, na, n0 = 1, n, n
        for _ in range(4 * n + 8):
            l0 = F32(0.0)
            f = F32(math.pow(float(F32(ha / hb)), 1.0 / (n0 + 1)) - 1.0)
            for k in range(1, nb):
                h[k] = hb
            for k in range(na + 1, n + 2):
                h[k] = ha
            for k in range(nb, na + 1):
                h[k] = F32(h[k - 1] * F32(F32(1.0) + f))
                l0 = F32(l0 + h[k])
            ha1 = F32(float(ha) * min(1.0, 1.0 - float(f)))
            ha2 = F32(float(ha) * max(1.0, 1.0 - float(f)))
            hb1 = F32(float(hb) * min(1.0, 1.0 + float(f)))
            hb2 = F32(float(hb) * max(1.0, 1.0 + float(f)))
            la1, la2 = F32(F32(n - na) * ha1), F32(F32(n - na) * ha2)
            lb1, lb2 = F32(F32(nb - 1) * hb1), F32(F32(nb - 1) * hb2)
            rest = F32(l - l0)
            if F32(la1 + lb1) <= rest <= F32(la2 + lb2):
                ha0, hb0 = _set_twoside_space(rest, n - na, nb - 1, ha1, ha2, hb1, hb2)
                for k in range(na + 1, n + 1):
                    h[k] = ha0
                for k in range(1, nb):
                    h[k] = hb0
                break
            if rest < F32(la1 + lb1):
                if ha < hb:
                    na, n0 = na - 1, n0 - 1
                else:
                    nb, n0 = nb + 1, n0 - 1
            else:
                if ha < hb:
                    nb, n0 = nb + 1, n0 - 1
                else:
                    na, n0 = na - 1, n0 - 1
            if n0 < 0 or na < nb - 1:
                break
        else:
            raise ConfError("smooth_zdepth: no transition spacing fits this layer")
        if n0 < 0 or na < nb - 1:
            raise ConfError("smooth_zdepth: no transition spacing fits this layer")
    else:
        na = nb = n // 4
        n0 = n - na - nb
        la1 = F32((na + 1) * na / 2.0)
        lb1 = F32((nb + 1) * nb / 2.0)
        la2 = F32(F32(la1 + F32(F32(F32(na) * lb1) / F32(nb))) + F32(na * n0))
        t = F32(l - F32(F32(na) * ha))
        t = F32(t - F32(F32(nb) * hb))
        t = F32(t - F32(F32(n0) * ha))
        l0 = F32(t - F32(F32(la1 * F32(ha - hb)) / F32(nb)))
        ha0 = F32(l0 / la2)
        hb0 = F32(F32(F32(F32(na) * ha0) + ha - hb) / F32(nb))
        for k in range(1, nb + 1):
            h[k] = F32(hb + F32(F32(k) * hb0))
        for k in range(nb + 1, nb + n0 + 1):
            h[k] = h[k - 1]
        for k in range(nb + n0 + 1, n + 1):
            h[k] = F32(ha + F32(F32(n - k + 1) * ha0))
    z = [F32(0.0)] * (n + 1)
    for k in range(n):
        z[k + 1] = F32(z[k] + h[k + 1])
    return np.array([F32(v + zb) for v in z], F32)


def read_vmap(path, ni, nj, nk):
    """Anchors of a vmap grid file: (posx[ni][nj], posy[ni][nj],
    posz[ni][nj][nsurf], gridpoints, equalspacing), with the checks of
    read_vmapnew (gridmesh.cpp:72-156)."""
    c = ConfFile(path, "gridvmap data")
    dims = [c.int("vmap_dims", 3 + i) for i in range(3)]
    ns = dims[2]
    if ns < 2:
        raise ConfError(f"{path}: vmap_dims needs at least 2 surfaces")
    gp = [c.int("vmap_gridpoints", 3 + i) for i in range(ns - 1)]
    eq = [c.bool("vmap_equalspacing", 3 + i) for i in range(ns - 1)]
    s = c.chunk("<vmap_anchor>")
    vals = s.reals(dims[0] * dims[1] * (ns + 2)).reshape(dims[0], dims[1], ns + 2)
    posx, posy, posz = vals[:, :, 0].copy(), vals[:, :, 1].copy(), vals[:, :, 2:].copy()
    if dims[0] != ni:
        raise ConfError("The X value number(dims[0]) obtained from gridfile is different with configure number(cdx.ni)")
    if dims[1] != nj:
        raise ConfError("The Y value number(dims[0]) obtained from gridfile is different with configure number(cdx.nj)")
    if not eq[0] or not eq[ns - 2]:
        raise ConfError("The first layer and last layer should be equal spacing")
    for i in range(1, ns - 1):
        if not eq[i - 1] and not eq[i]:
            raise ConfError("There is only one layer(in two adjacent layers) need to adjust the grid spacing")
    if sum(gp) + 1 != nk:
        raise ConfError("The Z value number[ sum(gridpoints)+1 ] obtained from gridfile is different with "
                        "configure number(cdx.nk)")
    bad = np.argwhere(posz[:, :, 1:] > posz[:, :, :-1])
    if len(bad):
        i, j, k = bad[0]
        raise ConfError(f"The Z direction depth should be monotonically decreasing, while position[{i}][{j}][{k + 1}]"
                        f" should deeper than [{i}][{j}][{k}]")
    return posx, posy, posz, gp, eq


def hybrid_layers(posz, gp, eq, nk, conv_depth):
    """The hybrid grid's z layering (gridmesh.cpp:243-375): the layer that
    holds the conversion depth ends there, and one more layer runs flat from
    it down to the shallowest point of the last interface; every layer's
    share of the nk grid planes is recomputed from its thickness.  Returns
    (Zdep [ni][nj][L+1], GPs [L], SpaceEQ [L], indexZ [L+1], ConIndex) for
    L = ConInterface + 1 layers; the planes below ConIndex lie on flat
    equally spaced levels."""
    ni, nj, ns = posz.shape
    gint = [(F32(posz[:, :, k].min()), F32(posz[:, :, k].max())) for k in range(ns)]
    glay = [(F32(posz[:, :, k + 1].min()), F32(posz[:, :, k].max())) for k in range(ns - 1)]
    cd = F32(conv_depth)
    con = None
    for k in range(ns - 1):
        if glay[k][0] <= cd <= glay[k][1]:
            con = k + 1
            break
    if con is None:
        raise ConfError(f"conversion_interface_depth {float(cd)} m lies in no layer of the vmap grid")
    top, bottom = cd, gint[ns - 1][0]
    space = [bool(eq[k]) for k in range(con)] + [True]
    zdep = np.empty((ni, nj, con + 2), F32)
    zdep[:, :, :con] = posz[:, :, :con]
    zdep[:, :, con] = top
    zdep[:, :, con + 1] = bottom
    layer = [list(glay[k]) for k in range(con)] + [[bottom, top]]
    layer[con - 1][0] = top
    total = F32(0.0)
    for k in range(con + 1):
        total = F32(total + F32(layer[k][1] - layer[k][0]))
    gps = [0] * (con + 1)
    gps[con] = nk - 1
    for k in range(con):
        part = F32(layer[k][1] - layer[k][0])
        gps[k] = int(F32(F32(F32(nk) * part) / total))    # int = int*Real/Real
        gps[con] -= gps[k]
    index = [0] * (con + 2)
    index[con + 1] = LFD
    for k in range(con, -1, -1):
        index[k] = index[k + 1] + gps[k]
    return zdep, gps, space, index, index[con + 1] + gps[con]


def _fill_z(zdep, gps, space, index, nk):
    """Column depths zvec (ni, nj, nk+6) between the interfaces zdep at the
    plane indices index (gridmesh.cpp:395-414, 469-496): equally spaced
    layers by the float expression of the source, transition layers by
    smooth_zdepth."""
    ni, nj, _ = zdep.shape
    zvec = np.zeros((ni, nj, nk + 2 * LFD), F32)
    for k in range(len(gps)):
        hi, lo = index[k], index[k + 1]
        if space[k]:
            step = ((zdep[:, :, k] - zdep[:, :, k + 1]).astype(np.float64) / (1.0 * gps[k])).astype(F32)
            for m in range(lo, hi + 1):
                zvec[:, :, m] = zdep[:, :, k] + F32(m - hi) * step
        else:
            for i in range(ni):
                for j in range(nj):
                    p = zdep[i, j]
                    zs0 = F32(float(F32(p[k - 1] - p[k])) / (1.0 * gps[k - 1]))
                    zs1 = F32(float(F32(p[k + 1] - p[k + 2])) / (1.0 * gps[k + 1]))
                    zvec[i, j, lo:lo + gps[k] + 1] = smooth_zdepth(p[k], p[k + 1], zs0, zs1, gps[k])
        zvec[:, :, hi] = zdep[:, :, k]
        zvec[:, :, lo] = zdep[:, :, k + 1]
    return zvec


def vmap_coords(posx, posy, posz, gp, eq, nk, conv_depth=None):
    """Coordinates with ghosts (x, y, z each (ni+6, nj+6, nk+6)) from vmap
    anchors, and ConIndex: the all-curvilinear branch of read_vmapnew
    (gridmesh.cpp:439-513, ConIndex = nk1) or, with a conversion depth, its
    hybrid branch (243-431); then the ghost extrapolation of readdata
    (gridmesh.cpp:1467-1469)."""
    ni, nj, ns = posz.shape
    shape = (ni + 2 * LFD, nj + 2 * LFD, nk + 2 * LFD)
    x = np.zeros(shape, F32)
    y = np.zeros(shape, F32)
    z = np.zeros(shape, F32)
    v = (slice(LFD, LFD + ni), slice(LFD, LFD + nj))
    x[v] = posx[:, :, None]
    y[v] = posy[:, :, None]
    if conv_depth is None:
        zindex = [0] * ns
        zindex[ns - 1] = LFD
        for i in range(ns - 2, -1, -1):
            zindex[i] = zindex[i + 1] + gp[i]
        zvec = _fill_z(posz, list(gp), [bool(e) for e in eq], zindex, nk)
        conindex = LFD
    else:
        zdep, gps, space, index, conindex = hybrid_layers(posz, gp, eq, nk, conv_depth)
        zvec = _fill_z(zdep, gps, space, index, nk)
    z[v + (slice(LFD, LFD + nk),)] = zvec[:, :, LFD:LFD + nk]
    for a in (x, y, z):
        _sc.interpolated_extend(a)
    return x, y, z, conindex


# ---------------------------------------------------------------------------
# absorbing layers (absorb.cpp:20-46)
# ---------------------------------------------------------------------------
def read_absorb(path):
    c = ConfFile(path, "absorb configure")
    nabs = [c.int("abs_number", 3 + i) for i in range(6)]
    vel = [c.real("abs_velocity", 3 + i) for i in range(6)]
    bmax = [c.real("bmax", 3 + i) for i in range(6)]
    amax = [c.real("amax", 3 + i) for i in range(6)]
    return nabs, vel, bmax, amax


# ---------------------------------------------------------------------------
# source time functions (mathfunc.cpp:481-610), float build
# ---------------------------------------------------------------------------
_SQRT_PI_F = F32(np.sqrt(PI_F))      # sqrt(PI) binds sqrtf: PI is a float


def _stf_value(name, x, t0, f0):
    """One sample of a source time function as cal_stf_force/cal_stf_moment
    call it (source.cpp:75-96): x = i*dt - start; Bell-type functions get
    x - t0 and width f0."""
    x, t0, f0 = F32(x), F32(t0), F32(f0)
    if name in ("Gauss", "gauss"):
        if abs(t0) >= SEIS_ZERO and (x <= 0.0 or float(x) >= 2.0 * float(t0)):
            return F32(0.0)
        d = F32(x - t0)
        g = F32(_expf(F32(-F32(F32(d * d) / F32(f0 * f0)))) / F32(_SQRT_PI_F * f0))
        if name == "Gauss":
            return g
        return F32(float(g) * (-2.0 * float(d) / float(F32(f0 * f0))))
    if name in ("Ricker", "ricker"):
        if x <= 0.0:
            return F32(0.0)
        u = F32(float(F32(x - t0)) * 2.0 * float(PI_F) * float(f0))
        uu = float(F32(u * u))
        e = math.exp(-uu / 4.0)
        if name == "Ricker":
            return F32((uu / 4.0 - 0.5) * e * float(_SQRT_PI_F) / 2.0)
        return F32(float(u) * (1.5 - uu / 4.0) * e * float(_SQRT_PI_F) / 2.0 * float(PI_F) * float(f0))
    if name in ("Bell", "bell", "BELL"):
        xb = F32(x - t0)
        w = 2.0 * float(PI_F) * float(xb) / float(f0)
        if name == "Bell":
            return F32((1.0 - math.cos(w)) / float(f0)) if 0.0 < xb < f0 else F32(0.0)
        if name == "bell":
            return F32(2.0 * float(PI_F) * math.sin(w) / float(f0)) if 0.0 < xb < f0 else F32(0.0)
        if xb <= 0.0:
            return F32(0.0)
        if xb < f0:
            return F32(float(F32(xb / f0)) - math.sin(w) / (2.0 * float(PI_F)))
        return F32(1.0)
    if name == "Step":
        return F32(1.0) if x > 0.0 else F32(0.0)
    raise ConfError(f"Wrong type({name}) for source time function")


def stf_samples(name, start, t0, f0, nt, stept, moment):
    """The 2(nt+1) half-step samples of one source (source.cpp:72-132):
    t = i*(stept/2) - start; moment tables flush |s| < 1e-20 to 0."""
    nstf = 2 * (nt + 1)
    dt = F32(float(F32(stept)) / 2.0)
    out = np.empty(nstf, F32)
    for i in range(nstf):
        s = _stf_value(name, F32(F32(i) * dt) - F32(start), t0, f0)
        out[i] = F32(0.0) if (moment and abs(s) < SEIS_ZERO) else s
    return out


def angle2moment_f(strike, dip, rake):
    """source::angle2moment (source.cpp:134-155) in the float build: the
    angles are floats, so sin/cos of a float bind sinf/cosf and the 2x-angle
    terms are double."""
    S = F32(float(F32(strike)) / 180.0 * float(PI_F))
    D = F32(float(F32(dip)) / 180.0 * float(PI_F))
    R = F32(float(F32(rake)) / 180.0 * float(PI_F))
    sD, cD, sR, cR, sS, cS = _sinf(D), _cosf(D), _sinf(R), _cosf(R), _sinf(S), _cosf(S)
    s2S, c2S = math.sin(2.0 * float(S)), math.cos(2.0 * float(S))
    s2D, c2D = math.sin(2.0 * float(D)), math.cos(2.0 * float(D))
    m11 = F32(-(float(F32(sD * cR)) * s2S + s2D * float(sR) * float(sS) * float(sS)))
    m22 = F32(float(F32(sD * cR)) * s2S - s2D * float(sR) * float(cS) * float(cS))
    m33 = F32(-F32(m11 + m22))
    m12 = F32(float(F32(sD * cR)) * c2S + s2D * float(sR) * s2S * 0.5)
    m13 = F32(-(float(F32(F32(cD * cR) * cS)) + c2D * float(sR) * float(sS)))
    m23 = F32(-(float(F32(F32(cD * cR) * sS)) - c2D * float(sR) * float(cS)))
    return np.array([m11, m22, m33, m12, m13, m23], F32)


# ---------------------------------------------------------------------------
# point location (common.cpp:225-268)
# ---------------------------------------------------------------------------
def reposition(px, py, pz, x, y, z, hyp):
    """common::reposition: the grid point nearest to (px, py, pz) in float
    squared distance, first in (i, j, k) order among equals.  Points above
    `hyp` are placed on the surface row by their horizontal distance, and
    get that point's z.  Returns ghost-inclusive (ix, iy, iz) and the new z."""
    px, py, pz = F32(px), F32(py), F32(pz)
    ni, nj, nk = (s - 2 * LFD for s in x.shape)
    v = (slice(LFD, LFD + ni), slice(LFD, LFD + nj))
    if pz > hyp:
        kz = LFD + nk - 1
        dx = x[v + (kz,)] - px
        dy = y[v + (kz,)] - py
        p = dx * dx + dy * dy
        n = int(np.argmin(p))
        if not p.flat[n] < SEIS_INF:
            return 0, 0, kz, z[0, 0, kz]
        ix, iy = divmod(n, nj)
        return ix + LFD, iy + LFD, kz, z[ix + LFD, iy + LFD, kz]
    best, loc = SEIS_INF, (0, 0, 0)
    vz = slice(LFD, LFD + nk)
    for i in range(ni):  # one x plane at a time keeps the temporaries small
        dx = x[LFD + i, v[1], vz] - px
        dy = y[LFD + i, v[1], vz] - py
        dz = z[LFD + i, v[1], vz] - pz
        p = (dx * dx + dy * dy) + dz * dz
        n = int(np.argmin(p))
        if p.flat[n] < best:
            best = p.flat[n]
            j, k = divmod(n, nk)
            loc = (i + LFD, j + LFD, k + LFD)
    return loc[0], loc[1], loc[2], pz


def _inside(loc, ni, nj, nk):
    ix, iy, iz = loc
    return LFD <= ix < LFD + ni and LFD <= iy < LFD + nj and LFD <= iz < LFD + nk


# ---------------------------------------------------------------------------
# sources (source.cpp:12-70, 671-849, 985-1289)
# ---------------------------------------------------------------------------
_NC4 = None


def _nc4_library():
    """libcgfd_nc4.so (built where libhdf5 is installed), or None."""
    global _NC4
    if _NC4 is None:
        path = os.environ.get("CGFD_NC4_LIB", os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                          "libcgfd_nc4.so"))
        try:
            lib = ctypes.CDLL(path)
        except OSError:
            _NC4 = False
            return None
        C, LL, VP, IP = ctypes.c_char_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)
        lib.cgfd_nc4_probe.argtypes = [C]
        lib.cgfd_nc4_var_dims.argtypes = [C, C, IP, ctypes.POINTER(LL)]
        lib.cgfd_nc4_get_var_float.argtypes = [C, C, VP, LL]
        lib.cgfd_nc4_get_att_float.argtypes = [C, C, ctypes.POINTER(ctypes.c_float)]
        for fn in ("cgfd_nc4_probe", "cgfd_nc4_var_dims", "cgfd_nc4_get_var_float", "cgfd_nc4_get_att_float"):
            getattr(lib, fn).restype = ctypes.c_int
        _NC4 = lib
    return _NC4 or None


class _NC4File:
    """The parts of a netCDF file the front end reads, from a netCDF-4
    (HDF5) file: dimension lengths, float variables and float global
    attributes, as nc_inq_dimlen, nc_get_var_float and nc_get_att_float
    return them (seisnc.cpp:1148-1228)."""

    def __init__(self, lib, path, what):
        self._lib, self._path, self._what = lib, os.fsencode(path), what
        self.dimensions = _NC4Dims(self)
        self.variables = _NC4Vars(self)

    def _shape(self, name):
        r = ctypes.c_int(0)
        d = (ctypes.c_longlong * 8)()
        if self._lib.cgfd_nc4_var_dims(self._path, name.encode(), ctypes.byref(r), d) != 0:
            raise ConfError(f"{self._what} {self._path.decode()} has no variable or dimension {name}")
        return tuple(int(d[i]) for i in range(r.value))

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        v = ctypes.c_float(0)
        if self._lib.cgfd_nc4_get_att_float(self._path, name.encode(), ctypes.byref(v)) != 0:
            raise AttributeError(f"{self._what} has no global attribute {name}")
        return F32(v.value)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


class _NC4Dims:
    def __init__(self, f):
        self._f = f

    def __getitem__(self, name):
        shape = self._f._shape(name)
        return shape[0] if shape else 1


class _NC4Var:
    def __init__(self, f, name):
        self._f, self._name = f, name

    def __getitem__(self, key):
        shape = self._f._shape(self._name)
        out = np.empty(shape, F32)
        if self._f._lib.cgfd_nc4_get_var_float(self._f._path, self._name.encode(),
                                               out.ctypes.data_as(ctypes.c_void_p), out.size) != 0:
            raise ConfError(f"fail to read variable {self._name} of {self._f._what}")
        return out[key]


class _NC4Vars:
    def __init__(self, f):
        self._f = f

    def __getitem__(self, name):
        self._f._shape(name)
        return _NC4Var(self._f, name)


def _read_netcdf(path, what):
    """A netCDF input file: the classic format through scipy's reader, the
    netCDF-4 format (HDF5) through libcgfd_nc4.so and libhdf5."""
    from scipy.io import netcdf_file

    try:
        with open(path, "rb") as f:
            magic = f.read(8)
    except OSError as e:
        raise ConfError(f"fail to open {what} {path}: {e.strerror}") from None
    if magic[:3] == b"CDF":
        return netcdf_file(path, "r", mmap=False)
    if magic == b"\x89HDF\r\n\x1a\n":
        lib = _nc4_library()
        if lib is None or lib.cgfd_nc4_probe(os.fsencode(path)) != 1:
            if lib is None:
                raise ConfError(f"{what} {path} is a netCDF-4 (HDF5) file, and the netCDF-4 reader "
                                f"(libcgfd_nc4.so, built against libhdf5) is not available; convert it to the "
                                f"classic format")
            raise ConfError(f"{what} {path} has the HDF5 signature but does not open as an HDF5 file")
        return _NC4File(lib, path, what)
    raise ConfError(f"{what} {path} is neither a netCDF classic nor a netCDF-4 file")


def read_sources(path, seispath, nt, stept):
    """Sources of the source file with positions in metres (before
    location).  Returns a dict with force / moment / focal lists."""
    c = ConfFile(path, "source configure")
    stype = c.int("source_type")
    out = {"force": None, "moment": None, "focal": None}
    dst2m = c.real("distance2meter")
    out["hyp"] = F32(c.real("source_hyper_height") * dst2m)
    fd0 = c.real("simulate_damp")  # read for every source type (source.cpp:786-788)
    if stype == 1:
        fn = f"{seispath}/{c.str('focal_process_NC_file')}"
        with _read_netcdf(fn, "focal process file") as f:
            npnt, fnt = f.dimensions["NP"], f.dimensions["NT"]
            fdt = F32(f.DT if np.ndim(f.DT) == 0 else f.DT[0])
            ser = np.stack([np.asarray(f.variables[nm][:], F32).reshape(npnt, fnt) * fd0
                            for nm in ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz")])
            pos = np.stack([np.asarray(f.variables[nm][:], F32).reshape(npnt) for nm in ("posx", "posy", "posz")], 1)
        out["focal"] = {"pos": pos, "series": ser.astype(F32), "dt": fdt, "nt": fnt}
        return out
    if stype not in (2, 3):
        return out  # any other type leaves the run without sources (source.cpp:39-66)
    nfrc, nmnt = c.int("number_of_force_source"), c.int("number_of_moment_source")
    if nfrc:
        nwin = c.int("force_stf_window")
        s = c.chunk("<anchor_force>")
        rows = [s.reals(8) for _ in range(nfrc)]
        pos = np.array([[F32(r[0] * dst2m), F32(r[1] * dst2m), F32(r[2] * dst2m)] for r in rows], F32)
        vec = np.array([[F32(r[5] * r[4]), F32(r[6] * r[4]), F32(r[7] * r[4])] for r in rows], F32)
        stf = np.zeros((nfrc, 2 * (nt + 1)), F32)
        for i, r in enumerate(rows):
            for k in range(nwin):  # the last window is what the reference keeps
                stf[i] = stf_samples(c.str("force_stf_type", 3 + k), r[3], c.real("force_stf_timefactor", 3 + k),
                                     c.real("force_stf_freqfactor", 3 + k), nt, stept, moment=False)
        out["force"] = {"pos": pos, "vec": vec, "stf": stf}
    if nmnt:
        nwin = c.int("moment_stf_window")
        mech = c.str("moment_mech_input")
        if mech not in ("moment", "angle"):
            raise ConfError(f"moment mechanics input wrong = {mech}")
        s = c.chunk("<anchor_moment>")
        pos, vec, start = [], [], []
        for _ in range(nmnt):
            r = s.reals(11 if mech == "moment" else 8)
            m = r[5:11] if mech == "moment" else angle2moment_f(r[5], r[6], r[7])
            pos.append([F32(r[0] * dst2m), F32(r[1] * dst2m), F32(r[2] * dst2m)])
            vec.append([F32(v * r[4]) for v in m])
            start.append(r[3])
        stf = np.zeros((nmnt, 2 * (nt + 1)), F32)
        for i in range(nmnt):
            for k in range(nwin):
                stf[i] = stf_samples(c.str("moment_stf_type", 3 + k), start[i], c.real("moment_stf_timefactor", 3 + k),
                                     c.real("moment_stf_freqfactor", 3 + k), nt, stept, moment=True)
        out["moment"] = {"pos": np.array(pos, F32), "vec": np.array(vec, F32), "stf": stf, "angle": mech == "angle"}
    return out


def locate_sources(src, x, y, z, rho, beta):
    """source::cal_index (source.cpp:985-1289) without SrcSmooth: each source
    goes to its nearest grid point; angle-mode moments and focal series are
    scaled by mu = rho*beta*beta there.  Force and moment sources outside
    the valid grid are errors; focal points outside it are dropped (the
    reference zeroes their series)."""
    ni, nj, nk = (s - 2 * LFD for s in x.shape)
    hyp = src["hyp"]
    res = {}
    for kind in ("force", "moment"):
        s = src[kind]
        if s is None:
            continue
        locs = []
        for i, p in enumerate(s["pos"]):
            ix, iy, iz, pz = reposition(p[0], p[1], p[2], x, y, z, hyp)
            if not _inside((ix, iy, iz), ni, nj, nk):
                raise ConfError(f"The {i + 1} {kind} source at position({p[0]},{p[1]},{p[2]}), exceed the coord boundary")
            s["pos"][i, 2] = pz
            locs.append((ix, iy, iz))
        locs = np.array(locs, np.int32)
        vec = s["vec"].copy()
        if kind == "moment" and s["angle"]:
            for i, (ix, iy, iz) in enumerate(locs):
                vec[i] *= F32(F32(rho[ix, iy, iz] * beta[ix, iy, iz]) * beta[ix, iy, iz])
        res[kind + "_locs"], res[kind + "_vec"], res[kind + "_stf"] = locs, vec, s["stf"]
    f = src["focal"]
    if f is not None:
        keep, locs = [], []
        for i, p in enumerate(f["pos"]):
            loc = reposition(p[0], p[1], p[2], x, y, z, hyp)[:3]
            if _inside(loc, ni, nj, nk):
                keep.append(i)
                locs.append(loc)
        ser = f["series"][:, keep, :].copy()
        for n, (ix, iy, iz) in enumerate(locs):
            ser[:, n, :] = F32(F32(rho[ix, iy, iz] * beta[ix, iy, iz]) * beta[ix, iy, iz]) * ser[:, n, :]
        res["focal_locs"] = np.array(locs, np.int32).reshape(-1, 3)
        res["focal_series"] = ser
        res["focal_dt"] = float(f["dt"])
    return res


# ---------------------------------------------------------------------------
# stations (seisplot.cpp:12-55, 330-457)
# ---------------------------------------------------------------------------
def read_stations(path):
    """Snapshot specs, receiver/line points (metres) with their Pnum/Lnum,
    topo_hyper_height and tinv_of_seismo of a station file."""
    c = ConfFile(path, "seisplot configure")
    nsnap, nline, nrecv = c.int("number_of_snap"), c.int("number_of_inline"), c.int("number_of_recv")
    hyp = c.real("topo_hyper_height")
    tinv = c.int("tinv_of_seismo")
    snaps = np.array([[c.int(f"snap_{i + 1:03d}", col) for col in range(3, 14)] for i in range(nsnap)],
                     np.int32).reshape(-1, 11)
    pos, pnum, lnum = [], [], []
    for i in range(nrecv):
        key = f"recv_{i + 1:03d}"
        pos.append([c.real(key, 3), c.real(key, 4), c.real(key, 5)])
        pnum.append(i)
        lnum.append(0)
    for i in range(nline):
        key = f"line_{i + 1:03d}"
        x0, y0, z0 = c.real(key, 3), c.real(key, 4), c.real(key, 5)
        dx, dy, dz = c.real(key, 7), c.real(key, 8), c.real(key, 9)
        num = c.real(key, 11)
        j = 0
        while j < num:
            pos.append([F32(x0 + F32(F32(j) * dx)), F32(y0 + F32(F32(j) * dy)), F32(z0 + F32(F32(j) * dz))])
            pnum.append(j)
            lnum.append(i + 1)
            j += 1
    return {"snapshots": snaps, "pos": np.array(pos, F32).reshape(-1, 3), "pnum": np.array(pnum, np.int32),
            "lnum": np.array(lnum, np.int32), "hyp": hyp, "tinv": tinv}


def locate_stations(st, x, y, z):
    """seisplot::locpoint (seisplot.cpp:412-440): every receiver and line
    point goes to its nearest valid grid point; surface points take that
    point's z."""
    ni, nj, nk = (s - 2 * LFD for s in x.shape)
    locs = []
    for i, p in enumerate(st["pos"]):
        ix, iy, iz, pz = reposition(p[0], p[1], p[2], x, y, z, st["hyp"])
        if not _inside((ix, iy, iz), ni, nj, nk):
            raise ConfError(f"No.{i} reciver is out of Computing Boundary (No Virtual)!")
        st["pos"][i, 2] = pz
        locs.append((ix, iy, iz))
    return np.array(locs, np.int32).reshape(-1, 3)


# ---------------------------------------------------------------------------
# volume media (mediaparas.cpp:294-441 of mathfunc/mediapar, 1273-1425, 2140-2219)
# ---------------------------------------------------------------------------
def loc_value_1d(p, arr):
    """mathfunc::LocValue1d (mathfunc.cpp:294-328), vectorised: left index L
    of the bracketing pair and the value clamped to the array's ends."""
    p = np.asarray(p, F32)
    n = arr.size
    L = np.searchsorted(arr, p, side="right") - 1
    L = np.clip(L, 0, n - 2)
    val = p.copy()
    lo, hi = p <= arr[0], p >= arr[n - 1]
    L[lo], val[lo] = 0, arr[0]
    L[hi & ~lo], val[hi & ~lo] = n - 2, arr[n - 1]
    return L, val


def _lagrange2(X0, X1, x0):
    """Weights of interp2d/interp3d with two nodes (mathfunc.cpp:353-447):
    L0 = (X1-x0)/(X1-X0), L1 = (X0-x0)/(X0-X1), float."""
    return ((X1 - x0) / (X1 - X0)).astype(F32), ((X0 - x0) / (X0 - X1)).astype(F32)


def _volume_table(path, x, y, z, zwmedia):
    """read_volume_new (mediaparas.cpp:1273-1425): the volume (netCDF
    classic: coordinate variables x, y and z -- or depth when the reference
    is built with ZWmedia, as its Makefile does -- and Vp, Vs, rho on
    [x][y][z]) cut to the grid's coordinate range."""
    zname = "depth" if zwmedia else "z"
    with _read_netcdf(path, "media volume file") as f:
        X = np.asarray(f.variables["x"][:], F32)
        Y = np.asarray(f.variables["y"][:], F32)
        Z = np.asarray(f.variables[zname][:], F32)
        vp, vs, den = (np.asarray(f.variables[nm][:], F32).reshape(X.size, Y.size, Z.size) for nm in ("Vp", "Vs", "rho"))
    ni, nj, nk = (s - 2 * LFD for s in x.shape)
    v = (slice(LFD, LFD + ni), slice(LFD, LFD + nj), slice(LFD, LFD + nk))
    cut = []
    for arr, c in ((X, x), (Y, y), (Z, z)):
        lo, _ = loc_value_1d(np.array([c[v].min()], F32), arr)
        hi, _ = loc_value_1d(np.array([c[v].max()], F32), arr)
        cut.append(slice(int(lo[0]), int(hi[0]) + 2))
    return {"X": X[cut[0]], "Y": Y[cut[1]], "Z": Z[cut[2]],
            "F": [a[cut[0], cut[1], cut[2]] for a in (vp, vs, den)]}


def _volume_sample(vt, px, py, pz):
    """volume_discrete (mediaparas.cpp:376-441) at arrays of points:
    (Vp, Vs, Dc) by two-node Lagrange interpolation in each direction."""
    X, Y, Z = vt["X"], vt["Y"], vt["Z"]
    i1, xv = loc_value_1d(px, X)
    j1, yv = loc_value_1d(py, Y)
    k1, zv = loc_value_1d(pz, Z)
    lx = _lagrange2(X[i1], X[i1 + 1], xv)
    ly = _lagrange2(Y[j1], Y[j1 + 1], yv)
    lz = _lagrange2(Z[k1], Z[k1 + 1], zv)
    out = []
    for F in vt["F"]:
        acc = np.zeros(xv.shape, F32)
        for a in range(2):
            for b in range(2):
                for c in range(2):
                    w = (lx[a] * ly[b]) * lz[c]
                    acc = acc + w * F[i1 + a, j1 + b, k1 + c]
        out.append(acc)
    return out


def _get2in1(Vp, Vs, Dc):
    """get2in1 (mediaparas.cpp:872-886) with equal upper and lower values."""
    mu = (Dc * Vs) * Vs
    lam = (((Vp * Vp) * Dc).astype(np.float64) - 2.0 * mu.astype(np.float64)).astype(F32)
    density = ((Dc.astype(np.float64) + Dc) / 2.0).astype(F32)
    miu = ((mu.astype(np.float64) + mu) / 2.0).astype(F32)
    lamb = ((lam.astype(np.float64) + lam) / 2.0).astype(F32)
    return density, miu, lamb


def _vel3d(vp0, vp1, crit_value, crit_perc):
    """set_vel3d (mediaparas.cpp:916-923): the mean Vp, pushed up by 9000 km
    where the two sides differ by more than the thresholds."""
    vp0, vp1 = np.asarray(vp0, F32), np.asarray(vp1, F32)
    d = np.abs(vp0 - vp1)
    with np.errstate(divide="ignore", invalid="ignore"):
        smooth = (d < crit_value) & (d / np.minimum(vp0, vp1) < crit_perc)
    mean = (vp0 + vp1).astype(np.float64) / 2.0
    return np.where(smooth, mean, mean + 9000e3).astype(F32)


def _lag3(x0):
    """Three-node Lagrange weights on nodes 0, 2, 4 (interp3d with vec3,
    mathfunc.cpp:395-447), float."""
    X = (F32(0.0), F32(2.0), F32(4.0))
    L = []
    for i in range(3):
        w = F32(1.0)
        for j in range(3):
            q = F32(1.0) if j == i else F32(F32(X[j] - x0) / F32(X[j] - X[i]))
            w = F32(w * q)
        L.append(w)
    return L


def _two_in_one_acc(samples):
    """TwoinOne over samples, then average (mediaparas.cpp:838-914),
    element-wise over arrays of points.  `samples` is a list of
    (vp0, vs0, dc0, vp1, vs1, dc1) arrays (upper and lower side of each
    sample).  The reference's test `nsamp>0 && muval==0` compares pointers
    and is never true, so mu's harmonic sum restarts at 0 after any fluid
    sample.  Returns (density, mu, lambda)."""
    f64 = np.float64
    shape = np.shape(samples[0][0])
    rhoval, lamval, muval = np.zeros(shape, F32), np.zeros(shape, F32), np.zeros(shape, F32)
    for vp0, vs0, dc0, vp1, vs1, dc1 in samples:
        m0 = (dc0 * vs0) * vs0
        m1 = (dc1 * vs1) * vs1
        l0 = (((vp0 * vp0) * dc0).astype(f64) - 2.0 * m0.astype(f64)).astype(F32)
        l1 = (((vp1 * vp1) * dc1).astype(f64) - 2.0 * m1.astype(f64)).astype(F32)
        rhoval = (rhoval.astype(f64) + 0.5 * (dc0 + dc1).astype(f64)).astype(F32)
        fluid = (m0 < SEIS_ZERO) | (m1 < SEIS_ZERO)
        with np.errstate(divide="ignore"):
            acc = ((muval.astype(f64) + 0.5 / m0.astype(f64)) + 0.5 / m1.astype(f64)).astype(F32)
            lamval = ((lamval.astype(f64) + 0.5 / l0.astype(f64)) + 0.5 / l1.astype(f64)).astype(F32)
        muval = np.where(fluid, F32(0.0), acc).astype(F32)
    n = F32(len(samples))
    density = (rhoval / n).astype(F32)
    with np.errstate(divide="ignore"):
        mu = np.where(muval > 0, (f64(n) * 1.0 / muval.astype(f64)), 0.0).astype(F32)
        lamb = (f64(n) * 1.0 / lamval.astype(f64)).astype(F32)
    return density, mu, lamb


def _finish_media(density, miu, lamb, vel3d, x, y, z, smooth, sample_pairs):
    """ApplySmooth (mediaparas.cpp:1712-1824) when a sampling factor is above
    1, MediaStatistics (1837-1905) and the ghost extension of readdata.
    `smooth` = (sampx, sampy, sampz, crit_value, crit_perc);
    sample_pairs(x0, y0, z0) gives the (vp0, vs0, dc0, vp1, vs1, dc1) arrays
    of the upper and lower side at arrays of points."""
    ni, nj, nk = (s - 2 * LFD for s in x.shape)
    v = (slice(LFD, LFD + ni), slice(LFD, LFD + nj), slice(LFD, LFD + nk))
    if smooth is not None and max(smooth[:3]) > 1:
        sampx, sampy, sampz, cv, cp = smooth

        def over(a, b):
            d = np.abs(a - b)
            with np.errstate(divide="ignore", invalid="ignore"):
                return (d >= cv) | (d / np.minimum(a, b) >= cp)

        c = vel3d[v]
        ns = []
        for ax, samp in enumerate((sampx, sampy, sampz)):
            lo = [slice(LFD, LFD + ni), slice(LFD, LFD + nj), slice(LFD, LFD + nk)]
            hi = list(lo)
            lo[ax] = slice(LFD - 1, LFD - 1 + (ni, nj, nk)[ax])
            hi[ax] = slice(LFD + 1, LFD + 1 + (ni, nj, nk)[ax])
            ns.append(np.where(over(c, vel3d[tuple(lo)]) | over(c, vel3d[tuple(hi)]), samp, 1))
        flagged = (ns[0] > 1) | (ns[1] > 1) | (ns[2] > 1)
        counts = np.stack([n_[flagged] for n_ in ns], 1)
        cells = np.argwhere(flagged) + LFD
        for key in {tuple(r) for r in counts.tolist()}:
            sel = cells[(counts == key).all(1)]
            I, J, K = sel[:, 0], sel[:, 1], sel[:, 2]
            nx, ny, nz = key
            samples = []
            for mi in range(1, nx + 1):
                for mj in range(1, ny + 1):
                    for mk in range(1, nz + 1):
                        lx = _lag3(F32(1.0 + 2.0 / nx * (mi - 1 + 0.5)))
                        ly = _lag3(F32(1.0 + 2.0 / ny * (mj - 1 + 0.5)))
                        lz = _lag3(F32(1.0 + 2.0 / nz * (mk - 1 + 0.5)))
                        pos = []
                        for crd in (x, y, z):  # interp3d over the 3x3x3 neighbourhood, float
                            acc = np.zeros(len(sel), F32)
                            for a in range(3):
                                for b in range(3):
                                    for g in range(3):
                                        w = F32(F32(lx[a] * ly[b]) * lz[g])
                                        acc = acc + w * crd[I - 1 + a, J - 1 + b, K - 1 + g]
                            pos.append(acc)
                        samples.append(sample_pairs(*pos))
            density[I, J, K], miu[I, J, K], lamb[I, J, K] = _two_in_one_acc(samples)
    shape = x.shape
    alpha, beta, rho = np.zeros(shape, F32), np.zeros(shape, F32), np.zeros(shape, F32)
    beta[v] = np.sqrt(miu[v] / density[v])
    alpha[v] = np.sqrt((lamb[v].astype(np.float64) + 2.0 * miu[v].astype(np.float64)) / density[v]).astype(F32)
    rho[v] = density[v]
    for a in (alpha, beta, rho):
        _sc.equivalence_extend(a)
    return alpha, beta, rho


def volume_media(path, x, y, z, zwmedia=True, smooth=None):
    """Media of a volume file on the grid (readdata, mediaparas.cpp:2140-2219):
    volume_discrete at every point, get2in1, the optional smoothing, then
    MediaStatistics and the ghost extension.  Returns (alpha, beta, rho)."""
    vt = _volume_table(path, x, y, z, zwmedia)
    Vp, Vs, Dc = _volume_sample(vt, x, y, z)
    density, miu, lamb = _get2in1(Vp, Vs, Dc)
    vel3d = _vel3d(Vp, Vp, *smooth[3:]) if smooth is not None else None

    def sample_pairs(px, py, pz):
        vp, vs, dc = _volume_sample(vt, px, py, pz)
        return vp, vs, dc, vp, vs, dc

    return _finish_media(density, miu, lamb, vel3d, x, y, z, smooth, sample_pairs)


# ---------------------------------------------------------------------------
# 3-D interface media (mediaparas.cpp:52-374, 442-836, 838-914, 1837-1905)
# ---------------------------------------------------------------------------
def _interp2d_w(X, Y, x0, y0):
    """Two-node Lagrange weights of interp2d (mathfunc.cpp:353-393), float."""
    lx = (F32(F32(X[1] - x0) / F32(X[1] - X[0])), F32(F32(X[0] - x0) / F32(X[0] - X[1])))
    ly = (F32(F32(Y[1] - y0) / F32(Y[1] - Y[0])), F32(F32(Y[0] - y0) / F32(Y[0] - Y[1])))
    return lx, ly


def _interp2d(lx, ly, Z):
    acc = F32(0.0)
    for a in range(2):
        for b in range(2):
            acc = F32(acc + F32(F32(lx[a] * ly[b]) * Z[a, b]))
    return acc


def read_interface3d(path, nx, ny, x0, y0, dx, dy, nint):
    """An interface3D media file and the tables read_interface3D derives
    from it (mediaparas.cpp:52-260): depths Zpos[i][j][n] and Vp/Vs/Dc[m][i][j][n]
    (m = 0 above, 1 below interface n), flatness, constancy, layer depth
    ranges and the coincident-interface index table indexZ[n][i][j][2].
    x0/y0/dx/dy are the integers `get_conf` reads (atoi)."""
    c = ConfFile(path, "media interface")
    s = _Scanner([t for p in c.pieces for t in p.split()], path)
    Z = s.reals(nint * ny * nx).reshape(nint, ny, nx).transpose(2, 1, 0).copy()
    V = {}
    for nm in ("Vp", "Vs", "Dc"):
        V[nm] = s.reals(nint * 2 * ny * nx).reshape(nint, 2, ny, nx).transpose(1, 3, 2, 0).copy()
    bad = np.argwhere(Z[:, :, :-1] < Z[:, :, 1:])
    if len(bad):
        i, j, n = bad[0]
        raise ConfError(f"at location ({i},{j},{n}) interface depth[{n}] is lower than depth[{n + 1}]")
    vp, vs = V["Vp"].astype(np.float64), V["Vs"].astype(np.float64)
    air = np.zeros(vp.shape, bool)
    air[0, :, :, 0] = (V["Vp"][0, :, :, 0] == 0) & (V["Vs"][0, :, :, 0] == 0)
    if ((vp * vp <= 2.0 * vs * vs) & ~air).any():
        raise ConfError("there is a velocity error: Vp^2 <= 2 Vs^2")
    t = {"Z": Z, "V": V, "nint": nint,
         "X": np.array([F32(x0 + i * dx) for i in range(nx)], F32),
         "Y": np.array([F32(y0 + j * dy) for j in range(ny)], F32)}
    t["flat"] = [bool((np.abs(Z[0, 0, n] - Z[:, :, n]) <= SEIS_ZERO).all()) for n in range(nint)]
    const = np.zeros((nint, 2), bool)
    for n in range(nint):
        for m in range(2):
            const[n, m] = all((np.abs(V[nm][m, 0, 0, n] - V[nm][m, :, :, n]) <= SEIS_ZERO).all()
                              for nm in ("Vp", "Vs", "Dc"))
    t["const"] = const
    lconst = []
    for n in range(nint):
        if n == nint - 1 and const[n, 1]:
            lconst.append(True)
        elif const[n, 1] and n + 1 < nint and const[n + 1, 0] and all(
                abs(V[nm][1, 0, 0, n] - V[nm][0, 0, 0, n + 1]) < SEIS_ZERO for nm in ("Vp", "Vs", "Dc")):
            lconst.append(True)
        else:
            lconst.append(False)
    t["layer_const"] = lconst
    rng = np.zeros((nint, 2), F32)
    for n in range(nint - 1):
        rng[n] = (max(Z[0, 0, n], Z[:, :, n].max()), min(Z[0, 0, n + 1], Z[:, :, n + 1].min()))
    n = nint - 1
    rng[n] = (Z[:, :, n].max(), F32(float(Z[:, :, n].min()) - 6371e3))
    t["layer_range"] = rng
    idx = np.zeros((nint, nx, ny, 2), np.int64)
    for i in range(nx - 1):
        for j in range(ny - 1):
            corners = Z[i:i + 2, j:j + 2, :]
            idx[nint - 1, i, j, 1] = nint
            for n in range(nint - 2, -1, -1):
                h = np.abs(corners[:, :, n] - corners[:, :, n + 1]).max()
                idx[n, i, j, 1] = idx[n + 1, i, j, 1] if h < SEIS_ZERO else n + 1
            top_air = const[0, 0] and V["Vp"][0, 0, 0, 0] == 0
            idx[0, i, j, 0] = -idx[0, i, j, 1] if top_air else 1
            for n in range(1, nint):
                h = np.abs(corners[:, :, n] - corners[:, :, n - 1]).max()
                idx[n, i, j, 0] = idx[n - 1, i, j, 0] if h < SEIS_ZERO else n + 1
    idx[:, :, ny - 1, :] = idx[:, :, ny - 2, :]
    idx[:, nx - 1, :, :] = idx[:, nx - 2, :, :]
    t["indexZ"] = idx
    return t


def _i3d_column(t, x0, y0):
    """Everything interface3d_discrete needs at horizontal position (x0, y0):
    the cell (i1, j1), the interface depths there and each interface side's
    (Vp, Vs, Dc)."""
    i1, xv = loc_value_1d(np.array([x0], F32), t["X"])
    j1, yv = loc_value_1d(np.array([y0], F32), t["Y"])
    i1, j1, xv, yv = int(i1[0]), int(j1[0]), xv[0], yv[0]
    lx, ly = _interp2d_w(t["X"][i1:i1 + 2], t["Y"][j1:j1 + 2], xv, yv)
    nint = t["nint"]
    zs = [t["Z"][i1, j1, n] if t["flat"][n] else _interp2d(lx, ly, t["Z"][i1:i1 + 2, j1:j1 + 2, n])
          for n in range(nint)]
    side = {}
    for n in range(nint):
        for m in range(2):
            if t["const"][n, m]:
                side[n, m] = tuple(t["V"][nm][m, i1, j1, n] for nm in ("Vp", "Vs", "Dc"))
            else:
                side[n, m] = tuple(_interp2d(lx, ly, t["V"][nm][m, i1:i1 + 2, j1:j1 + 2, n])
                                   for nm in ("Vp", "Vs", "Dc"))
    inner = {}
    for n in range(nint):  # values interpolated regardless of constancy (layer interior)
        for m in range(2):
            inner[n, m] = tuple(_interp2d(lx, ly, t["V"][nm][m, i1:i1 + 2, j1:j1 + 2, n]) for nm in ("Vp", "Vs", "Dc"))
    return {"i1": i1, "j1": j1, "z": zs, "side": side, "inner": inner}


def _i3d_point(t, col, z0, prev):
    """interface3d_discrete (mediaparas.cpp:442-836) at depth z0 of a column:
    the (upper, lower) (Vp, Vs, Dc) pair; `prev` is returned when no layer
    matches, as the reference keeps the previous point's values."""
    nint, idx, i1, j1 = t["nint"], t["indexZ"], col["i1"], col["j1"]

    def upper(n):
        k, s = idx[n, i1, j1, 0], 0
        if k < 0:
            k, s = -k, 1
        return col["side"][k - 1, s]

    def lower(n):
        return col["side"][idx[n, i1, j1, 1] - 1, 1]

    n1, n2 = nint + 1, -1
    for n in range(nint):
        if t["layer_range"][n, 1] <= z0 <= t["layer_range"][n, 0]:
            n1, n2 = min(n1, n), max(n2, n + 1)
    n2 = min(n2, nint - 1)
    if n1 > n2:
        v = upper(0)
        return v, v
    for n in range(n1, n2 + 1):
        if idx[n, i1, j1, 1] != n + 1:
            continue
        z1 = col["z"][n]
        if abs(F32(z0 - z1)) < SEIS_ZERO:
            return upper(n), lower(n)
        if z0 > z1:
            v = upper(n)
            return v, v
        if n == nint - 1:
            v = lower(n)
            return v, v
        z2 = col["z"][n + 1]
        if abs(F32(z0 - z2)) < SEIS_ZERO:
            return upper(n + 1), lower(n + 1)
        if z0 > z2:
            if t["layer_const"][n]:
                v = tuple(t["V"][nm][1, i1, j1, n] for nm in ("Vp", "Vs", "Dc"))
                return v, v
            L = F32(z1 - z2)
            L1 = F32(F32(z0 - z2) / L)
            L2 = F32(F32(1.0) - L1)
            a, b = col["inner"][n, 1], col["inner"][n + 1, 0]
            v = tuple(F32(F32(a[q] * L1) + F32(b[q] * L2)) for q in range(3))
            return v, v
    return prev


def interface3d_media(path, nx, ny, x0, y0, dx, dy, nint, x, y, z, smooth=None):
    """Media of an interface3D file on the grid: interface3d_discrete at every
    point in the reference's loop order (a point with no matching layer keeps
    the previous point's values), TwoinOne/average, the optional smoothing,
    MediaStatistics and the ghost extension.  Returns (alpha, beta, rho).
    Pure Python per point: meant for grids up to a few million cells."""
    t = read_interface3d(path, nx, ny, x0, y0, dx, dy, nint)
    shape = x.shape
    side = [np.zeros(shape, F32) for _ in range(6)]   # vp0 vs0 dc0 vp1 vs1 dc1
    prev = ((F32(0.0),) * 3, (F32(0.0),) * 3)
    cols = {}
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                key = (x[i, j, k], y[i, j, k])
                if key not in cols:
                    cols[key] = _i3d_column(t, *key)
                prev = _i3d_point(t, cols[key], z[i, j, k], prev)
                for q in range(6):
                    side[q][i, j, k] = prev[q // 3][q % 3]
        cols.clear()
    density, miu, lamb = _two_in_one_acc([tuple(side)])
    vel3d = _vel3d(side[0], side[3], *smooth[3:]) if smooth is not None else None

    def sample_pairs(px, py, pz):
        out = [np.zeros(len(px), F32) for _ in range(6)]
        for n in range(len(px)):
            pr = _i3d_point(t, _i3d_column(t, px[n], py[n]), pz[n], prev)
            for q in range(6):
                out[q][n] = pr[q // 3][q % 3]
        return tuple(out)

    return _finish_media(density, miu, lamb, vel3d, x, y, z, smooth, sample_pairs)


# ---------------------------------------------------------------------------
# the whole run
# ---------------------------------------------------------------------------
def load(conf_path, seispath=None, media=None, zwmedia=True, simplecar=False):
    """Scenario dict for `cgfd3d.Simulation.setup` from a SeisFD3D.conf.

    `seispath` overrides the directory the conf file names; `media` =
    (alpha, beta, rho) arrays with ghosts replace the configured media;
    `simplecar` asks for the reference's SimpleCar build (a compile-time
    option there: xi/eta metrics taken as the identity).
    The dict also holds the receiver metadata (`recv_pos`, `recv_pnum`,
    `recv_lnum`, `seismo_tinv`) that `Simulation.write_seismo` takes."""
    m = read_main(conf_path, seispath)
    ni, nj, nk, nt = m["ni"], m["nj"], m["nk"], m["nt"]
    if m["grid_type"] == "point":
        raise ConfError("grid_type point: the reference's read_point stores x, y and z all into crd.x "
                        "(gridmesh.cpp:52-68), so it yields no valid grid; use a vmap grid")
    if m["grid_type"] != "vmap":
        raise ConfError(f"gridtype configure {m['grid_type']} is wrong when reading grid data")
    posx, posy, posz, gp, eq = read_vmap(m["grid_filename"], ni, nj, nk)
    x, y, z, conindex = vmap_coords(posx, posy, posz, gp, eq, nk,
                                    m["conversion_depth"] if m["hybrid_grid"] else None)
    met = _sc.metrics_from_coords(x, y, z, m["steph"])

    smooth = (*m["sampling"], m["crit_value"], m["crit_perc"])
    if media is not None:
        alpha, beta, rho = (np.asarray(a, F32) for a in media)
    elif m["media_type"] == "volume":
        alpha, beta, rho = volume_media(m["media_filename"], x, y, z, zwmedia, smooth)
    elif m["media_type"] == "interface3D":
        c = ConfFile(conf_path, "main par")
        par = [c.int("interface3dpar", 3 + q) for q in range(7)]
        alpha, beta, rho = interface3d_media(m["media_filename"], *par, x, y, z, smooth)
    elif m["media_type"] == "interface":
        raise ConfError("media_type interface: the reference's readdata overwrites the interface media with zero "
                        "Vp/Vs/rho (get2in1, mediaparas.cpp:2163-2200) and MediaStatistics turns them into NaN; "
                        "use volume media or pass media arrays")
    else:
        raise ConfError(f"media_type {m['media_type']} is not supported by this front end; pass media arrays")

    nabs, vel, bmax, amax = read_absorb(m["absorb_filename"])
    pml = _sc.cfs_pml_tables(ni, nj, nk, nabs, vel, bmax, amax, m["steph"], met)

    src = read_sources(m["source_filename"], m["seispath"], nt, m["stept"])
    sc = dict(ni=ni, nj=nj, nk=nk, nt=nt, dh=float(m["steph"]), dt=float(m["stept"]), metrics=met,
              alpha=alpha, beta=beta, rho=rho, nabs=nabs, pml=pml, pgv=bool(m["output_peakvel"]),
              coords=(x, y, z), name=f"conf-{os.path.basename(os.path.dirname(os.path.abspath(conf_path)))}")
    if m["hybrid_grid"]:
        sc["conindex"] = conindex
    if simplecar:
        sc["simplecar"] = True
    sc.update(locate_sources(src, x, y, z, rho, beta))
    st = read_stations(m["station_filename"])
    sc["recv_locs"] = locate_stations(st, x, y, z)
    sc["recv_pos"], sc["recv_pnum"], sc["recv_lnum"], sc["seismo_tinv"] = st["pos"], st["pnum"], st["lnum"], st["tinv"]
    sc["snapshots"] = st["snapshots"]
    return sc
