"""ctypes binding of libcgfd3d.so, the gfx950 time-stepping engine.

This is the host-side mirror of the reference's ChildProcs interface
(src/typedata.h:114-217 of fes478/CoPara-CGFD3D) for Python callers: the same
inputs (metrics, media, CFS-PML tables, point sources, receivers) in the
reference's [x][y][z]+3-ghost layout, the same stepping entry points and the
same outputs (receiver traces, peak ground velocity, wavefield).  All compute
runs in the HIP library; loading fails loudly if it has not been built.

A process that also uses torch on the GPU (the surrogate) must import torch
before this library is loaded: both bind libamdhip64.so.7, the first one
loaded serves the whole process, and torch initialises only on the runtime it
ships with (the engine runs on either).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CGFD3D_LIB", os.path.join(HERE, "libcgfd3d.so"))

FP32, FP64 = 0, 1
COMPONENTS = ("Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")
METRICS = ("xix", "xiy", "xiz", "etax", "etay", "etaz", "zetax", "zetay", "zetaz", "jac")


class CgfdError(RuntimeError):
    pass


class Dims(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int), ("nk", ctypes.c_int),
                ("dh", ctypes.c_double), ("dt", ctypes.c_double), ("nt", ctypes.c_int)]


_LIB = None

# exported symbol -> (restype, argtypes); the list is the C-ABI of include/cgfd3d.h
_VP = ctypes.c_void_p
_IP = ctypes.POINTER(ctypes.c_int)
SIGNATURES = {
    "cgfd_create": (ctypes.c_int, [ctypes.POINTER(Dims), ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP)]),
    "cgfd_destroy": (None, [_VP]),
    "cgfd_last_error": (ctypes.c_char_p, [_VP]),
    "cgfd_set_metrics": (ctypes.c_int, [_VP, ctypes.POINTER(_VP)]),
    "cgfd_set_media": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "cgfd_set_pml": (ctypes.c_int, [_VP, _IP, ctypes.POINTER(_VP)]),
    "cgfd_set_forces": (ctypes.c_int, [_VP, ctypes.c_int, _IP, _VP, _VP, ctypes.c_int]),
    "cgfd_set_moments": (ctypes.c_int, [_VP, ctypes.c_int, _IP, _VP, _VP, ctypes.c_int]),
    "cgfd_set_focal": (ctypes.c_int, [_VP, ctypes.c_int, _IP, _VP, ctypes.c_int, ctypes.c_double]),
    "cgfd_set_receivers": (ctypes.c_int, [_VP, ctypes.c_int, _IP]),
    "cgfd_set_snapshots": (ctypes.c_int, [_VP, ctypes.c_int, _IP]),
    "cgfd_snap_shape": (ctypes.c_int, [_VP, ctypes.c_int, _IP, _IP]),
    "cgfd_get_snap": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _VP]),
    "cgfd_set_pgv": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cgfd_set_hybrid": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cgfd_set_simplecar": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cgfd_share_medium": (ctypes.c_int, [_VP, _VP]),
    "cgfd_set_tile_width": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cgfd_tile_width": (ctypes.c_int, [_VP]),
    "cgfd_init": (ctypes.c_int, [_VP]),
    "cgfd_run": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cgfd_run_stages": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cgfd_sync": (ctypes.c_int, [_VP]),
    "cgfd_set_graph": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cgfd_run_batch": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, ctypes.c_int]),
    "cgfd_run_batch_static": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, ctypes.c_int]),
    "cgfd_batch_fusable": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int]),
    "cgfd_autotune": (ctypes.c_int, [_VP, ctypes.c_int, _IP]),
    "cgfd_current_step": (ctypes.c_int, [_VP]),
    "cgfd_current_stage": (ctypes.c_int, [_VP]),
    "cgfd_get_field": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _VP]),
    "cgfd_set_field": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _VP]),
    "cgfd_aux_size": (ctypes.c_long, [_VP, ctypes.c_int]),
    "cgfd_get_aux": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP]),
    "cgfd_set_aux": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP]),
    "cgfd_get_seismo": (ctypes.c_int, [_VP, _VP]),
    "cgfd_get_pgv": (ctypes.c_int, [_VP, _VP]),
    "cgfd_get_cov": (ctypes.c_int, [_VP, _VP, _VP]),
    "cgfd_profile": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cgfd_profile_read": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long),
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]),
    "cgfd_set_tuning": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    "cgfd_set_variant": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cgfd_stage_bytes": (ctypes.c_double, [_VP]),
    "cgfd_metrics_specialized": (ctypes.c_int, [_VP]),
    "cgfd_set_specialize": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cgfd_nc_seismo": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP),
                                      _VP, _IP, _IP, _VP, _VP, _VP]),
    "cgfd_nc_snap": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _IP, ctypes.c_int, ctypes.POINTER(_VP), _VP]),
    "cgfd_nc_peakvel": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _VP, _VP]),
    "cgfd_write_seismo": (ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_int, _IP, _IP, _VP, _VP, _VP]),
    "cgfd_write_snap": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_char_p]),
    "cgfd_write_peakvel": (ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_int]),
    "cgfd_stream": (_VP, [_VP]),
}


def load_library(path=LIB_PATH):
    """Load libcgfd3d.so; raises if it is missing (there is no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise CgfdError(f"{path} not built: run copara-cgfd3d_amd/build.sh (or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _prec_of(a):
    return FP64 if np.asarray(a).dtype == np.float64 else FP32


def nc_seismo(path, wave, currtime, pnum, lnum, posx, posy, posz):
    """seismo.nc from host arrays: wave [9][nrec][npnt], currtime [nrec]."""
    lib = load_library()
    dt = np.float64 if _prec_of(wave) == FP64 else np.float32
    w = np.ascontiguousarray(wave, dtype=dt)
    nrec, npnt = w.shape[1], w.shape[2]
    ptrs = (_VP * 9)(*[w[c].ctypes.data for c in range(9)])
    arrs = [np.ascontiguousarray(a, dtype=dt) for a in (currtime, posx, posy, posz)]
    ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (pnum, lnum)]
    rc = lib.cgfd_nc_seismo(os.fsencode(path), _prec_of(wave), nrec, npnt, ptrs, _ptr(arrs[0]),
                            ints[0].ctypes.data_as(_IP), ints[1].ctypes.data_as(_IP),
                            _ptr(arrs[1]), _ptr(arrs[2]), _ptr(arrs[3]))
    if rc:
        raise CgfdError(f"cgfd_nc_seismo failed ({rc})")


def nc_snap(path, spec, wave, currtime):
    """snap_NNN.nc from host arrays: wave maps component -> [nrec][xn][yn][zn]."""
    lib = load_library()
    first = next(iter(wave.values()))
    dt = np.float64 if _prec_of(first) == FP64 else np.float32
    keep = {c: np.ascontiguousarray(a, dtype=dt) for c, a in wave.items()}
    ptrs = (_VP * 9)(*[keep[c].ctypes.data if c in keep else None for c in range(9)])
    sp = np.ascontiguousarray(spec, dtype=np.int32)
    t = np.ascontiguousarray(currtime, dtype=dt)
    nrec = t.shape[0]
    rc = lib.cgfd_nc_snap(os.fsencode(path), _prec_of(first), sp.ctypes.data_as(_IP), nrec, ptrs, _ptr(t))
    if rc:
        raise CgfdError(f"cgfd_nc_snap failed ({rc})")


def nc_peakvel(path, vx, vy, vz):
    """peakvel.nc from host [ni][nj] maps."""
    lib = load_library()
    dt = np.float64 if _prec_of(vx) == FP64 else np.float32
    a = [np.ascontiguousarray(v, dtype=dt) for v in (vx, vy, vz)]
    ni, nj = a[0].shape
    rc = lib.cgfd_nc_peakvel(os.fsencode(path), _prec_of(vx), ni, nj, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]))
    if rc:
        raise CgfdError(f"cgfd_nc_peakvel failed ({rc})")


def run_batch(sims, nsteps):
    """Advance several Simulations by nsteps together: one launch per kernel
    for all of them when batch_fusable(sims), else one stream each."""
    if not sims:
        return
    lib = sims[0].lib
    arr = (_VP * len(sims))(*[s.h for s in sims])
    rc = lib.cgfd_run_batch(arr, len(sims), int(nsteps))
    if rc:
        msgs = "; ".join(lib.cgfd_last_error(s.h).decode() for s in sims)
        raise CgfdError(f"cgfd_run_batch failed ({rc}): {msgs}")


def run_batch_static(sims, nsteps):
    """run_batch's fused path without any device allocation (the library's
    fixed per-device array holds the batch's argument images): at most 64
    Simulations that batch_fusable accepts."""
    if not sims:
        return
    lib = sims[0].lib
    arr = (_VP * len(sims))(*[s.h for s in sims])
    rc = lib.cgfd_run_batch_static(arr, len(sims), int(nsteps))
    if rc:
        raise CgfdError(f"cgfd_run_batch_static failed ({rc}): {lib.cgfd_last_error(sims[0].h).decode()}")


def batch_fusable(sims):
    """True if run_batch advances these Simulations with one launch per
    kernel (same device and precision, default stage variant, same metric
    mode and step phase, no kernel timing; sizes may differ)."""
    if not sims:
        return False
    lib = sims[0].lib
    arr = (_VP * len(sims))(*[s.h for s in sims])
    rc = lib.cgfd_batch_fusable(arr, len(sims))
    if rc < 0:
        raise CgfdError(f"cgfd_batch_fusable failed ({rc})")
    return bool(rc)


class Simulation:
    """One simulation context on one GPU (the reference's ChildProcs)."""

    def __init__(self, ni, nj, nk, dh, dt, nt, precision=FP32, device=0):
        self.lib = load_library()
        self.ni, self.nj, self.nk, self.nt = ni, nj, nk, nt
        self.dtype = np.float32 if precision == FP32 else np.float64
        self.shape = (ni + 6, nj + 6, nk + 6)
        dims = Dims(ni, nj, nk, float(dh), float(dt), nt)
        h = ctypes.c_void_p()
        rc = self.lib.cgfd_create(ctypes.byref(dims), precision, device, ctypes.byref(h))
        if rc != 0:
            raise CgfdError(f"cgfd_create failed with code {rc}")
        self.h = h
        self.nrecv = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.cgfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.cgfd_last_error(self.h)
            raise CgfdError(f"{what}: code {rc}: {msg.decode() if msg else ''}")

    def _arr(self, a, shape=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if shape is not None and a.shape != shape:
            raise CgfdError(f"expected shape {shape}, got {a.shape}")
        return a

    # ---- inputs -----------------------------------------------------------
    def set_metrics(self, drv):
        arrs = [self._arr(drv[k], self.shape) for k in METRICS]
        ptrs = (ctypes.c_void_p * 10)(*[a.ctypes.data for a in arrs])
        self._check(self.lib.cgfd_set_metrics(self.h, ptrs), "set_metrics")

    def set_media(self, alpha, beta, rho):
        a, b, r = (self._arr(v, self.shape) for v in (alpha, beta, rho))
        self._check(self.lib.cgfd_set_media(self.h, _ptr(a), _ptr(b), _ptr(r)), "set_media")

    def set_pml(self, nabs, tabs):
        """tabs: dict with APDx APDy APDz Bx By Bz DBx DBy DBz (lengths NX, NY, NZ)."""
        order = ("APDx", "APDy", "APDz", "Bx", "By", "Bz", "DBx", "DBy", "DBz")
        arrs = [self._arr(tabs[k]) for k in order]
        nab = (ctypes.c_int * 6)(*[int(v) for v in nabs])
        ptrs = (ctypes.c_void_p * 9)(*[a.ctypes.data for a in arrs])
        self._check(self.lib.cgfd_set_pml(self.h, nab, ptrs), "set_pml")

    def set_forces(self, locs, vec, stf):
        locs = np.ascontiguousarray(locs, dtype=np.int32).reshape(-1, 3)
        n = locs.shape[0]
        vec, stf = self._arr(np.reshape(vec, (n, 3))), self._arr(np.reshape(stf, (n, -1)))
        self._check(self.lib.cgfd_set_forces(self.h, n, locs.ctypes.data_as(_IP), _ptr(vec), _ptr(stf),
                                             stf.shape[1]), "set_forces")

    def set_moments(self, locs, vec, stf):
        locs = np.ascontiguousarray(locs, dtype=np.int32).reshape(-1, 3)
        n = locs.shape[0]
        vec, stf = self._arr(np.reshape(vec, (n, 6))), self._arr(np.reshape(stf, (n, -1)))
        self._check(self.lib.cgfd_set_moments(self.h, n, locs.ctypes.data_as(_IP), _ptr(vec), _ptr(stf),
                                              stf.shape[1]), "set_moments")

    def set_focal(self, locs, series, fdt):
        """Focal points (the reference's source_type 1): series[6][n][fnt]
        moment-rate samples (mxx myy mzz mxy mxz myz) every fdt s from t = 0."""
        locs = np.ascontiguousarray(locs, dtype=np.int32).reshape(-1, 3)
        n = locs.shape[0]
        ser = self._arr(np.reshape(series, (6, n, -1)))
        self._check(self.lib.cgfd_set_focal(self.h, n, locs.ctypes.data_as(_IP), _ptr(ser), ser.shape[2],
                                            float(fdt)), "set_focal")

    def set_receivers(self, locs):
        locs = np.ascontiguousarray(locs, dtype=np.int32).reshape(-1, 3)
        self.nrecv = locs.shape[0]
        self._check(self.lib.cgfd_set_receivers(self.h, self.nrecv, locs.ctypes.data_as(_IP)), "set_receivers")

    def set_snapshots(self, specs):
        """specs: one row per snapshot box, xs ys zs xn yn zn xi yi zi tinv cmp
        (the snap_NNN line of the reference's seismic-output configuration)."""
        spec = np.ascontiguousarray(specs, dtype=np.int32).reshape(-1, 11)
        self.nsnap = spec.shape[0]
        self._check(self.lib.cgfd_set_snapshots(self.h, self.nsnap, spec.ctypes.data_as(_IP)), "set_snapshots")

    def snap_shape(self, isnap):
        n, t = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.cgfd_snap_shape(self.h, isnap, ctypes.byref(n), ctypes.byref(t)), "snap_shape")
        return n.value, t.value

    def get_snap(self, isnap, comp):
        """Component comp of snapshot isnap as [npoints][ntime]."""
        out = np.empty(self.snap_shape(isnap), dtype=self.dtype)
        self._check(self.lib.cgfd_get_snap(self.h, isnap, comp, _ptr(out)), "get_snap")
        return out

    def set_pgv(self, on=True):
        self._check(self.lib.cgfd_set_pgv(self.h, 1 if on else 0), "set_pgv")

    def set_hybrid(self, conindex):
        """hybrid_grid = 1: the reference's ConIndex (z index with the 3 ghost
        planes); below it the Cartesian contraction applies."""
        self._check(self.lib.cgfd_set_hybrid(self.h, int(conindex)), "set_hybrid")

    def share_medium(self, src):
        """Use src's metrics and media (device arrays) instead of a copy,
        before this context's own metrics and media.  The arrays stay until
        the last context holding them is closed."""
        self._check(self.lib.cgfd_share_medium(self.h, src.h), "share_medium")

    def set_simplecar(self, on=True):
        """The reference's SimpleCar build: xi/eta metrics taken as the
        identity (call before set_metrics)."""
        self._check(self.lib.cgfd_set_simplecar(self.h, int(bool(on))), "set_simplecar")

    def init(self):
        self._check(self.lib.cgfd_init(self.h), "init")

    def setup(self, sc):
        """Load every input of a scenario dict (see scenario.py) and init."""
        if sc.get("simplecar"):
            self.set_simplecar(True)
        if sc.get("medium_of") is not None:
            self.share_medium(sc["medium_of"])
        else:
            self.set_metrics(sc["metrics"])
            self.set_media(sc["alpha"], sc["beta"], sc["rho"])
        self.set_pml(sc["nabs"], sc["pml"])
        if sc.get("force_locs") is not None and len(sc["force_locs"]):
            self.set_forces(sc["force_locs"], sc["force_vec"], sc["force_stf"])
        if sc.get("moment_locs") is not None and len(sc["moment_locs"]):
            self.set_moments(sc["moment_locs"], sc["moment_vec"], sc["moment_stf"])
        if sc.get("focal_locs") is not None and len(sc["focal_locs"]):
            self.set_focal(sc["focal_locs"], sc["focal_series"], sc["focal_dt"])
        if sc.get("recv_locs") is not None and len(sc["recv_locs"]):
            self.set_receivers(sc["recv_locs"])
        if sc.get("snapshots") is not None and len(sc["snapshots"]):
            self.set_snapshots(sc["snapshots"])
        self.set_pgv(sc.get("pgv", True))
        if sc.get("conindex") is not None:
            self.set_hybrid(sc["conindex"])
        self.init()

    # ---- stepping ---------------------------------------------------------
    def run(self, nsteps):
        self._check(self.lib.cgfd_run(self.h, int(nsteps)), "run")

    def autotune(self, trial_steps=3):
        """Pick the fastest z-chunk count for this grid (and, with trial_steps >= 4,
        whether graph replay is faster); re-initialises."""
        z = ctypes.c_int(0)
        self._check(self.lib.cgfd_autotune(self.h, int(trial_steps), ctypes.byref(z)), "autotune")
        return z.value

    def run_stages(self, nstages):
        self._check(self.lib.cgfd_run_stages(self.h, int(nstages)), "run_stages")

    def set_graph(self, on=True):
        """Replay captured 4-step cycles in run() (small, launch-bound grids)."""
        self._check(self.lib.cgfd_set_graph(self.h, int(bool(on))), "set_graph")

    def sync(self):
        self._check(self.lib.cgfd_sync(self.h), "sync")

    @property
    def step(self):
        return self.lib.cgfd_current_step(self.h)

    @property
    def stage(self):
        return self.lib.cgfd_current_stage(self.h)

    # ---- state and outputs -----------------------------------------------
    def get_field(self, comp, which=0):
        out = np.empty(self.shape, dtype=self.dtype)
        self._check(self.lib.cgfd_get_field(self.h, which, comp, _ptr(out)), "get_field")
        return out

    def set_field(self, comp, arr, which=0):
        a = self._arr(arr, self.shape)
        self._check(self.lib.cgfd_set_field(self.h, which, comp, _ptr(a)), "set_field")

    def aux_size(self, axis):
        return self.lib.cgfd_aux_size(self.h, axis)

    def get_aux(self, axis, comp, which=0):
        out = np.empty(max(self.aux_size(axis), 1), dtype=self.dtype)
        self._check(self.lib.cgfd_get_aux(self.h, which, axis, comp, _ptr(out)), "get_aux")
        return out[: self.aux_size(axis)]

    def set_aux(self, axis, comp, arr, which=0):
        a = self._arr(arr)
        if a.size != self.aux_size(axis):
            raise CgfdError("aux size mismatch")
        self._check(self.lib.cgfd_set_aux(self.h, which, axis, comp, _ptr(a)), "set_aux")

    def get_seismo(self):
        out = np.zeros((9, self.nrecv, self.nt), dtype=self.dtype)
        if self.nrecv:
            self._check(self.lib.cgfd_get_seismo(self.h, _ptr(out)), "get_seismo")
        return out

    def get_pgv(self):
        out = np.empty((3, self.shape[0], self.shape[1]), dtype=self.dtype)
        self._check(self.lib.cgfd_get_pgv(self.h, _ptr(out)), "get_pgv")
        return out

    def get_cov(self):
        cx = np.empty((self.shape[0], self.shape[1], 9), dtype=self.dtype)
        cy = np.empty_like(cx)
        self._check(self.lib.cgfd_get_cov(self.h, _ptr(cx), _ptr(cy)), "get_cov")
        return cx, cy

    # ---- measurement ------------------------------------------------------
    def profile(self, on=True):
        self._check(self.lib.cgfd_profile(self.h, 1 if on else 0), "profile")

    def profile_read(self):
        ms_s, n_s, ms_t, n_t = ctypes.c_double(), ctypes.c_long(), ctypes.c_double(), ctypes.c_long()
        self._check(self.lib.cgfd_profile_read(self.h, ctypes.byref(ms_s), ctypes.byref(n_s), ctypes.byref(ms_t),
                                               ctypes.byref(n_t)), "profile_read")
        return {"stage_ms": ms_s.value, "stage_n": n_s.value, "timg_ms": ms_t.value, "timg_n": n_t.value}

    def set_tuning(self, zchunks=0, rows=0):
        self._check(self.lib.cgfd_set_tuning(self.h, int(zchunks), int(rows)), "set_tuning")

    def set_variant(self, variant):
        self._check(self.lib.cgfd_set_variant(self.h, int(variant)), "set_variant")

    def set_tile_width(self, width):
        """Tile width of the default stage kernel: 64, 32 or 0 (the one that
        leaves fewer lanes idle on this grid's rows)."""
        self._check(self.lib.cgfd_set_tile_width(self.h, int(width)), "set_tile_width")

    def tile_width(self):
        return int(self.lib.cgfd_tile_width(self.h))

    # ---- NetCDF outputs (seismo.nc, snap_NNN.nc, peakvel.nc) ---------------
    def write_seismo(self, path, tinv=1, pnum=None, lnum=None, posx=None, posy=None, posz=None):
        ints = [None if a is None else np.ascontiguousarray(a, dtype=np.int32) for a in (pnum, lnum)]
        reals = [None if a is None else self._arr(a) for a in (posx, posy, posz)]
        self._check(self.lib.cgfd_write_seismo(
            self.h, os.fsencode(path), int(tinv),
            *[None if a is None else a.ctypes.data_as(_IP) for a in ints],
            *[None if a is None else _ptr(a) for a in reals]), "write_seismo")

    def write_snap(self, isnap, path):
        self._check(self.lib.cgfd_write_snap(self.h, isnap, os.fsencode(path)), "write_snap")

    def write_peakvel(self, path, vz_from_vy=True):
        self._check(self.lib.cgfd_write_peakvel(self.h, os.fsencode(path), 1 if vz_from_vy else 0), "write_peakvel")

    def metrics_specialized(self):
        """True when the stage kernel runs with constant horizontal metrics."""
        return bool(self.lib.cgfd_metrics_specialized(self.h))

    def set_specialize(self, on=True):
        self._check(self.lib.cgfd_set_specialize(self.h, 1 if on else 0), "set_specialize")

    def stage_bytes(self):
        return self.lib.cgfd_stage_bytes(self.h)
