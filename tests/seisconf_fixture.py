"""Writes a small, complete parameter-file set in the reference's formats
(SeisFD3D.conf, vmap grid, absorb, source, station files and a netCDF
classic volume-media file) for the parameter-file front end's tests."""
import numpy as np
from scipy.io import netcdf_file

NI, NJ, NK = 24, 20, 16
DH = 100.0


def surfaces(ni=NI, nj=NJ):
    """Four vmap surfaces per column: a topography with a bump, two flat
    interfaces and a flat bottom."""
    x = np.arange(ni) * DH
    y = np.arange(nj) * DH
    X, Y = np.meshgrid(x, y, indexing="ij")
    top = 150.0 * np.exp(-((X - 1200.0) ** 2 + (Y - 900.0) ** 2) / (600.0 ** 2))
    return X, Y, [top, np.full_like(X, -500.0), np.full_like(X, -900.0), np.full_like(X, -1500.0)]


def write_set(d, nt=12, media=True, source="moment", hybrid=0, gp=(5, 4, 6), eq=(1, 0, 1), extra_conf="",
              conv="-3.0E3"):
    d.mkdir(parents=True, exist_ok=True)
    X, Y, surf = surfaces()
    (d / "SeisFD3D.conf").write_text(f"""# main parameters
ni    = {NI}
nj    = {NJ}
nk    = {NK}
nt    = {nt}
stept = 0.008
steph = {DH}
seispath = {d}
grid_type = vmap  # vmap or point
grid_filename = grid.dat
hybrid_grid = {hybrid}
conversion_interface_depth = {conv}
reference_velocity = 2000.0
media_type = volume # interface, interface3D or volume
media_filename = vol.nc
threshold_velocity_jump_of_interface = 100
threshold_percent_change_of_interface = 0.1
sampling_point_per_cell = 0 0 0
source_filename = source.dat
station_filename = station.dat
output_peakvel = 1 # 1 YES 0 NO
absorb_filename = absorb.dat
{extra_conf}
""")
    lines = ["#grid vmap ascii data", f"vmap_dims = {NI} {NJ} 4",
             "vmap_gridpoints = " + " ".join(map(str, gp)),
             "vmap_equalspacing = " + " ".join(map(str, eq)), "<vmap_anchor>"]
    for i in range(NI):
        for j in range(NJ):
            lines.append(f"{X[i, j]:.6f} {Y[i, j]:.6f} " + " ".join(f"{s[i, j]:.6f}" for s in surf))
    (d / "grid.dat").write_text("\n".join(lines) + "\n")
    (d / "absorb.dat").write_text("""#X1 X2 Y1 Y2 Z1 Z2
abs_number = 4 4 4 4 4 0
abs_velocity  = 4000 4000 4000 4000 4000 4000
bmax = 2.0 2.0 2.0 2.0 2.0 2.0
amax = 6.0 6.0 6.0 6.0 6.0 6.0
""")
    if source == "focal":
        (d / "source.dat").write_text("""source_hyper_height = 9.0
distance2meter = 1.0E3
source_type = 1 # 1-> focal ; 2-> force ; 3-> moment
simulate_damp = 0.5
focal_process_NC_file = focal.nc
""")
        write_focal(d / "focal.nc")
    if source == "moment":
        (d / "source.dat").write_text("""source_hyper_height = 9.0
distance2meter = 1.0E3
source_type = 3 # 1-> focal ; 2-> force ; 3-> moment
simulate_damp = 1.0
number_of_force_source = 1
force_stf_window = 1
force_stf_type = Ricker
force_stf_timefactor = 0.05
force_stf_freqfactor = 12.0
<anchor_force>
  1.1 0.9 -0.7   0.0    1.0e+12    0.0 0.0 1.0
number_of_moment_source = 2
moment_stf_window = 1
moment_stf_type = gauss
moment_stf_timefactor = 0.04
moment_stf_freqfactor = 0.012
moment_mech_input = angle  # moment, angle
<anchor_moment>
  0.8 0.7 -0.6  0.0   1.0e+10    30 60 90
  1.5 1.2 -1.0  0.008   2.0e+10    120 45 -30
""")
    (d / "station.dat").write_text("""number_of_snap = 2
# id        start          count           stride  tinv  cmp
snap_001 =  1   1   16    24 20 1       1 1 1   3     3
snap_002 =  2   3   4     5  4  6       4 4 2   2     1
topo_hyper_height = 8000.0E3
tinv_of_seismo = 1
number_of_inline = 1
line_001 =  0.3E3  0.5E3  9000E3 | 0.4E3 0.2E3 0.0 | 3
number_of_recv = 3
recv_001 =  1.2E3  0.9E3  9000E3
recv_002 =  0.6E3  1.3E3  9000E3
recv_003 =  1.0E3  1.0E3 -0.55E3
""")
    if media:
        write_volume(d / "vol.nc")
    return d / "SeisFD3D.conf"


def vol_fields(x, y, z):
    """Linear in x, y and z, so trilinear interpolation reproduces it."""
    vp = 3000.0 + 0.4 * x - 0.3 * y - 1.0 * z
    vs = vp / 1.8
    rho = 2200.0 + 0.1 * x + 0.05 * y - 0.2 * z
    return vp, vs, rho


def write_volume(path, zname="depth"):
    xs = np.arange(-300.0, 2700.0, 250.0)
    ys = np.arange(-300.0, 2300.0, 250.0)
    zs = np.arange(-2000.0, 600.0, 125.0)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    vp, vs, rho = vol_fields(X, Y, Z)
    with netcdf_file(str(path), "w", version=2) as f:
        for nm, a in (("x", xs), ("y", ys), (zname, zs)):
            f.createDimension(nm, a.size)
            v = f.createVariable(nm, "f4", (nm,))
            v[:] = a.astype(np.float32)
        for nm, a in (("Vp", vp), ("Vs", vs), ("rho", rho)):
            v = f.createVariable(nm, "f4", ("x", "y", zname))
            v[:] = a.astype(np.float32)


def write_focal(path, nfnt=40, fdt=0.005):
    """A focal-process file: three points (one beyond the grid's x range), moment-rate
    series on [NP][NT], DT as a global attribute, positions in metres."""
    pos = np.array([[900.0, 800.0, -700.0], [1400.0, 1000.0, -1100.0], [9000.0, 800.0, -700.0]], np.float32)
    t = np.arange(nfnt) * fdt
    rate = np.exp(-((t - 0.03) / 0.01) ** 2)
    comps = {"Mxx": 1.0, "Myy": -0.5, "Mzz": -0.5, "Mxy": 0.3, "Mxz": 0.0, "Myz": 0.2}
    with netcdf_file(str(path), "w", version=2) as f:
        f.createDimension("NP", len(pos))
        f.createDimension("NT", nfnt)
        f.DT = np.float32(fdt)
        for n, (nm, c) in enumerate(comps.items()):
            v = f.createVariable(nm, "f4", ("NP", "NT"))
            v[:] = (c * 1e3 * np.outer(np.arange(1, len(pos) + 1), rate)).astype(np.float32)
        for c, nm in enumerate(("posx", "posy", "posz")):
            v = f.createVariable(nm, "f4", ("NP",))
            v[:] = pos[:, c]


I3D = dict(nx=7, ny=6, x0=-500, y0=-500, dx=500, dy=500)
ROCK_A_TOP = (4000.0, 2300.0, 2500.0)    # lower side of interface 0
ROCK_A_BOT = (4400.0, 2500.0, 2550.0)    # upper side of interface 1: layer 0 is a gradient
ROCK_B = (5000.0, 2900.0, 2700.0)        # layer 1, constant
ROCK_C = (6000.0, 3400.0, 2900.0)        # below interface 2


def i3d_plane(x, y):
    """Interface 1: a tilted plane (bilinear interpolation reproduces it)."""
    return -600.0 + 0.05 * x - 0.08 * y


def write_interface3d(d, z2=-900.0, sampling=None, thresholds=None):
    """An interface3D media file (three interfaces: flat at +400 m with air
    above, a tilted plane, flat at -900 m) and a conf that points to it."""
    p = I3D
    X = p["x0"] + np.arange(p["nx"]) * p["dx"]
    Y = p["y0"] + np.arange(p["ny"]) * p["dy"]
    XX, YY = np.meshgrid(X, Y, indexing="ij")
    zs = [np.full_like(XX, 400.0, dtype=float), i3d_plane(XX, YY), np.full_like(XX, z2, dtype=float)]
    sides = [((0.0, 0.0, 0.0), ROCK_A_TOP), (ROCK_A_BOT, ROCK_B), (ROCK_B, ROCK_C)]
    out = []
    for n in range(3):
        for j in range(p["ny"]):
            out.append(" ".join(f"{zs[n][i, j]:.6f}" for i in range(p["nx"])))
    for q in range(3):  # Vp, Vs, Dc
        for n in range(3):
            for m in range(2):
                for j in range(p["ny"]):
                    out.append(" ".join(f"{sides[n][m][q]:.1f}" for _ in range(p["nx"])))
    (d / "media.int3d").write_text("\n".join(out) + "\n")
    conf = d / "SeisFD3D.conf"
    txt = conf.read_text().replace("media_type = volume", "media_type = interface3D").replace(
        "media_filename = vol.nc", "media_filename = media.int3d")
    txt += f"interface3dpar = {p['nx']} {p['ny']} {p['x0']} {p['y0']} {p['dx']} {p['dy']} 3\n"
    if sampling is not None:
        txt = txt.replace("sampling_point_per_cell = 0 0 0", "sampling_point_per_cell = " + " ".join(map(str, sampling)))
    if thresholds is not None:
        txt = txt.replace("threshold_velocity_jump_of_interface = 100",
                          f"threshold_velocity_jump_of_interface = {thresholds[0]}")
        txt = txt.replace("threshold_percent_change_of_interface = 0.1",
                          f"threshold_percent_change_of_interface = {thresholds[1]}")
    conf.write_text(txt)
    return conf
