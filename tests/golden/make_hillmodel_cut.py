"""Writes tests/golden/hillmodel_cut/: the reference's example model
(EXAMPLE/hillmodel: 201x201x200 Gaussian hill, dh 50 m, hybrid grid with the
conversion depth at -3 km, one focal point, the example's absorbing layers
and stations) cut to the 48 x 48 columns around its focal point, as input
data for the parameter-file front end's tests.

- gridvmap_G.dat: the example's own anchor lines of those columns (text kept
  as is) under a vmap_dims header of 48 48 2;
- SeisFD3D.conf, absorb.dat, source.dat: the example's files with ni/nj,
  seispath and the media and focal file names changed;
- station.dat: the example's receivers and lines, with one snapshot box that
  fits the cut grid;
- stf-classic.nc: the example's netCDF-4 focal file (stf-GPU.nc) rewritten
  in the netCDF classic format (same dimensions, variables and DT), read
  with the front end's netCDF-4 reader;
- media_standin.nc: the example's media file (Vsim_gpu.nc) is not part of
  the reference repository, so a synthetic three-layer volume (Vp 4000 /
  5500 / 6500 m/s, Vs = Vp/sqrt(3), rho 2400 / 2600 / 2800 kg/m^3, as the
  hill benchmark's medium) stands in for it.

Run from the repository root with the reference at /root/reference:
    python tests/golden/make_hillmodel_cut.py
"""
import os
import re
import sys

import numpy as np
from scipy.io import netcdf_file

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "copara-cgfd3d_amd"))
import seisconf  # noqa: E402

SRC = "/root/reference/EXAMPLE/hillmodel"
OUT = os.path.join(HERE, "hillmodel_cut")
I0, J0, N = 76, 76, 48       # columns 76..123 in x and y: x, y = 3800 .. 6150 m


def main():
    os.makedirs(OUT, exist_ok=True)
    with open(f"{SRC}/gridvmap_G.dat") as f:
        lines = f.read().splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.strip().startswith("<vmap_anchor>"))
    head, anchors = lines[:k + 1], [ln for ln in lines[k + 1:] if ln.strip()]
    assert len(anchors) == 201 * 201
    head = [re.sub(r"^vmap_dims\s*=.*", f"vmap_dims = {N} {N} 2", ln) for ln in head]
    cut = [anchors[i * 201 + j] for i in range(I0, I0 + N) for j in range(J0, J0 + N)]
    with open(f"{OUT}/gridvmap_G.dat", "w") as f:
        f.write("\n".join(head + cut) + "\n")

    conf = open(f"{SRC}/SeisFD3D.conf").read()
    conf = re.sub(r"(?m)^ni\s*=.*$", f"ni    = {N}", conf)
    conf = re.sub(r"(?m)^nj\s*=.*$", f"nj    = {N}", conf)
    conf = re.sub(r"(?m)^seispath\s*=.*$", "seispath = .", conf)
    conf = re.sub(r"(?m)^media_filename\s*=.*$", "media_filename = media_standin.nc", conf)
    with open(f"{OUT}/SeisFD3D.conf", "w") as f:
        f.write(conf)
    with open(f"{OUT}/absorb.dat", "w") as f:
        f.write(open(f"{SRC}/absorb.dat").read())
    src = open(f"{SRC}/source.dat").read()
    src = re.sub(r"(?m)^focal_process_NC_file\s*=.*$", "focal_process_NC_file = stf-classic.nc", src)
    with open(f"{OUT}/source.dat", "w") as f:
        f.write(src)
    st = open(f"{SRC}/station.dat").read()
    st = re.sub(r"(?m)^number_of_snap\s*=.*$", "number_of_snap = 1", st)
    st = re.sub(r"(?m)^snap_001\s*=.*$", f"snap_001 =  1   1   100    {N} {N} 1       1 1 1   20     3", st)
    with open(f"{OUT}/station.dat", "w") as f:
        f.write(st)

    # the focal file, through the netCDF-4 reader, in the classic format
    nc4 = seisconf._read_netcdf(f"{SRC}/stf-GPU.nc", "focal process file")
    npnt, fnt = nc4.dimensions["NP"], nc4.dimensions["NT"]
    with netcdf_file(f"{OUT}/stf-classic.nc", "w", version=2) as f:
        f.createDimension("NP", npnt)
        f.createDimension("NT", fnt)
        f.DT = np.float32(nc4.DT)
        for nm in ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz"):
            v = f.createVariable(nm, "f4", ("NP", "NT"))
            v[:] = nc4.variables[nm][:]
        for nm in ("posx", "posy", "posz"):
            v = f.createVariable(nm, "f4", ("NP",))
            v[:] = nc4.variables[nm][:]

    # stand-in media on the cut's footprint (coordinate variables x, y, depth)
    xs = np.arange(3500.0, 6600.0, 250.0)
    ys = np.arange(3500.0, 6600.0, 250.0)
    zs = np.arange(-10500.0, 2001.0, 500.0)
    Z = np.broadcast_to(zs[None, None, :], (xs.size, ys.size, zs.size))
    vp = np.where(Z > -3000.0, 4000.0, np.where(Z > -8000.0, 5500.0, 6500.0))
    rho = np.where(Z > -3000.0, 2400.0, np.where(Z > -8000.0, 2600.0, 2800.0))
    with netcdf_file(f"{OUT}/media_standin.nc", "w", version=2) as f:
        for nm, a in (("x", xs), ("y", ys), ("depth", zs)):
            f.createDimension(nm, a.size)
            v = f.createVariable(nm, "f4", (nm,))
            v[:] = a.astype(np.float32)
        for nm, a in (("Vp", vp), ("Vs", vp / np.sqrt(3.0)), ("rho", rho)):
            v = f.createVariable(nm, "f4", ("x", "y", "depth"))
            v[:] = a.astype(np.float32)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
