#!/bin/bash
# Builds libcgfd3d.so (gfx950) in-tree.  Kernels are compiled without FP
# contraction (bit-exact arithmetic contract), with IEEE denormals, and with
# packed FP32 instructions disabled.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
root="$(cd "$here/.." && pwd)"
ARCH="${CGFD_ARCH:-gfx950}"
FLAGS=(-O3 -std=c++17 -fPIC --offload-arch="$ARCH" -ffp-contract=off -fno-gpu-flush-denormals-to-zero
       -fhip-fp32-correctly-rounded-divide-sqrt
       -Xclang -target-feature -Xclang -packed-fp32-ops
       -I"$root/include" -I"$here/csrc")
mkdir -p "$here/_build"
hipcc "${FLAGS[@]}" ${CGFD_EXTRA_FLAGS:-} -c "$here/csrc/cgfd_kernels.hip" -o "$here/_build/cgfd_kernels.o"
hipcc "${FLAGS[@]}" -c "$here/csrc/cgfd_capi.hip" -o "$here/_build/cgfd_capi.o"
g++ -O2 -std=c++17 -fPIC -ffp-contract=off -Wall -I"$root/include" -c "$here/csrc/cgfd_ncio.cpp" -o "$here/_build/cgfd_ncio.o"
hipcc -shared -fPIC --offload-arch="$ARCH" -o "$here/libcgfd3d.so.tmp" "$here/_build/cgfd_kernels.o" "$here/_build/cgfd_capi.o" \
      "$here/_build/cgfd_ncio.o"
mv -f "$here/libcgfd3d.so.tmp" "$here/libcgfd3d.so"
echo "built $here/libcgfd3d.so"
# the surrogate's BF16 MFMA kernels
hipcc "${FLAGS[@]}" -shared "$here/csrc/fno_kernels.hip" -o "$here/libcgfd_fno.so.tmp"
mv -f "$here/libcgfd_fno.so.tmp" "$here/libcgfd_fno.so"
echo "built $here/libcgfd_fno.so"
# the netCDF-4 (HDF5) input reader of the front end, where libhdf5 is installed
HDF5_ROOT="${CGFD_HDF5_ROOT:-/opt/conda}"
if [ -f "$HDF5_ROOT/include/hdf5.h" ]; then
    gcc -O2 -fPIC -shared -Wall -I"$root/include" -I"$HDF5_ROOT/include" "$here/csrc/cgfd_nc4.c" \
        -L"$HDF5_ROOT/lib" -lhdf5 -Wl,-rpath,"$HDF5_ROOT/lib" -o "$here/libcgfd_nc4.so.tmp"
    mv -f "$here/libcgfd_nc4.so.tmp" "$here/libcgfd_nc4.so"
    echo "built $here/libcgfd_nc4.so"
else
    echo "no hdf5.h under $HDF5_ROOT: libcgfd_nc4.so (netCDF-4 inputs) not built"
fi
# the C++ host driver (one child-rank run through the C-ABI)
g++ -O2 -std=c++17 -Wall -ffp-contract=off -I"$root/include" "$here/host/cgfd_child.cpp" "$here/host/cgfd_conf.cpp" \
    -L"$here" -lcgfd3d -ldl -Wl,-rpath,'$ORIGIN' -o "$here/cgfd_child"
echo "built $here/cgfd_child"
