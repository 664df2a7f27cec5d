/* cgfd_nc4.c -- reader of netCDF-4 (HDF5-based) input files: the variables
 * and global attributes the front end takes from focal-process and volume
 * media files (the reference reads them with netCDF-C's nc_get_var_float
 * and nc_get_att_float, seisnc.cpp:1148-1228).  A netCDF-4 variable is an
 * HDF5 dataset of the same name in the root group, a dimension is a
 * dataset of the dimension's name whose extent is the dimension's length,
 * and a global attribute is an attribute of the root group.  Values are
 * converted to float (or double) by the HDF5 library, as netCDF-C converts
 * them.  Built against libhdf5 where that library is installed. */
#include "cgfd_nc4.h"

#include <hdf5.h>
#include <string.h>

static hid_t open_ro(const char *path)
{
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL); /* errors come back as status codes */
    if (H5Fis_hdf5(path) <= 0) return -1;
    return H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
}

int cgfd_nc4_probe(const char *path)
{
    if (!path) return -1;
    /* the signature alone is not enough: a damaged file fails to open */
    hid_t f = open_ro(path);
    if (f < 0) return 0;
    H5Fclose(f);
    return 1;
}

int cgfd_nc4_var_dims(const char *path, const char *name, int *rank, long long *dims)
{
    if (!path || !name || !rank || !dims) return -1;
    hid_t f = open_ro(path);
    if (f < 0) return -2;
    int rc = -3;
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d >= 0) {
        hid_t sp = H5Dget_space(d);
        int r = H5Sget_simple_extent_ndims(sp);
        if (r >= 0 && r <= CGFD_NC4_MAX_RANK) {
            hsize_t e[CGFD_NC4_MAX_RANK];
            H5Sget_simple_extent_dims(sp, e, NULL);
            *rank = r;
            for (int i = 0; i < r; i++) dims[i] = (long long)e[i];
            rc = 0;
        }
        H5Sclose(sp);
        H5Dclose(d);
    }
    H5Fclose(f);
    return rc;
}

static int read_var(const char *path, const char *name, hid_t memtype, void *out, long long n)
{
    if (!path || !name || !out || n < 0) return -1;
    hid_t f = open_ro(path);
    if (f < 0) return -2;
    int rc = -3;
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d >= 0) {
        hid_t sp = H5Dget_space(d);
        hssize_t np = H5Sget_simple_extent_npoints(sp);
        if (np != (hssize_t)n) rc = -4;
        else if (H5Dread(d, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0) rc = 0;
        H5Sclose(sp);
        H5Dclose(d);
    }
    H5Fclose(f);
    return rc;
}

int cgfd_nc4_get_var_float(const char *path, const char *name, float *out, long long n)
{
    return read_var(path, name, H5T_NATIVE_FLOAT, out, n);
}

int cgfd_nc4_get_var_double(const char *path, const char *name, double *out, long long n)
{
    return read_var(path, name, H5T_NATIVE_DOUBLE, out, n);
}

int cgfd_nc4_get_att_float(const char *path, const char *name, float *out)
{
    if (!path || !name || !out) return -1;
    hid_t f = open_ro(path);
    if (f < 0) return -2;
    int rc = -3;
    if (H5Aexists(f, name) > 0) {
        hid_t a = H5Aopen(f, name, H5P_DEFAULT);
        hid_t sp = H5Aget_space(a);
        hssize_t np = H5Sget_simple_extent_npoints(sp);
        if (np >= 1) {
            float buf[64];
            if (np <= 64 && H5Aread(a, H5T_NATIVE_FLOAT, buf) >= 0) {
                *out = buf[0];
                rc = 0;
            }
        }
        H5Sclose(sp);
        H5Aclose(a);
    }
    H5Fclose(f);
    return rc;
}
