/* cgfd_nc4.h -- netCDF-4 (HDF5) input files of the front end.
 *
 * The reference reads its focal-process and volume-media files with
 * netCDF-C (seisnc.cpp:1148-1228: nc_open, nc_inq_dimlen, nc_get_var_float,
 * nc_get_att_float); the example's focal file EXAMPLE/hillmodel/stf-GPU.nc is
 * netCDF-4.  These entry points read the same values from such a file
 * through libhdf5.  All return 0 on success and a negative code otherwise
 * (-1 bad argument, -2 not an HDF5/netCDF-4 file or unreadable, -3 no such
 * variable or attribute, -4 size mismatch). */
#ifndef CGFD_NC4_H
#define CGFD_NC4_H
#ifdef __cplusplus
extern "C" {
#endif

#define CGFD_NC4_MAX_RANK 8

/* 1 if path is an HDF5 file (a netCDF-4 file), 0 otherwise. */
int cgfd_nc4_probe(const char *path);
/* nc_inq_var dims of variable name (for a dimension's own dataset, its
 * length): rank and up to CGFD_NC4_MAX_RANK extents. */
int cgfd_nc4_var_dims(const char *path, const char *name, int *rank, long long *dims);
/* nc_get_var_float / nc_get_var_double of all n values of variable name. */
int cgfd_nc4_get_var_float(const char *path, const char *name, float *out, long long n);
int cgfd_nc4_get_var_double(const char *path, const char *name, double *out, long long n);
/* nc_get_att_float of global attribute name (its first value). */
int cgfd_nc4_get_att_float(const char *path, const char *name, float *out);

#ifdef __cplusplus
}
#endif
#endif
