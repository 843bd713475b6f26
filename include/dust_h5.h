/*
 * dust_h5.h -- C ABI of libdust_h5.so: the HDF5 result and restart files of
 * DUST in the reference's layout, for the caller that keeps the particle
 * pool and the panel wake on the device and writes or reads them at an
 * output step.
 *
 * Replaces the HDF5 helpers the reference calls from save_status
 * (src/io/mod_dust_io.f90:127-437) and load_wake
 * (src/aeroel/mod_wake.f90:597-778): new_hdf5_file / open_hdf5_file,
 * new_hdf5_group, write_hdf5 / write_hdf5_attr, read_hdf5(_al) and
 * check_dset_hdf5 of src/io/mod_hdf5_io.f90.  Datasets are IEEE F64LE or
 * STD I32LE as the reference writes them, scalars in a scalar dataspace,
 * strings fixed-length and space-padded.  Dimensions are given in the file's
 * (C) order: a Fortran array x(3,n) is the dataset [n][3].  Dataset names are
 * paths ("ParticleWake/WakePoints"); missing groups are created.
 *
 * Status: 0 on success, DUST_H5_E_ARG for a bad argument, a missing dataset
 * or a size mismatch, DUST_H5_E_IO for an HDF5 failure;
 * dust_h5_last_error() gives the text.  Host memory only; no GPU calls.
 */
#ifndef DUST_H5_H
#define DUST_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DUST_H5_E_ARG (-1)
#define DUST_H5_E_IO (-7)

#define DUST_H5_F64 0   /* double, H5T_IEEE_F64LE in the file          */
#define DUST_H5_I32 1   /* int32, H5T_STD_I32LE in the file            */
#define DUST_H5_STR 2   /* fixed-length string (dust_h5_info only)     */
#define DUST_H5_MAX_RANK 8

const char *dust_h5_last_error(void);

/* new_hdf5_file (truncates) / open_hdf5_file (read-only) / close_hdf5_file */
int dust_h5_create(const char *path, int64_t *file);
int dust_h5_open(const char *path, int64_t *file);
int dust_h5_close(int64_t file);

/* write_hdf5: a dataset of `rank` dimensions (0 = scalar). */
int dust_h5_write(int64_t file, const char *name, int type, int rank, const int64_t *dims,
                  const void *data);
/* write_hdf5 of a character string (CompName, RefTag, Tag). */
int dust_h5_write_string(int64_t file, const char *name, const char *s);
/* write_hdf5_attr on the root group (sim_param%save_param); len <= 0 means
 * strlen(s), else the string is space-padded to len (the reference's
 * character(len=max_char_len) attributes). */
int dust_h5_write_attr(int64_t file, const char *name, int type, int rank, const int64_t *dims,
                       const void *data);
int dust_h5_write_attr_string(int64_t file, const char *name, const char *s, int64_t len);

/* check_dset_hdf5: *exists = 1 if the path exists. */
int dust_h5_exists(int64_t file, const char *name, int *exists);
/* Element type (DUST_H5_*), rank and dimensions (dims[DUST_H5_MAX_RANK]) of
 * a dataset; a string reports rank 1 and its length. */
int dust_h5_info(int64_t file, const char *name, int *type, int *rank, int64_t *dims);
/* read_hdf5: the whole dataset into `data` as `type`; count must equal the
 * number of elements. */
int dust_h5_read(int64_t file, const char *name, int type, void *data, int64_t count);
/* A string dataset without its space padding, NUL-terminated. */
int dust_h5_read_string(int64_t file, const char *name, char *out, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* DUST_H5_H */
