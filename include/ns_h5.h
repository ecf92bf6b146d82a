/* ns_h5 -- the reference's HDF5 files read and written from C (host only).
 *
 * The reference writes with the Fortran HDF5 high-level API
 * (h5ltmake_dataset_double_f, h5ltset_attribute_*_f; big_periodic_3D.f90:
 * 589-712, re_simulation_sub_tstep_interBC_combine.f90:403-504,
 * NS_functions.f90:1633-1664).  A Fortran array of dims (n1,...,nr) is stored
 * as a dataset whose HDF5 (C-order) dims are (nr,...,n1), with the same bytes;
 * the functions here take Fortran-order dims and column-major data, so a
 * Fortran or numpy-'F' array round-trips unchanged and files interoperate
 * with the reference and its MATLAB scripts.  Status: 0 ok, negative error
 * (NSH5_ERR_*); HDF5's own error printing is switched off. */
#ifndef NS_H5_H
#define NS_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define NSH5_OK 0
#define NSH5_ERR_INVALID (-1)
#define NSH5_ERR_IO (-2)     /* HDF5 call failed (missing file/object, exists on create, ...) */
#define NSH5_ERR_SHAPE (-3)  /* dataset rank/dims differ from the expected ones */

typedef long long nsh5_file; /* hid_t of an open file */

/* create a new file; fails if it exists (H5F_ACC_EXCL, as h5fcreate_f(...,H5F_ACC_EXCL_F,...)) */
int nsh5_create(const char *path, nsh5_file *f);
/* open an existing file read-only (readwrite = 0) or read-write */
int nsh5_open(const char *path, int readwrite, nsh5_file *f);
int nsh5_close(nsh5_file f);
/* 1 if the link (group or dataset path, e.g. "t_1.0000/u_sub") exists, 0 if not */
int nsh5_exists(nsh5_file f, const char *path);
/* FP64 dataset at path (intermediate groups created), Fortran dims[rank] */
int nsh5_write_double(nsh5_file f, const char *path, int rank, const long *dims, const double *data);
/* read into data; rank and Fortran dims must match the dataset */
int nsh5_read_double(nsh5_file f, const char *path, int rank, const long *dims, double *data);
/* rank and Fortran-order dims (up to 8) of a dataset */
int nsh5_get_dims(nsh5_file f, const char *path, int *rank, long *dims);
/* scalar attributes on an object ("/" = the root group), h5ltset_attribute_*_f */
int nsh5_set_attr_int(nsh5_file f, const char *obj, const char *name, int v);
int nsh5_set_attr_double(nsh5_file f, const char *obj, const char *name, double v);
int nsh5_set_attr_string(nsh5_file f, const char *obj, const char *name, const char *v);
int nsh5_get_attr_int(nsh5_file f, const char *obj, const char *name, int *v);
int nsh5_get_attr_double(nsh5_file f, const char *obj, const char *name, double *v);
/* string attribute into buf[cap] (NUL-terminated, trailing blanks/NULs removed) */
int nsh5_get_attr_string(nsh5_file f, const char *obj, const char *name, char *buf, long cap);
/* names of the members of a group, '\n'-separated, into buf[cap] */
int nsh5_list(nsh5_file f, const char *group, char *buf, long cap);

#ifdef __cplusplus
}
#endif
#endif
