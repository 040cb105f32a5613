/*
 * satt_h5.h -- the reference's HDF5 model and kernel files, read and written
 * by libsatt_h5.so (host code over the HDF5 C library, 1.10).
 *
 * The reference writes Fortran arrays through the Fortran HDF5 API
 * (src/shared/hdf5_interface.f90:362-417, h5ltmake_dataset_f), so a Fortran
 * array of shape (n1, ..., nk) is a dataset of C dimensions (nk, ..., n1)
 * holding the array's elements in memory order, as IEEE 64-bit floats.
 * Arrays here are passed in that same Fortran order, with `fdims` the
 * Fortran shape: satt_h5_add(f, "/x", 2, {nx, ny}, x) writes what the
 * reference's hf%add('/x', x) writes for real(dp) x(nx, ny).  Parent groups
 * of a path are created as the reference's hdf_add_group does
 * (:220-250).  Every entry point returns SATT_H5_OK (0) or a negative
 * status: SATT_H5_E_ARG for bad arguments, a missing dataset or a shape that
 * does not match, SATT_H5_E_H5 when the HDF5 library reports an error.
 */
#ifndef SATT_H5_H
#define SATT_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define SATT_H5_OK 0
#define SATT_H5_E_ARG (-1)
#define SATT_H5_E_H5 (-9)
#define SATT_H5_MAX_RANK 4

typedef struct satt_h5 satt_h5;

/* hdf5_file%open (src/shared/hdf5_interface.f90:100-151): status "new" or
 * "replace" creates (truncating an existing file); "old" with action "read"
 * opens read-only, with "write"/"rw" opens for writing (creating the file
 * when it does not exist). */
int satt_h5_open(satt_h5 **file, const char *path, const char *status, const char *action);
/* hdf5_file%close (:153-175). */
int satt_h5_close(satt_h5 *file);
/* hdf5_file%exist (:199-208): 1 if `name` exists, 0 if not. */
int satt_h5_exist(satt_h5 *file, const char *name);
/* hdf5_file%add for real(dp) arrays of rank 1-4 (:362-417). */
int satt_h5_add(satt_h5 *file, const char *name, int rank, const long long *fdims,
                const double *data);
/* hdf5_file%add of transpose_k(data) for a rank-k array data (k = 2..4,
 * src/shared/utils.f90:820-850: the axes reversed), the form in which the
 * reference stores every model and kernel array. */
int satt_h5_add_transposed(satt_h5 *file, const char *name, int rank, const long long *fdims,
                           const double *data);
/* The Fortran shape of a dataset (hdf5_file%get's allocation, :545-620). */
int satt_h5_shape(satt_h5 *file, const char *name, int *rank, long long *fdims);
/* hdf5_file%get for real(dp) arrays: the dataset's elements in memory order
 * into data (room for the product of its shape). */
int satt_h5_get(satt_h5 *file, const char *name, double *data);

/* ---- the 2-D inversion's files (src/acqui_2d.f90:210-264) ----
 * svel is (nperiod, nx, ny), the att_grid layout; the dataset holds
 * transpose_3(svel) (src/shared/utils.f90:820-833), of C dimensions
 * (nperiod, nx, ny). */

/* att_acqui_2d%write_model (:228-245) with stations, or
 * att_acqui_2d%write_target_model (:247-264) when stlo and stla are NULL:
 * a new file with /stlo, /stla (nsta), /period (nperiod), /lon (nx),
 * /lat (ny) and /vel. */
int satt_h5_write_model_2d(const char *path, const double *stlo, const double *stla, int nsta,
                           const double *periods, int nperiod, const double *lon, int nx,
                           const double *lat, int ny, const double *svel);
/* att_acqui_2d%write_iter (:210-226) into an open model_iter file: at
 * iter 0 first /stlo, /stla, /period, /lon, /lat, then
 * '/vel' // gr_name // '_' // iter as i3.3 (e.g. /velph_000). */
int satt_h5_write_iter_2d(satt_h5 *file, int iter, const char *gr_name, const double *stlo,
                          const double *stla, int nsta, const double *periods, int nperiod,
                          const double *lon, int nx, const double *lat, int ny,
                          const double *svel);

/* ---- the 3-D inversion's model file (src/tomo.f90:72-98, 350-385) ----
 * Model arrays are (nx, ny, nz); the datasets hold transpose_3 of them, of
 * C dimensions (nx, ny, nz). */

/* att_tomo%initialize_inv's first records (:79-94): /lon, /lat, /dep and
 * /vs_000 (with /vp_000, /rho_000 when vp3d and rho3d are given). */
int satt_h5_write_model_3d_init(satt_h5 *file, const double *lon, int nx, const double *lat,
                                int ny, const double *dep, int nz, const double *vs3d,
                                const double *vp3d, const double *rho3d);
/* write_tmp_model (:350-365): /vs_iii (and /vp_iii, /rho_iii). */
int satt_h5_write_model_3d_iter(satt_h5 *file, int iter, int nx, int ny, int nz,
                                const double *vs3d, const double *vp3d, const double *rho3d);
/* write_gradient (:367-383): /gradient_iii of transpose_4(gradient_s) with
 * gradient_s (nker, nx, ny, nz). */
int satt_h5_write_gradient_3d(satt_h5 *file, int iter, int nker, int nx, int ny, int nz,
                              const double *gradient_s);
/* write_gradient's density kernels (:376-381): '/kdensity_' // gr_name //
 * '_' // iii of transpose_4(ker_density), ker_density (nker, nx, ny, nz). */
int satt_h5_write_kdensity_3d(satt_h5 *file, int iter, const char *gr_name, int nker, int nx,
                              int ny, int nz, const double *ker_density);
/* The initial model of init_model_type 2 (src/model.f90:95-105): dataset
 * `name` (the reference reads '/vs') of a file as transpose_3 of it,
 * vs3d (nx, ny, nz); a dataset of another shape is SATT_H5_E_ARG (the
 * reference stops). */
int satt_h5_read_model_3d(const char *path, const char *name, int nx, int ny, int nz,
                          double *vs3d);

#ifdef __cplusplus
}
#endif
#endif /* SATT_H5_H */
