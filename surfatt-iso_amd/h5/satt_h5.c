/*
 * The reference's HDF5 model and kernel files (include/satt_h5.h), host code
 * over the HDF5 C library.  Datasets are written with the C dimensions
 * reversed from the Fortran shape and the data in memory order, which is
 * what the reference's Fortran writers (src/shared/hdf5_interface.f90,
 * h5ltmake_dataset_f) put in the file; transpose_k (src/shared/utils.f90:
 * 820-850) reverses the axes of a model or kernel array before it is
 * written.
 */
#include "../../include/satt_h5.h"

#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct satt_h5 {
  hid_t fid;
};

static int quiet_done = 0;

/* The reference ignores the library's error stack; so do we (errors come
 * back as status codes). */
static void quiet(void) {
  if (!quiet_done) {
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    quiet_done = 1;
  }
}

int satt_h5_open(satt_h5 **file, const char *path, const char *status, const char *action) {
  quiet();
  if (!file || !path) return SATT_H5_E_ARG;
  *file = NULL;
  const char *st = status ? status : "old";
  const char *ac = action ? action : "rw";
  hid_t fid = -1;
  if (!strcmp(st, "new") || !strcmp(st, "replace")) {
    fid = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  } else if (!strcmp(st, "old")) {
    if (!strcmp(ac, "read") || !strcmp(ac, "r")) {
      fid = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    } else if (!strcmp(ac, "write") || !strcmp(ac, "readwrite") || !strcmp(ac, "w") ||
               !strcmp(ac, "rw")) {
      FILE *probe = fopen(path, "rb");
      if (probe) {
        fclose(probe);
        fid = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
      } else {
        fid = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
      }
    } else {
      return SATT_H5_E_ARG;
    }
  } else {
    return SATT_H5_E_ARG;
  }
  if (fid < 0) return SATT_H5_E_H5;
  satt_h5 *f = (satt_h5 *)malloc(sizeof *f);
  if (!f) {
    H5Fclose(fid);
    return SATT_H5_E_H5;
  }
  f->fid = fid;
  *file = f;
  return SATT_H5_OK;
}

int satt_h5_close(satt_h5 *file) {
  if (!file) return SATT_H5_E_ARG;
  const herr_t e = H5Fclose(file->fid);
  free(file);
  return e < 0 ? SATT_H5_E_H5 : SATT_H5_OK;
}

int satt_h5_exist(satt_h5 *file, const char *name) {
  if (!file || !name) return SATT_H5_E_ARG;
  const htri_t e = H5Lexists(file->fid, name, H5P_DEFAULT);
  return e > 0 ? 1 : 0;
}

/* hdf_add_group: every parent group of the path, created when missing. */
static int add_groups(hid_t fid, const char *name) {
  const size_t n = strlen(name);
  char *buf = (char *)malloc(n + 1);
  if (!buf) return SATT_H5_E_H5;
  for (size_t k = 1; k < n; ++k) {
    if (name[k] != '/') continue;
    memcpy(buf, name, k);
    buf[k] = '\0';
    if (H5Lexists(fid, buf, H5P_DEFAULT) <= 0) {
      const hid_t g = H5Gcreate2(fid, buf, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
      if (g < 0) {
        free(buf);
        return SATT_H5_E_H5;
      }
      H5Gclose(g);
    }
  }
  free(buf);
  return SATT_H5_OK;
}

int satt_h5_add(satt_h5 *file, const char *name, int rank, const long long *fdims,
                const double *data) {
  if (!file || !name || !fdims || !data || rank < 1 || rank > SATT_H5_MAX_RANK)
    return SATT_H5_E_ARG;
  hsize_t cd[SATT_H5_MAX_RANK];
  for (int k = 0; k < rank; ++k) {
    if (fdims[k] < 1) return SATT_H5_E_ARG;
    cd[k] = (hsize_t)fdims[rank - 1 - k];
  }
  if (H5Lexists(file->fid, name, H5P_DEFAULT) > 0) return SATT_H5_E_ARG;
  const int g = add_groups(file->fid, name);
  if (g != SATT_H5_OK) return g;
  const hid_t sp = H5Screate_simple(rank, cd, NULL);
  if (sp < 0) return SATT_H5_E_H5;
  const hid_t ds = H5Dcreate2(file->fid, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT,
                              H5P_DEFAULT);
  int st = SATT_H5_E_H5;
  if (ds >= 0) {
    if (H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0)
      st = SATT_H5_OK;
    H5Dclose(ds);
  }
  H5Sclose(sp);
  return st;
}

/* transpose_k of a Fortran array of shape fdims: y(i_k, ..., i_1) =
 * x(i_1, ..., i_k), returned with its shape in ydims. */
static double *transpose_k(int rank, const long long *fdims, const double *x, long long *ydims) {
  size_t n = 1;
  for (int k = 0; k < rank; ++k) {
    n *= (size_t)fdims[k];
    ydims[k] = fdims[rank - 1 - k];
  }
  double *y = (double *)malloc((n ? n : 1) * sizeof(double));
  if (!y) return NULL;
  long long idx[SATT_H5_MAX_RANK] = {0, 0, 0, 0};
  for (size_t e = 0; e < n; ++e) {
    /* e runs over x in memory order: idx is x's multi-index (first fastest) */
    size_t ye = 0;
    for (int k = 0; k < rank; ++k) ye = ye * (size_t)fdims[k] + (size_t)idx[k];
    y[ye] = x[e];
    for (int k = 0; k < rank; ++k) {
      if (++idx[k] < fdims[k]) break;
      idx[k] = 0;
    }
  }
  return y;
}

int satt_h5_add_transposed(satt_h5 *file, const char *name, int rank, const long long *fdims,
                           const double *data) {
  if (!file || !name || !fdims || !data || rank < 2 || rank > SATT_H5_MAX_RANK)
    return SATT_H5_E_ARG;
  for (int k = 0; k < rank; ++k)
    if (fdims[k] < 1) return SATT_H5_E_ARG;
  long long yd[SATT_H5_MAX_RANK];
  double *y = transpose_k(rank, fdims, data, yd);
  if (!y) return SATT_H5_E_H5;
  const int st = satt_h5_add(file, name, rank, yd, y);
  free(y);
  return st;
}

int satt_h5_shape(satt_h5 *file, const char *name, int *rank, long long *fdims) {
  if (!file || !name || !rank || !fdims) return SATT_H5_E_ARG;
  if (H5Lexists(file->fid, name, H5P_DEFAULT) <= 0) return SATT_H5_E_ARG;
  const hid_t ds = H5Dopen2(file->fid, name, H5P_DEFAULT);
  if (ds < 0) return SATT_H5_E_ARG;
  const hid_t sp = H5Dget_space(ds);
  const int r = sp < 0 ? -1 : H5Sget_simple_extent_ndims(sp);
  int st = SATT_H5_E_H5;
  if (r >= 1 && r <= SATT_H5_MAX_RANK) {
    hsize_t cd[SATT_H5_MAX_RANK];
    H5Sget_simple_extent_dims(sp, cd, NULL);
    *rank = r;
    for (int k = 0; k < r; ++k) fdims[k] = (long long)cd[r - 1 - k];
    st = SATT_H5_OK;
  } else if (r >= 0) {
    st = SATT_H5_E_ARG;
  }
  if (sp >= 0) H5Sclose(sp);
  H5Dclose(ds);
  return st;
}

int satt_h5_get(satt_h5 *file, const char *name, double *data) {
  if (!file || !name || !data) return SATT_H5_E_ARG;
  if (H5Lexists(file->fid, name, H5P_DEFAULT) <= 0) return SATT_H5_E_ARG;
  const hid_t ds = H5Dopen2(file->fid, name, H5P_DEFAULT);
  if (ds < 0) return SATT_H5_E_ARG;
  const herr_t e = H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  H5Dclose(ds);
  return e < 0 ? SATT_H5_E_H5 : SATT_H5_OK;
}

/* '/stem' // tag // '_' // iter as i3.3 */
static void sec_name(char *out, size_t cap, const char *stem, const char *tag, int iter) {
  snprintf(out, cap, "/%s%s_%03d", stem, tag ? tag : "", iter);
}

/* The records the 2-D files open with: /stlo, /stla, /period, /lon, /lat. */
static int write_axes_2d(satt_h5 *f, const double *stlo, const double *stla, int nsta,
                         const double *periods, int nperiod, const double *lon, int nx,
                         const double *lat, int ny) {
  int st;
  if (stlo && stla) {
    const long long ns = nsta;
    if ((st = satt_h5_add(f, "/stlo", 1, &ns, stlo)) < 0) return st;
    if ((st = satt_h5_add(f, "/stla", 1, &ns, stla)) < 0) return st;
  }
  const long long np = nperiod, lx = nx, ly = ny;
  if ((st = satt_h5_add(f, "/period", 1, &np, periods)) < 0) return st;
  if ((st = satt_h5_add(f, "/lon", 1, &lx, lon)) < 0) return st;
  return satt_h5_add(f, "/lat", 1, &ly, lat);
}

static int bad_2d(const double *periods, int nperiod, const double *lon, int nx, const double *lat,
                  int ny, const double *svel, const double *stlo, const double *stla, int nsta) {
  return !periods || nperiod < 1 || !lon || nx < 1 || !lat || ny < 1 || !svel ||
         ((stlo || stla) && (!stlo || !stla || nsta < 1));
}

int satt_h5_write_model_2d(const char *path, const double *stlo, const double *stla, int nsta,
                           const double *periods, int nperiod, const double *lon, int nx,
                           const double *lat, int ny, const double *svel) {
  if (!path || bad_2d(periods, nperiod, lon, nx, lat, ny, svel, stlo, stla, nsta))
    return SATT_H5_E_ARG;
  satt_h5 *f = NULL;
  int st = satt_h5_open(&f, path, "new", "write");
  if (st < 0) return st;
  st = write_axes_2d(f, stlo, stla, nsta, periods, nperiod, lon, nx, lat, ny);
  if (st == SATT_H5_OK) {
    const long long d[3] = {nperiod, nx, ny};
    st = satt_h5_add_transposed(f, "/vel", 3, d, svel);
  }
  const int sc = satt_h5_close(f);
  return st < 0 ? st : sc;
}

int satt_h5_write_iter_2d(satt_h5 *file, int iter, const char *gr_name, const double *stlo,
                          const double *stla, int nsta, const double *periods, int nperiod,
                          const double *lon, int nx, const double *lat, int ny,
                          const double *svel) {
  if (!file || iter < 0 || iter > 999 || !gr_name ||
      bad_2d(periods, nperiod, lon, nx, lat, ny, svel, stlo, stla, nsta) ||
      (iter == 0 && (!stlo || !stla)))
    return SATT_H5_E_ARG;
  int st;
  if (iter == 0 &&
      (st = write_axes_2d(file, stlo, stla, nsta, periods, nperiod, lon, nx, lat, ny)) < 0)
    return st;
  char name[256];
  sec_name(name, sizeof name, "vel", gr_name, iter);
  const long long d[3] = {nperiod, nx, ny};
  return satt_h5_add_transposed(file, name, 3, d, svel);
}

static int write_model_set(satt_h5 *f, int iter, int nx, int ny, int nz, const double *vs3d,
                           const double *vp3d, const double *rho3d) {
  const long long d[3] = {nx, ny, nz};
  char name[64];
  int st;
  sec_name(name, sizeof name, "vs", NULL, iter);
  if ((st = satt_h5_add_transposed(f, name, 3, d, vs3d)) < 0) return st;
  if (vp3d && rho3d) {
    sec_name(name, sizeof name, "vp", NULL, iter);
    if ((st = satt_h5_add_transposed(f, name, 3, d, vp3d)) < 0) return st;
    sec_name(name, sizeof name, "rho", NULL, iter);
    if ((st = satt_h5_add_transposed(f, name, 3, d, rho3d)) < 0) return st;
  }
  return SATT_H5_OK;
}

int satt_h5_write_model_3d_init(satt_h5 *file, const double *lon, int nx, const double *lat,
                                int ny, const double *dep, int nz, const double *vs3d,
                                const double *vp3d, const double *rho3d) {
  if (!file || !lon || nx < 1 || !lat || ny < 1 || !dep || nz < 1 || !vs3d ||
      (!vp3d) != (!rho3d))
    return SATT_H5_E_ARG;
  const long long lx = nx, ly = ny, lz = nz;
  int st;
  if ((st = satt_h5_add(file, "/lon", 1, &lx, lon)) < 0) return st;
  if ((st = satt_h5_add(file, "/lat", 1, &ly, lat)) < 0) return st;
  if ((st = satt_h5_add(file, "/dep", 1, &lz, dep)) < 0) return st;
  return write_model_set(file, 0, nx, ny, nz, vs3d, vp3d, rho3d);
}

int satt_h5_write_model_3d_iter(satt_h5 *file, int iter, int nx, int ny, int nz,
                                const double *vs3d, const double *vp3d, const double *rho3d) {
  if (!file || iter < 0 || iter > 999 || nx < 1 || ny < 1 || nz < 1 || !vs3d ||
      (!vp3d) != (!rho3d))
    return SATT_H5_E_ARG;
  return write_model_set(file, iter, nx, ny, nz, vs3d, vp3d, rho3d);
}

int satt_h5_write_gradient_3d(satt_h5 *file, int iter, int nker, int nx, int ny, int nz,
                              const double *gradient_s) {
  if (!file || iter < 0 || iter > 999 || nker < 1 || nx < 1 || ny < 1 || nz < 1 || !gradient_s)
    return SATT_H5_E_ARG;
  char name[64];
  sec_name(name, sizeof name, "gradient", NULL, iter);
  const long long d[4] = {nker, nx, ny, nz};
  return satt_h5_add_transposed(file, name, 4, d, gradient_s);
}

int satt_h5_write_kdensity_3d(satt_h5 *file, int iter, const char *gr_name, int nker, int nx,
                              int ny, int nz, const double *ker_density) {
  if (!file || iter < 0 || iter > 999 || !gr_name || nker < 1 || nx < 1 || ny < 1 || nz < 1 ||
      !ker_density)
    return SATT_H5_E_ARG;
  char name[256];
  snprintf(name, sizeof name, "/kdensity_%s_%03d", gr_name, iter);
  const long long d[4] = {nker, nx, ny, nz};
  return satt_h5_add_transposed(file, name, 4, d, ker_density);
}

int satt_h5_read_model_3d(const char *path, const char *name, int nx, int ny, int nz,
                          double *vs3d) {
  if (!path || !name || nx < 1 || ny < 1 || nz < 1 || !vs3d) return SATT_H5_E_ARG;
  satt_h5 *f = NULL;
  int st = satt_h5_open(&f, path, "old", "read");
  if (st < 0) return st;
  int rank = 0;
  long long fd[SATT_H5_MAX_RANK];
  st = satt_h5_shape(f, name, &rank, fd);
  /* vstmp has the dataset's Fortran shape; transpose_3(vstmp) must be
   * (nx, ny, nz) */
  if (st == SATT_H5_OK && (rank != 3 || fd[0] != nz || fd[1] != ny || fd[2] != nx))
    st = SATT_H5_E_ARG;
  double *tmp = NULL;
  if (st == SATT_H5_OK) {
    tmp = (double *)malloc((size_t)nx * ny * nz * sizeof(double));
    if (!tmp) st = SATT_H5_E_H5;
  }
  if (st == SATT_H5_OK) st = satt_h5_get(f, name, tmp);
  if (st == SATT_H5_OK) {
    long long yd[3];
    double *y = transpose_k(3, fd, tmp, yd);
    if (!y) {
      st = SATT_H5_E_H5;
    } else {
      memcpy(vs3d, y, (size_t)nx * ny * nz * sizeof(double));
      free(y);
    }
  }
  free(tmp);
  const int sc = satt_h5_close(f);
  return st < 0 ? st : sc;
}
