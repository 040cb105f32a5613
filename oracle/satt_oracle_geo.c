/*
 * CPU restatement of the geographic smoothing, gradient and topographic
 * metric routines on either side of the solver path ("the oracle").
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, loaded by tests/
 * and bench.py's cpu_baseline leg, never by the product.
 *
 * Follows the reference's statement and evaluation order (Fortran
 * left-to-right, no FMA contraction; built with -ffp-contract=off), with the
 * constants of src/shared/shared_par.f90:11-15:
 *   pi      = 3.14159265359, a default-real literal widened to double
 *             (3.1415927410125732);
 *   deg2rad = pi/180.0d0, km2deg = 1/(6371.0d0*pi/180.0d0), radius 6371.0d0.
 * The Fortran intrinsics cosd/atand are cos(x*pi/180)/atan(x)*180/pi with
 * the true pi.  sum() over a window is accumulated in array element order
 * (x fastest), one accumulator, as gfortran does.
 *
 * Layout: a field f(nx,ny) is f[i + nx*j]; period-major arrays
 * (nperiod,nx,ny) are v[p + nperiod*(i + nx*j)].  Indices are 0-based.
 */
#include <math.h>
#include <stdlib.h>

#include "satt_oracle.h"

static const double GEO_PI = (double)3.14159265359f;
static const double GEO_RADIUS = 6371.0;
#define GEO_DEG2RAD (GEO_PI / 180.0)
#define GEO_KM2DEG (1.0 / (6371.0 * GEO_PI / 180.0))

static const double TRUE_PI = 3.14159265358979323846;
static double cosd_(double x) { return cos(x * (TRUE_PI / 180.0)); }
static double atand_(double x) { return atan(x) * (180.0 / TRUE_PI); }

/* gps2dist (src/shared/utils.f90:1250-1294): haversine distance in km */
double orc_gps2dist(double lat0, double lon0, double lat1, double lon1) {
  const double dlat = (lat0 - lat1) * GEO_DEG2RAD;
  const double dlon = (lon1 - lon0) * GEO_DEG2RAD;
  const double s1 = sin(dlat * 0.5), s2 = sin(dlon * 0.5);
  const double a = s1 * s1 + s2 * s2 * cos(lat0 * GEO_DEG2RAD) * cos(lat1 * GEO_DEG2RAD);
  const double deg = 2.0 * atan2(sqrt(a), sqrt(1.0 - a));
  return GEO_RADIUS * deg;
}

/* Window half-widths of gaussian_smooth_geo_2 (src/shared/utils.f90:1204-
 * 1213): nx = sigma3/dx truncated, and per row ny = sigma3/cosd(lat)/dy
 * truncated (the row's latitude scales the y half-width). */
int orc_smooth_halfwidth_x(const double *lons, double sigma) {
  return (int)(sigma * 3 / (lons[1] - lons[0]));
}
int orc_smooth_halfwidth_y(const double *lats, int j, double sigma) {
  return (int)(sigma * 3 / cosd_(lats[j]) / (lats[1] - lats[0]));
}

/* gaussian_smooth_geo_2 (src/shared/utils.f90:1194-1223) on rows j0..j1-1
 * of one field: out = sum(w*data)/sum(w) over the window, w =
 * exp(-delta**2/(2 sigma**2)), delta the great-circle distance in degrees
 * (gps2dist*km2deg) from the cell to each window node. */
int orc_smooth_geo2_rows(const double *lons, const double *lats, int nx, int ny,
                         const double *data, double sigma, int j0, int j1, double *out) {
  if (nx < 2 || ny < 2 || j0 < 0 || j1 > ny || j0 > j1 || !(sigma > 0)) return ORC_E_ARG;
  const double sigma_sq = sigma * sigma;
  const int hx = orc_smooth_halfwidth_x(lons, sigma);
  for (int j = j0; j < j1; ++j) {
    const int hy = orc_smooth_halfwidth_y(lats, j, sigma);
    for (int i = 0; i < nx; ++i) {
      const int m1 = i - hx > 0 ? i - hx : 0, m2 = i + hx < nx - 1 ? i + hx : nx - 1;
      const int n1 = j - hy > 0 ? j - hy : 0, n2 = j + hy < ny - 1 ? j + hy : ny - 1;
      double swd = 0, sw = 0;
      for (int jj = n1; jj <= n2; ++jj)
        for (int ii = m1; ii <= m2; ++ii) {
          const double delta = orc_gps2dist(lats[j], lons[i], lats[jj], lons[ii]) * GEO_KM2DEG;
          const double w = exp(-(delta * delta) / (2.0 * sigma_sq));
          swd += w * data[ii + (size_t)nx * jj];
          sw += w;
        }
      out[i + (size_t)nx * j] = (0.0 + swd) / sw;
    }
  }
  return ORC_OK;
}

int orc_smooth_geo2(const double *lons, const double *lats, int nx, int ny, const double *data,
                    double sigma, double *out) {
  return orc_smooth_geo2_rows(lons, lats, nx, ny, data, sigma, 0, ny, out);
}

/* gradient_2_geo (src/shared/utils.f90:1118-1160): central differences in
 * km on the sphere, one-sided at the edges (the first column divides by the
 * second column's spacing, :1153). */
int orc_gradient_geo2(const double *f, const double *lon, const double *lat, int nx, int ny,
                      double *tx, double *ty) {
  if (nx < 2 || ny < 2) return ORC_E_ARG;
  for (int j = 0; j < ny; ++j) {
    const double cl = cosd_(lat[j]);
    const double dyj = GEO_RADIUS * (j > 0 ? lat[j] - lat[j - 1] : 0.0) * GEO_DEG2RAD;
    const double dy2 = GEO_RADIUS * (lat[1] - lat[0]) * GEO_DEG2RAD;
    for (int i = 0; i < nx; ++i) {
      const size_t k = i + (size_t)nx * j;
      const double dxi = GEO_RADIUS * (i > 0 ? lon[i] - lon[i - 1] : 0.0) * GEO_DEG2RAD * cl;
      const double gfx = i > 0 ? f[k] - f[k - 1] : 0.0;
      const double gbx = i < nx - 1 ? f[k + 1] - f[k] : 0.0;
      const double gfy = j > 0 ? f[k] - f[k - nx] : 0.0;
      const double gby = j < ny - 1 ? f[k + nx] - f[k] : 0.0;
      if (i == 0) {
        const double dx2 = GEO_RADIUS * (lon[1] - lon[0]) * GEO_DEG2RAD * cl;
        tx[k] = gbx / dx2;
      } else if (i == nx - 1) {
        tx[k] = gfx / dxi;
      } else {
        tx[k] = (gfx + gbx) / (2 * dxi);
      }
      if (j == 0)
        ty[k] = gby / dy2;
      else if (j == ny - 1)
        ty[k] = gfy / dyj;
      else
        ty[k] = (gfy + gby) / (2 * dyj);
    }
  }
  return ORC_OK;
}

/* sigma of one period (src/grid.f90:116) */
double orc_topo_sigma(double vel, double period, double wavelen_factor) {
  return vel * period * wavelen_factor * GEO_KM2DEG / (2.0 * GEO_PI);
}

/* The topographic metric of get_topo (src/grid.f90:114-123,137-141) and the
 * dip angle (src/topo.f90:123-130) per period: smooth z with the period's
 * sigma, take its gradient, m11 = (1+fy^2)/(1+fx^2+fy^2), m22 = (1+fx^2)/..,
 * m12 = -fx*fy/..; outputs period-major. */
int orc_topo_metric(const double *lons, const double *lats, int nx, int ny, const double *z,
                    int nperiod, const double *sigma, double *m11, double *m12, double *m22,
                    double *angle) {
  if (nx < 2 || ny < 2 || nperiod < 1) return ORC_E_ARG;
  const size_t N = (size_t)nx * ny;
  double *sm = malloc(3 * N * sizeof(double));
  if (!sm) return ORC_E_NOMEM;
  double *fx = sm + N, *fy = sm + 2 * N;
  int st = ORC_OK;
  for (int p = 0; p < nperiod && st == ORC_OK; ++p) {
    st = orc_smooth_geo2(lons, lats, nx, ny, z, sigma[p], sm);
    if (st != ORC_OK) break;
    orc_gradient_geo2(sm, lons, lats, nx, ny, fx, fy);
    for (size_t k = 0; k < N; ++k) {
      const double gx = fx[k], gy = fy[k];
      const double den = 1 + gx * gx + gy * gy;
      const size_t o = p + (size_t)nperiod * k;
      m11[o] = (1 + gy * gy) / den;
      m22[o] = (1 + gx * gx) / den;
      m12[o] = -(gx * gy / den);
      if (angle) angle[o] = atand_(sqrt(gx * gx + gy * gy));
    }
  }
  free(sm);
  return st;
}

/* optimize + steepest_descent (src/tomo2d.f90:203-251): each period of the
 * stacked kernel adj_s smoothed with sigma_2d into ker_s, then
 * svel = svel*(1 + updatemax*ker_s/maxval(abs(ker_s))). */
int orc_steepest_descent(const double *lons, const double *lats, int nx, int ny, int nperiod,
                         const double *adj_s, double sigma_2d, double updatemax, double *svel,
                         double *ker_s) {
  if (nx < 2 || ny < 2 || nperiod < 1) return ORC_E_ARG;
  const size_t N = (size_t)nx * ny;
  double *in = malloc(2 * N * sizeof(double));
  if (!in) return ORC_E_NOMEM;
  double *o = in + N;
  int st = ORC_OK;
  for (int p = 0; p < nperiod && st == ORC_OK; ++p) {
    for (size_t k = 0; k < N; ++k) in[k] = adj_s[p + (size_t)nperiod * k];
    st = orc_smooth_geo2(lons, lats, nx, ny, in, sigma_2d, o);
    for (size_t k = 0; k < N; ++k) ker_s[p + (size_t)nperiod * k] = o[k];
  }
  free(in);
  if (st != ORC_OK) return st;
  double mx = 0;
  for (size_t k = 0; k < N * nperiod; ++k) mx = fabs(ker_s[k]) > mx ? fabs(ker_s[k]) : mx;
  for (size_t k = 0; k < N * nperiod; ++k) svel[k] = svel[k] * (1 + updatemax * (ker_s[k] / mx));
  return ORC_OK;
}
