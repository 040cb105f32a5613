/*
 * CPU restatement of SurfATT's forward + adjoint hot path ("the oracle").
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 * Build: gcc -O2 -fPIC -ffp-contract=off (oracle/Makefile).  With
 * FLT_EVAL_METHOD == 0 (x86-64 SSE) float expressions are evaluated in
 * float and double expressions in double, which is what the reference's
 * implicitly typed REAL(4) temporaries need.
 *
 * Cited line numbers are in the reference tree.
 */
#include "satt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* src/eikon_solver_2d.f90:785,1297: the literal has no kind suffix, so it is
 * rounded to REAL(4) and then widened. */
static const double PI_F = (double)3.14159265358979323846264338327950288f;
/* src/eikon_solver_2d.f90:776,1296 */
static const double R_EARTH = 6371.0;
/* src/eikon_solver_2d.f90:775: tol=(10.0)**(-4), folded in REAL(4) */
static const double FWD_TOL = (double)1.0e-4f;
/* src/eikon_solver_2d.f90:1294: tol=(10.0)**(-9), eps=(10.0)**(-14) */
static const double ADJ_TOL = (double)1.0e-9f;
static const double ADJ_EPS = (double)1.0e-14f;
static const int FWD_MAXITER = 2000; /* :774 */
static const int ADJ_MAXITER = 1000; /* :1293 */

#define IDX(i, j) ((size_t)(i) + (size_t)nx * (size_t)(j))

/* xx = xx_deg/180.0*pi  (:801-804, :1308-1317) */
static double deg2rad_ref(double d) { return d / 180.0 * PI_F; }

/* a, b, c from the metric (:817-821, :1338-1344).  R_earth**2 is folded to
 * 6371*6371; cos(y)**2 is cos(y)*cos(y). */
static void metric_abc(int nx, int ny, const double *yy, const double *spha,
                       const double *sphb, const double *sphc, double *a,
                       double *b, double *c) {
  const double r2 = R_EARTH * R_EARTH;
  for (int j = 0; j < ny; ++j) {
    const double cy = cos(yy[j]);
    for (int i = 0; i < nx; ++i) {
      a[IDX(i, j)] = spha[IDX(i, j)] / (r2 * (cy * cy));
      b[IDX(i, j)] = sphb[IDX(i, j)] / r2;
      c[IDX(i, j)] = sphc[IDX(i, j)] / (r2 * cy);
    }
  }
}

/* r = min(1.0, (x0-xx(idx0))/d) assigned to an implicit REAL(4) (:829,:1330) */
static float weight_r4(double v) { return (float)(v < 1.0 ? v : 1.0); }

/* Forward setup shared by both entry points: source bilinear slowness fun0
 * (:825-841, only fun0 is used by the lon/lat T0), and optionally T0, T0x,
 * T0y (:845-857). */
static int fwd_source(int nx, int ny, const double *xx, const double *yy,
                      const double *fun, double x0, double y0, double dx,
                      double dy, double *fun0_out) {
  const int idx0 = (int)floor((x0 - xx[0]) / dx + 1); /* 1-based */
  const int idy0 = (int)floor((y0 - yy[0]) / dy + 1);
  /* a bracket outside the grid means the source sits on (or beyond) the
   * boundary; the reference reads out of bounds here and then stops at
   * :863-867 */
  if (idx0 < 1 || idy0 < 1 || idx0 + 1 > nx || idy0 + 1 > ny)
    return ORC_E_SRC_ON_BOUNDARY;
  const int i0 = idx0 - 1, j0 = idy0 - 1;
  const float r1 = weight_r4((x0 - xx[i0]) / dx);
  const float r2 = weight_r4((y0 - yy[j0]) / dy);
  const float w00 = (1.0f - r1) * (1.0f - r2);
  const float w01 = (1.0f - r1) * r2;
  const float w10 = r1 * (1.0f - r2);
  const float w11 = r1 * r2;
  *fun0_out = (double)w00 * fun[IDX(i0, j0)] + (double)w01 * fun[IDX(i0, j0 + 1)] +
              (double)w10 * fun[IDX(i0 + 1, j0)] +
              (double)w11 * fun[IDX(i0 + 1, j0 + 1)];
  return ORC_OK;
}

static void fwd_t0_fields(int nx, int ny, const double *xx, const double *yy,
                          double x0, double y0, double fun0, double *T0,
                          double *T0x, double *T0y) {
  const double sy0 = sin(y0), cy0 = cos(y0);
  for (int i = 0; i < nx; ++i) {
    const double sdx = sin(xx[i] - x0), cdx = cos(xx[i] - x0);
    for (int j = 0; j < ny; ++j) {
      const double sy = sin(yy[j]), cy = cos(yy[j]);
      const double t = sy * sy0 + cy * cy0 * cdx;
      double tc = t < 1.0 ? t : 1.0;
      tc = tc > -1.0 ? tc : -1.0;
      const size_t k = IDX(i, j);
      T0[k] = acos(tc) * R_EARTH * fun0;
      if (T0[k] == 0) {
        T0x[k] = 0;
        T0y[k] = 0;
      } else {
        const double s = sqrt(1 - t * t);
        T0x[k] = -(1.0 / s * cy * cy0 * (-1.0) * sdx * R_EARTH * fun0);
        T0y[k] = -(1.0 / s * (cy * sy0 - sy * cy0 * cdx) * R_EARTH * fun0);
      }
    }
  }
}

/* tau init and freeze (:859-872) */
static int fwd_tau_init(int nx, int ny, const double *xx, const double *yy,
                        double x0, double y0, double dx, double dy,
                        double *tau, unsigned char *ischange) {
  for (int i = 0; i < nx; ++i) {
    for (int j = 0; j < ny; ++j) {
      const size_t k = IDX(i, j);
      if (fabs((xx[i] - x0) / dx) <= 1 && fabs((yy[j] - y0) / dy) <= 1) {
        tau[k] = 1;
        ischange[k] = 0;
        if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1)
          return ORC_E_SRC_ON_BOUNDARY;
      } else {
        tau[k] = 10;
        ischange[k] = 1;
      }
    }
  }
  return ORC_OK;
}

/* One cell of the forward sweep body (:910-1079). */
static void fwd_cell(int nx, int ny, int i, int j, double dx, double dy,
                     const double *a, const double *b, const double *c,
                     const double *fun, const double *T0, const double *T0x,
                     const double *T0y, double *tau) {
  const size_t k = IDX(i, j);
  const double ak = a[k], bk = b[k], ck = c[k], fk = fun[k];
  const double t0 = T0[k], t0x = T0x[k], t0y = T0y[k];
  double cand[16];
  int nc = 0;
  double px, qx, py, qy;

  for (int icase = 1; icase <= 4; ++icase) {
    switch (icase) {
    case 1: /* -1, -1 */
      if (i == 0 || j == 0) continue;
      px = (t0x + t0 / dx); qx = -(t0 / dx * tau[IDX(i - 1, j)]);
      py = (t0y + t0 / dy); qy = -(t0 / dy * tau[IDX(i, j - 1)]);
      break;
    case 2: /* -1, +1 */
      if (i == 0 || j == ny - 1) continue;
      px = (t0x + t0 / dx); qx = -(t0 / dx * tau[IDX(i - 1, j)]);
      py = (t0y - t0 / dy); qy = t0 / dy * tau[IDX(i, j + 1)];
      break;
    case 3: /* +1, -1 */
      if (i == nx - 1 || j == 0) continue;
      px = (t0x - t0 / dx); qx = +(t0 / dx * tau[IDX(i + 1, j)]);
      py = (t0y + t0 / dy); qy = -(t0 / dy * tau[IDX(i, j - 1)]);
      break;
    default: /* +1, +1 */
      if (i == nx - 1 || j == ny - 1) continue;
      px = (t0x - t0 / dx); qx = t0 / dx * tau[IDX(i + 1, j)];
      py = (t0y - t0 / dy); qy = t0 / dy * tau[IDX(i, j + 1)];
      break;
    }
    const double ea = ak * (px * px) + bk * (py * py) - 2 * ck * px * py;
    const double eb = 2 * ak * px * qx + 2 * bk * py * qy - 2 * ck * (px * qy + py * qx);
    const double ec = ak * (qx * qx) + bk * (qy * qy) - 2 * ck * qx * qy - fk * fk;
    const double delta = eb * eb - 4 * ea * ec;
    if (delta >= 0) {
      for (int isol = 1; isol <= 2; ++isol) {
        const double tt = isol == 1 ? (-eb + sqrt(delta)) / (2 * ea)
                                    : (-eb - sqrt(delta)) / (2 * ea);
        const double Tx = px * tt + qx;
        const double Ty = py * tt + qy;
        const double chx = ak * Tx - ck * Ty;
        const double chy = bk * Ty - ck * Tx;
        int ok = 0;
        switch (icase) {
        case 1: ok = (chx >= 0 && chy >= 0); break;
        case 2: ok = (chx >= 0 && chy <= 0); break;
        case 3: ok = (chx <= 0 && chy >= 0); break;
        default: ok = (chx <= 0 && chy <= 0); break;
        }
        if (ok) cand[nc++] = tt;
      }
    }
  }

  for (int icase = 1; icase <= 4; ++icase) {
    size_t kn;
    double p, q, dis;
    switch (icase) {
    case 1:
      if (i == 0) continue;
      kn = IDX(i - 1, j);
      p = (t0x + t0 / dx); q = -(t0 / dx * tau[kn]);
      dis = sqrt(fk * fk * (bk / (ak * bk - ck * ck)));
      break;
    case 2:
      if (i == nx - 1) continue;
      kn = IDX(i + 1, j);
      p = (t0x - t0 / dx); q = t0 / dx * tau[kn];
      dis = sqrt(fk * fk * (bk / (ak * bk - ck * ck)));
      break;
    case 3:
      if (j == 0) continue;
      kn = IDX(i, j - 1);
      p = (t0y + t0 / dy); q = -(t0 / dy * tau[kn]);
      dis = sqrt(fk * fk * (ak / (ak * bk - ck * ck)));
      break;
    default:
      if (j == ny - 1) continue;
      kn = IDX(i, j + 1);
      p = (t0y - t0 / dy); q = t0 / dy * tau[kn];
      dis = sqrt(fk * fk * (ak / (ak * bk - ck * ck)));
      break;
    }
    const double tn = tau[kn], t0n = T0[kn];
    double tt = (dis - q) / p;
    if (tt * t0 >= tn * t0n && tt > tn * 0.5) cand[nc++] = tt;
    tt = (-dis - q) / p;
    if (tt * t0 >= tn * t0n && tt > tn * 0.5) cand[nc++] = tt;
  }

  double t = tau[k];
  for (int ic = 0; ic < nc; ++ic) t = cand[ic] < t ? cand[ic] : t;
  tau[k] = t;
}

/* Iterations (:881-1119) and T = tau*T0 (:1124).  a,b,c already metric. */
static int fwd_iterate(int nx, int ny, double dx, double dy, const double *a,
                       const double *b, const double *c, const double *fun,
                       const double *T0, const double *T0x, const double *T0y,
                       double *tau, const unsigned char *ischange,
                       double *tau_old, double *T, int *iters) {
  int iter, converged = 0;
  const size_t n = (size_t)nx * ny;
  for (iter = 1; iter <= FWD_MAXITER; ++iter) {
    memcpy(tau_old, tau, n * sizeof(double));
    for (int isw = 1; isw <= 4; ++isw) {
      const int xd = isw <= 2 ? 1 : -1;
      const int yd = (isw == 1 || isw == 3) ? 1 : -1;
      for (int ii = 0; ii < nx; ++ii) {
        const int i = xd > 0 ? ii : nx - 1 - ii;
        for (int jj = 0; jj < ny; ++jj) {
          const int j = yd > 0 ? jj : ny - 1 - jj;
          if (ischange[IDX(i, j)])
            fwd_cell(nx, ny, i, j, dx, dy, a, b, c, fun, T0, T0x, T0y, tau);
        }
      }
    }
    double l1 = 0, linf = 0;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        const double d = fabs(tau[IDX(i, j)] - tau_old[IDX(i, j)]) * T0[IDX(i, j)];
        l1 = l1 + d;
        linf = linf > d ? linf : d; /* max(Linf, d) */
      }
    l1 = l1 / (double)nx / (double)ny;
    if (fabs(l1) < FWD_TOL && linf < FWD_TOL) {
      converged = 1;
      break;
    }
  }
  if (!converged) iter = FWD_MAXITER;
  for (size_t k = 0; k < n; ++k) T[k] = tau[k] * T0[k];
  if (iters) *iters = iter;
  return converged ? ORC_OK : ORC_W_MAXITER;
}

static int fwd_common(const double *xx_deg, const double *yy_deg, int nx,
                      int ny, const double *spha, const double *sphb,
                      const double *sphc, const double *fun, double x0_deg,
                      double y0_deg, const double *abc_in, const double *T0_in,
                      const double *T0x_in, const double *T0y_in, double *T,
                      int *iters) {
  if (nx < 2 || ny < 2) return ORC_E_ARG;
  const size_t n = (size_t)nx * ny;
  double *xx = malloc(sizeof(double) * nx), *yy = malloc(sizeof(double) * ny);
  double *buf = malloc(sizeof(double) * n * 8);
  unsigned char *isc = malloc(n);
  if (!xx || !yy || !buf || !isc) {
    free(xx); free(yy); free(buf); free(isc);
    return ORC_E_NOMEM;
  }
  double *a = buf, *b = buf + n, *c = buf + 2 * n, *T0 = buf + 3 * n,
         *T0x = buf + 4 * n, *T0y = buf + 5 * n, *tau = buf + 6 * n,
         *tau_old = buf + 7 * n;
  for (int i = 0; i < nx; ++i) xx[i] = deg2rad_ref(xx_deg[i]);
  for (int j = 0; j < ny; ++j) yy[j] = deg2rad_ref(yy_deg[j]);
  const double x0 = deg2rad_ref(x0_deg), y0 = deg2rad_ref(y0_deg);
  const double dx = xx[1] - xx[0], dy = yy[1] - yy[0];
  if (abc_in) {
    memcpy(a, abc_in, n * sizeof(double));
    memcpy(b, abc_in + n, n * sizeof(double));
    memcpy(c, abc_in + 2 * n, n * sizeof(double));
  } else {
    metric_abc(nx, ny, yy, spha, sphb, sphc, a, b, c);
  }
  int st = ORC_OK;
  if (T0_in) {
    memcpy(T0, T0_in, n * sizeof(double));
    memcpy(T0x, T0x_in, n * sizeof(double));
    memcpy(T0y, T0y_in, n * sizeof(double));
  } else {
    double fun0;
    st = fwd_source(nx, ny, xx, yy, fun, x0, y0, dx, dy, &fun0);
    if (st == ORC_OK) fwd_t0_fields(nx, ny, xx, yy, x0, y0, fun0, T0, T0x, T0y);
  }
  if (st == ORC_OK)
    st = fwd_tau_init(nx, ny, xx, yy, x0, y0, dx, dy, tau, isc);
  if (st == ORC_OK)
    st = fwd_iterate(nx, ny, dx, dy, a, b, c, fun, T0, T0x, T0y, tau, isc,
                     tau_old, T, iters);
  free(xx); free(yy); free(buf); free(isc);
  return st;
}

int orc_fwd_lonlat(const double *xx_deg, const double *yy_deg, int nx, int ny,
                   const double *spha, const double *sphb, const double *sphc,
                   const double *fun, double x0_deg, double y0_deg, double *T,
                   int *iters) {
  return fwd_common(xx_deg, yy_deg, nx, ny, spha, sphb, sphc, fun, x0_deg,
                    y0_deg, NULL, NULL, NULL, NULL, T, iters);
}

int orc_fwd_lonlat_pre(const double *xx_deg, const double *yy_deg, int nx,
                       int ny, const double *a, const double *b,
                       const double *c, const double *fun, double x0_deg,
                       double y0_deg, const double *T0, const double *T0x,
                       const double *T0y, double *T, int *iters) {
  const size_t n = (size_t)nx * ny;
  double *abc = malloc(sizeof(double) * 3 * n);
  if (!abc) return ORC_E_NOMEM;
  memcpy(abc, a, n * sizeof(double));
  memcpy(abc + n, b, n * sizeof(double));
  memcpy(abc + 2 * n, c, n * sizeof(double));
  const int st = fwd_common(xx_deg, yy_deg, nx, ny, NULL, NULL, NULL, fun,
                            x0_deg, y0_deg, abc, T0, T0x, T0y, T, iters);
  free(abc);
  return st;
}

int orc_fwd_t0(const double *xx_deg, const double *yy_deg, int nx, int ny,
               const double *fun, double x0_deg, double y0_deg, double *T0,
               double *T0x, double *T0y) {
  if (nx < 2 || ny < 2) return ORC_E_ARG;
  double *xx = malloc(sizeof(double) * nx), *yy = malloc(sizeof(double) * ny);
  if (!xx || !yy) { free(xx); free(yy); return ORC_E_NOMEM; }
  for (int i = 0; i < nx; ++i) xx[i] = deg2rad_ref(xx_deg[i]);
  for (int j = 0; j < ny; ++j) yy[j] = deg2rad_ref(yy_deg[j]);
  const double x0 = deg2rad_ref(x0_deg), y0 = deg2rad_ref(y0_deg);
  const double dx = xx[1] - xx[0], dy = yy[1] - yy[0];
  double fun0;
  int st = fwd_source(nx, ny, xx, yy, fun, x0, y0, dx, dy, &fun0);
  if (st == ORC_OK) fwd_t0_fields(nx, ny, xx, yy, x0, y0, fun0, T0, T0x, T0y);
  free(xx); free(yy);
  return st;
}

/* ------------------------------------------------------------------ adjoint */

static void adj_kform(int nx, int ny, double dx, double dy, const double *T,
                      const double *a, const double *b, const double *c,
                      const double *delta, float *kk);
static int adj_phase32(int nx, int ny, const float *kk, double *Ta, int *done);

static int adj_common(const double *xx_deg, const double *yy_deg, int nx,
                      int ny, const double *spha, const double *sphb,
                      const double *sphc, const double *T,
                      const double *xrec_deg, const double *yrec_deg,
                      const double *src_adj, int nr, const double *pre,
                      int mixed, double *Ta, int *iters) {
  if (nx < 3 || ny < 3 || nr < 0) return ORC_E_ARG;
  const size_t n = (size_t)nx * ny;
  double *xx = malloc(sizeof(double) * nx), *yy = malloc(sizeof(double) * ny);
  double *buf = malloc(sizeof(double) * n * 9);
  if (!xx || !yy || !buf) {
    free(xx); free(yy); free(buf);
    return ORC_E_NOMEM;
  }
  double *a = buf, *b = buf + n, *c = buf + 2 * n, *a1m = buf + 3 * n,
         *a1p = buf + 4 * n, *a2m = buf + 5 * n, *a2p = buf + 6 * n,
         *delta = buf + 7 * n, *old = buf + 8 * n;
  /* b1m..b2p share storage with a, b, c after those are consumed below;
   * keep them separate for clarity instead. */
  double *bb = malloc(sizeof(double) * n * 4);
  if (!bb) { free(xx); free(yy); free(buf); return ORC_E_NOMEM; }
  double *b1m = bb, *b1p = bb + n, *b2m = bb + 2 * n, *b2p = bb + 3 * n;
  int st = ORC_OK;

  for (int i = 0; i < nx; ++i) xx[i] = deg2rad_ref(xx_deg[i]);
  for (int j = 0; j < ny; ++j) yy[j] = deg2rad_ref(yy_deg[j]);
  const double dx = xx[1] - xx[0], dy = yy[1] - yy[0];

  /* delta source (:1320-1336) */
  memset(delta, 0, n * sizeof(double));
  if (pre) {
    memcpy(delta, pre + 3 * n, n * sizeof(double));
    nr = 0;
  }
  for (int ir = 0; ir < nr; ++ir) {
    const double xr = deg2rad_ref(xrec_deg[ir]), yr = deg2rad_ref(yrec_deg[ir]);
    const int idx0 = (int)floor((xr - xx[0]) / dx + 1);
    const int idy0 = (int)floor((yr - yy[0]) / dy + 1);
    if (idx0 < 1 || idy0 < 1 || idx0 + 1 > nx || idy0 + 1 > ny) {
      st = ORC_E_REC_OUT_OF_GRID;
      break;
    }
    const int i0 = idx0 - 1, j0 = idy0 - 1;
    const float r1 = weight_r4((xr - xx[i0]) / dx);
    const float r2 = weight_r4((yr - yy[j0]) / dy);
    const double sa = src_adj[ir];
    const double den = dx * R_EARTH * cos(yr) * dy * R_EARTH;
    delta[IDX(i0, j0)] = delta[IDX(i0, j0)] + sa * (double)(1.0f - r1) * (double)(1.0f - r2) / den;
    delta[IDX(i0, j0 + 1)] = delta[IDX(i0, j0 + 1)] + sa * (double)(1.0f - r1) * (double)r2 / den;
    delta[IDX(i0 + 1, j0)] = delta[IDX(i0 + 1, j0)] + sa * (double)r1 * (double)(1.0f - r2) / den;
    delta[IDX(i0 + 1, j0 + 1)] = delta[IDX(i0 + 1, j0 + 1)] + sa * (double)r1 * (double)r2 / den;
  }
  if (st != ORC_OK) {
    free(xx); free(yy); free(buf); free(bb);
    return st;
  }

  if (pre) {
    memcpy(a, pre, n * sizeof(double));
    memcpy(b, pre + n, n * sizeof(double));
    memcpy(c, pre + 2 * n, n * sizeof(double));
  } else {
    metric_abc(nx, ny, yy, spha, sphb, sphc, a, b, c);
  }

  /* face coefficients (:1348-1370) */
  for (int i = 1; i < nx - 1; ++i) {
    for (int j = 1; j < ny - 1; ++j) {
      const double a1 =
          -((T[IDX(i, j)] - T[IDX(i - 1, j)]) / dx * (a[IDX(i, j)] + a[IDX(i - 1, j)]) / 2) +
          (T[IDX(i, j + 1)] - T[IDX(i, j - 1)] + T[IDX(i - 1, j + 1)] - T[IDX(i - 1, j - 1)]) /
              (4 * dy) * (c[IDX(i, j)] + c[IDX(i - 1, j)]) / 2;
      a1m[IDX(i, j)] = (a1 - fabs(a1)) / 2;
      a1p[IDX(i, j)] = (a1 + fabs(a1)) / 2;
      const double a2 =
          -((T[IDX(i + 1, j)] - T[IDX(i, j)]) / dx * (a[IDX(i + 1, j)] + a[IDX(i, j)]) / 2) +
          (T[IDX(i + 1, j + 1)] - T[IDX(i + 1, j - 1)] + T[IDX(i, j + 1)] - T[IDX(i, j - 1)]) /
              (4 * dy) * (c[IDX(i + 1, j)] + c[IDX(i, j)]) / 2;
      a2m[IDX(i, j)] = (a2 - fabs(a2)) / 2;
      a2p[IDX(i, j)] = (a2 + fabs(a2)) / 2;
      const double b1 =
          -((T[IDX(i, j)] - T[IDX(i, j - 1)]) / dy * (b[IDX(i, j)] + b[IDX(i, j - 1)]) / 2) +
          (T[IDX(i + 1, j)] - T[IDX(i - 1, j)] + T[IDX(i + 1, j - 1)] - T[IDX(i - 1, j - 1)]) /
              (4 * dx) * (c[IDX(i, j)] + c[IDX(i, j - 1)]) / 2;
      b1m[IDX(i, j)] = (b1 - fabs(b1)) / 2;
      b1p[IDX(i, j)] = (b1 + fabs(b1)) / 2;
      const double b2 =
          -((T[IDX(i, j + 1)] - T[IDX(i, j)]) / dy * (b[IDX(i, j + 1)] + b[IDX(i, j)]) / 2) +
          (T[IDX(i + 1, j + 1)] - T[IDX(i - 1, j + 1)] + T[IDX(i + 1, j)] - T[IDX(i - 1, j)]) /
              (4 * dx) * (c[IDX(i, j + 1)] + c[IDX(i, j)]) / 2;
      b2m[IDX(i, j)] = (b2 - fabs(b2)) / 2;
      b2p[IDX(i, j)] = (b2 + fabs(b2)) / 2;
      Ta[IDX(i, j)] = 100;
    }
  }
  /* boundary fixed at zero (:1373-1380) */
  for (int i = 0; i < nx; ++i) { Ta[IDX(i, 0)] = 0; Ta[IDX(i, ny - 1)] = 0; }
  for (int j = 0; j < ny; ++j) { Ta[IDX(0, j)] = 0; Ta[IDX(nx - 1, j)] = 0; }

  int it32 = 0, done32 = 0;
  /* mixed mode: the float k fields, the FP32 phase on them, then exactly
   * one FP64 polish iteration on the same fields, widened */
  float *kk = NULL;
  if (mixed) {
    kk = malloc(sizeof(float) * n * 5);
    if (!kk) {
      free(xx); free(yy); free(buf); free(bb);
      return ORC_E_NOMEM;
    }
    adj_kform(nx, ny, dx, dy, T, a, b, c, delta, kk);
    it32 = adj_phase32(nx, ny, kk, Ta, &done32);
    if (it32 < 0) {
      free(xx); free(yy); free(buf); free(bb); free(kk);
      return it32;
    }
  }
  const int maxit = mixed ? 1 : ADJ_MAXITER;
  int iter, converged = 0;
  for (iter = 1; iter <= maxit; ++iter) {
    memcpy(old, Ta, n * sizeof(double));
    /* do xdirec=-1,1,2; do ydirec=-1,1,2 with nint bounds (:1385-1390):
     * xdirec=-1 runs x over 2..nx-1 ascending, +1 descending. */
    for (int xdir = -1; xdir <= 1; xdir += 2) {
      for (int ydir = -1; ydir <= 1; ydir += 2) {
        for (int ii = 1; ii < nx - 1; ++ii) {
          const int i = xdir < 0 ? ii : nx - 1 - ii;
          for (int jj = 1; jj < ny - 1; ++jj) {
            const int j = ydir < 0 ? jj : ny - 1 - jj;
            const size_t k = IDX(i, j);
            if (mixed) {
              /* k0 + kW*TaW + kE*TaE + kS*TaS + kN*TaN, one rounding per
               * operation, in the FP32 phase's order */
              double t = (double)kk[k];
              t = (double)kk[n + k] * Ta[IDX(i - 1, j)] + t;
              t = (double)kk[2 * n + k] * Ta[IDX(i + 1, j)] + t;
              t = (double)kk[3 * n + k] * Ta[IDX(i, j - 1)] + t;
              Ta[k] = (double)kk[4 * n + k] * Ta[IDX(i, j + 1)] + t;
              continue;
            }
            const double d = (a2p[k] - a1m[k]) / dx + (b2p[k] - b1m[k]) / dy;
            if (fabs(d) < ADJ_EPS) {
              Ta[k] = 0;
            } else {
              const double e = (Ta[IDX(i - 1, j)] * a1p[k] - Ta[IDX(i + 1, j)] * a2m[k]) / dx +
                               (Ta[IDX(i, j - 1)] * b1p[k] - Ta[IDX(i, j + 1)] * b2m[k]) / dy;
              Ta[k] = (delta[k] + e) / d;
            }
          }
        }
      }
    }
    double l1 = 0, linf = 0;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        const double dd = fabs(Ta[IDX(i, j)] - old[IDX(i, j)]);
        l1 = l1 + dd;
        linf = linf > dd ? linf : dd;
      }
    l1 = l1 / (double)(nx * ny);
    if (fabs(l1) < ADJ_TOL && linf < ADJ_TOL) {
      converged = 1;
      break;
    }
  }
  if (mixed) {
    converged = done32;
    if (iters) { iters[0] = it32; iters[1] = 1; }
  } else {
    if (!converged) iter = ADJ_MAXITER;
    if (iters) *iters = iter;
  }
  free(xx); free(yy); free(buf); free(bb); free(kk);
  return converged ? ORC_OK : ORC_W_MAXITER;
}

int orc_adj_lonlat(const double *xx_deg, const double *yy_deg, int nx, int ny,
                   const double *spha, const double *sphb, const double *sphc,
                   const double *T, const double *xrec_deg,
                   const double *yrec_deg, const double *src_adj, int nr,
                   double *Ta, int *iters) {
  return adj_common(xx_deg, yy_deg, nx, ny, spha, sphb, sphc, T, xrec_deg,
                    yrec_deg, src_adj, nr, NULL, 0, Ta, iters);
}

int orc_adj_lonlat_pre(const double *xx_deg, const double *yy_deg, int nx,
                       int ny, const double *a, const double *b,
                       const double *c, const double *T, const double *delta,
                       double *Ta, int *iters) {
  const size_t n = (size_t)nx * ny;
  double *pre = malloc(sizeof(double) * 4 * n);
  if (!pre) return ORC_E_NOMEM;
  memcpy(pre, a, n * sizeof(double));
  memcpy(pre + n, b, n * sizeof(double));
  memcpy(pre + 2 * n, c, n * sizeof(double));
  memcpy(pre + 3 * n, delta, n * sizeof(double));
  const int st = adj_common(xx_deg, yy_deg, nx, ny, NULL, NULL, NULL, T, NULL,
                            NULL, NULL, 0, pre, 0, Ta, iters);
  free(pre);
  return st;
}

/* --------------------------------------------------------------- mask */

void orc_mask(const double *xx, const double *yy, int nx, int ny,
              double *table, double x0, double y0) {
  const double dx = xx[1] - xx[0], dy = yy[1] - yy[0];
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      if (fabs(xx[i] - x0) < 1.5 * dx && fabs(yy[j] - y0) < 1.5 * dy)
        table[IDX(i, j)] = 0;
}

/* ------------------------------------------------------------- interp2 */

/* minloc(xq - x, mask = xq >= x): first index of the smallest non-negative
 * difference; maxloc(xq - x, mask = xq < x): first index of the largest
 * negative difference (src/shared/utils.f90:925-928).  -1 when the mask is
 * empty. */
static int masked_minloc_ge(const double *x, int n, double q) {
  int best = -1;
  double bv = 0;
  for (int k = 0; k < n; ++k)
    if (q >= x[k]) {
      const double v = q - x[k];
      if (best < 0 || v < bv) { best = k; bv = v; }
    }
  return best;
}
static int masked_maxloc_lt(const double *x, int n, double q) {
  int best = -1;
  double bv = 0;
  for (int k = 0; k < n; ++k)
    if (q < x[k]) {
      const double v = q - x[k];
      if (best < 0 || v > bv) { best = k; bv = v; }
    }
  return best;
}

static int interp2_0(const double *x, const double *y, int nx, int ny,
                     const double *v, double xq, double yq, double *out) {
  const int x1 = masked_minloc_ge(x, nx, xq), y1 = masked_minloc_ge(y, ny, yq);
  const int x2 = masked_maxloc_lt(x, nx, xq), y2 = masked_maxloc_lt(y, ny, yq);
  if (x1 < 0 || y1 < 0 || x2 < 0 || y2 < 0) return ORC_E_REC_OUT_OF_GRID;
  const double vn = fabs((x[x2] - x[x1]) * (y[y2] - y[y1]));
  const double xr[2] = {x[x1], x[x2]}, yr[2] = {y[y1], y[y2]};
  static const int ix[4] = {1, 0, 1, 0}, iy[4] = {1, 1, 0, 0};
  double N[4];
  for (int q = 0; q < 4; ++q) N[q] = fabs((xr[ix[q]] - xq) * (yr[iy[q]] - yq));
  const double vr[4] = {v[IDX(x1, y1)], v[IDX(x2, y1)], v[IDX(x1, y2)], v[IDX(x2, y2)]};
  double s = 0;
  for (int q = 0; q < 4; ++q) s = s + vr[q] * (N[q] / vn);
  *out = s;
  return ORC_OK;
}

int orc_interp2(const double *x, const double *y, int nx, int ny,
                const double *v, const double *xq, const double *yq, int nq,
                double *vq) {
  for (int q = 0; q < nq; ++q) {
    const int st = interp2_0(x, y, nx, ny, v, xq[q], yq[q], vq + q);
    if (st != ORC_OK) return st;
  }
  return ORC_OK;
}

/* ------------------------------------------------- forward_simulate loop */

typedef struct {
  int dim, nx, ny, nperiod, nitems, npath, istotable, isadj;
  const double *xg, *yg, *svel, *m11, *m12, *m22;
  const int *item_pidx, *rec_off, *idx_path;
  const double *srcx, *srcy, *recx, *recy, *tt, *weight;
  /* per block outputs */
  int istart, iend; /* inclusive-exclusive item range */
  double *tt_fwd, *adj_s, *adj_den, *item_chi, *weight_out;
  int *iters;
  double chi;
  int status;
} fs_block;

/* One static block of items: the body of forward_simulate
 * (src/tomo2d.f90:170-194 / src/acqui.f90:122-151). */
static void *fs_run_block(void *arg) {
  fs_block *B = (fs_block *)arg;
  const int nx = B->nx, ny = B->ny, np = B->nperiod;
  const size_t n = (size_t)nx * ny;
  double *work = malloc(sizeof(double) * n * 9);
  double *tek = NULL, *w = NULL, *rhs = NULL;
  int maxr = 0;
  for (int it = B->istart; it < B->iend; ++it) {
    const int nr = B->rec_off[it + 1] - B->rec_off[it];
    if (nr > maxr) maxr = nr;
  }
  tek = malloc(sizeof(double) * (maxr + 1));
  w = malloc(sizeof(double) * (maxr + 1));
  rhs = malloc(sizeof(double) * (maxr + 1));
  B->chi = 0;
  B->status = ORC_OK;
  if (!work || !tek || !w || !rhs) {
    B->status = ORC_E_NOMEM;
    goto done;
  }
  double *fun = work, *spha = work + n, *sphb = work + 2 * n, *sphc = work + 3 * n,
         *T = work + 4 * n, *adj = work + 5 * n, *kden = work + 6 * n,
         *vsl = work + 7 * n;
  for (int it = B->istart; it < B->iend; ++it) {
    const int p = B->item_pidx[it];
    const int r0 = B->rec_off[it], nr = B->rec_off[it + 1] - r0;
    /* svel(pidx,:,:) etc. (copy-in of the strided section) */
    for (size_t k = 0; k < n; ++k) {
      const size_t g = (size_t)p + (size_t)np * k;
      vsl[k] = B->svel[g];
      fun[k] = 1.0 / B->svel[g];  /* vtmp = 1./vel2d (measadj.f90:67) */
      spha[k] = B->m11[g];        /* m11, m22, -m12 (measadj.f90:68-70) */
      sphb[k] = B->m22[g];
      sphc[k] = -B->m12[g];
    }
    int fit = 0, ait = 0, dit = 0, st;
    st = orc_fwd_lonlat(B->xg, B->yg, nx, ny, spha, sphb, sphc, fun,
                        B->srcx[it], B->srcy[it], T, &fit);
    if (st < 0) { B->status = st; goto done; }
    /* interp2 + chi (measadj.f90:71-72) */
    st = orc_interp2(B->xg, B->yg, nx, ny, T, B->recx + r0, B->recy + r0, nr, tek);
    if (st < 0) { B->status = st; goto done; }
    double s = 0;
    for (int r = 0; r < nr; ++r) {
      const double d = tek[r] - B->tt[r0 + r];
      s = s + B->weight[r0 + r] * (d * d);
    }
    const double chi_item = 0.5 * s;
    if (B->item_chi) B->item_chi[it] = chi_item;
    if (B->istotable && B->tt_fwd)
      for (int r = 0; r < nr; ++r) B->tt_fwd[B->idx_path[r0 + r]] = tek[r];
    for (int r = 0; r < nr; ++r) w[r] = B->weight[r0 + r];
    if (B->isadj) {
      /* outlier rejection and adjoint source (measadj.f90:82-88) */
      for (int r = 0; r < nr; ++r)
        if (fabs(tek[r] - B->tt[r0 + r]) / B->tt[r0 + r] > 0.5) w[r] = 0.;
      for (int r = 0; r < nr; ++r) rhs[r] = w[r] * (tek[r] - B->tt[r0 + r]);
      st = orc_adj_lonlat(B->xg, B->yg, nx, ny, spha, sphb, sphc, T,
                          B->recx + r0, B->recy + r0, rhs, nr, adj, &ait);
      if (st < 0) { B->status = st; goto done; }
      orc_mask(B->xg, B->yg, nx, ny, adj, B->srcx[it], B->srcy[it]);
      if (B->dim == 3) {
        /* run_adjoint_density (measadj.f90:92-104): RHS -1.0*weight */
        for (int r = 0; r < nr; ++r) rhs[r] = -1.0 * w[r];
        st = orc_adj_lonlat(B->xg, B->yg, nx, ny, spha, sphb, sphc, T,
                            B->recx + r0, B->recy + r0, rhs, nr, kden, &dit);
        if (st < 0) { B->status = st; goto done; }
        orc_mask(B->xg, B->yg, nx, ny, kden, B->srcx[it], B->srcy[it]);
        /* post_proc_eikokernel (acqui.f90:314-315): /svel**3 */
        for (size_t k = 0; k < n; ++k) {
          const size_t g = (size_t)p + (size_t)np * k;
          const double v3 = vsl[k] * vsl[k] * vsl[k];
          B->adj_s[g] = B->adj_s[g] + adj[k] / v3;
          B->adj_den[g] = B->adj_den[g] + kden[k] / v3;
        }
      } else {
        /* post_proc_eikokernel (tomo2d.f90:155): /svel**2 */
        for (size_t k = 0; k < n; ++k) {
          const size_t g = (size_t)p + (size_t)np * k;
          B->adj_s[g] = B->adj_s[g] + adj[k] / (vsl[k] * vsl[k]);
        }
      }
    }
    if (B->weight_out)
      for (int r = 0; r < nr; ++r) B->weight_out[r0 + r] = w[r];
    if (B->iters) {
      B->iters[3 * it] = fit;
      B->iters[3 * it + 1] = ait;
      B->iters[3 * it + 2] = dit;
    }
    B->chi = B->chi + chi_item;
  }
done:
  free(work); free(tek); free(w); free(rhs);
  return NULL;
}

/* scatter_all_i (src/shared/parallel.f90:1877-1903), 0-based half-open */
static void scatter_block(int count, int size, int rank, int *s, int *e) {
  if (count < size) {
    if (rank < count) { *s = rank; *e = rank + 1; }
    else { *s = 0; *e = 0; }
  } else {
    const int c = count / size, rem = count % size;
    if (rank < rem) { *s = rank * (c + 1); *e = *s + c + 1; }
    else { *s = rank * c + rem; *e = *s + c; }
  }
}

int orc_forward_simulate(int dim, int nx, int ny, const double *xgrids,
                         const double *ygrids, int nperiod, const double *svel,
                         const double *m11, const double *m12,
                         const double *m22, int nitems, const int *item_pidx,
                         const double *item_srcx, const double *item_srcy,
                         const int *rec_off, const double *recx,
                         const double *recy, const double *tt,
                         const double *weight, const int *idx_path, int npath,
                         int istotable, int isadj, int nthreads,
                         double *tt_fwd, double *chi, double *adj_s,
                         double *adj_density, double *item_chi, int *iters,
                         double *weight_out) {
  if ((dim != 2 && dim != 3) || nx < 3 || ny < 3 || nperiod < 1 || nitems < 0)
    return ORC_E_ARG;
  if (isadj && (!adj_s || (dim == 3 && !adj_density))) return ORC_E_ARG;
  if (nthreads < 1) nthreads = 1;
  const size_t nk = (size_t)nperiod * nx * ny;
  fs_block *blk = calloc((size_t)nthreads, sizeof(fs_block));
  pthread_t *th = calloc((size_t)nthreads, sizeof(pthread_t));
  if (!blk || !th) { free(blk); free(th); return ORC_E_NOMEM; }
  int st = ORC_OK;
  for (int r = 0; r < nthreads; ++r) {
    fs_block *B = &blk[r];
    B->dim = dim; B->nx = nx; B->ny = ny; B->nperiod = nperiod;
    B->nitems = nitems; B->npath = npath; B->istotable = istotable; B->isadj = isadj;
    B->xg = xgrids; B->yg = ygrids; B->svel = svel; B->m11 = m11; B->m12 = m12; B->m22 = m22;
    B->item_pidx = item_pidx; B->rec_off = rec_off; B->idx_path = idx_path;
    B->srcx = item_srcx; B->srcy = item_srcy; B->recx = recx; B->recy = recy;
    B->tt = tt; B->weight = weight;
    scatter_block(nitems, nthreads, r, &B->istart, &B->iend);
    B->tt_fwd = istotable ? calloc((size_t)(npath > 0 ? npath : 1), sizeof(double)) : NULL;
    B->adj_s = isadj ? calloc(nk, sizeof(double)) : NULL;
    B->adj_den = (isadj && dim == 3) ? calloc(nk, sizeof(double)) : NULL;
    B->item_chi = item_chi; B->iters = iters; B->weight_out = weight_out;
    if ((istotable && !B->tt_fwd) || (isadj && !B->adj_s) ||
        (isadj && dim == 3 && !B->adj_den))
      st = ORC_E_NOMEM;
  }
  if (st == ORC_OK) {
    if (nthreads == 1) {
      fs_run_block(&blk[0]);
    } else {
      for (int r = 0; r < nthreads; ++r) pthread_create(&th[r], NULL, fs_run_block, &blk[r]);
      for (int r = 0; r < nthreads; ++r) pthread_join(th[r], NULL);
    }
    for (int r = 0; r < nthreads; ++r)
      if (blk[r].status < 0 && st == ORC_OK) st = blk[r].status;
  }
  if (st == ORC_OK) {
    /* sum_all / MPI_Reduce(SUM) over ranks in rank order (tomo2d.f90:196-200) */
    double c = 0;
    for (int r = 0; r < nthreads; ++r) c = c + blk[r].chi;
    if (chi) *chi = c;
    if (istotable && tt_fwd) {
      for (int q = 0; q < npath; ++q) {
        double s = 0;
        for (int r = 0; r < nthreads; ++r) s = s + blk[r].tt_fwd[q];
        tt_fwd[q] = s;
      }
    }
    if (isadj) {
      for (size_t k = 0; k < nk; ++k) {
        double s = 0, sd = 0;
        for (int r = 0; r < nthreads; ++r) {
          s = s + blk[r].adj_s[k];
          if (dim == 3) sd = sd + blk[r].adj_den[k];
        }
        adj_s[k] = s;
        if (dim == 3) adj_density[k] = sd;
      }
    }
  }
  for (int r = 0; r < nthreads; ++r) {
    free(blk[r].tt_fwd); free(blk[r].adj_s); free(blk[r].adj_den);
  }
  free(blk); free(th);
  return st;
}

/* ======================================================================
 * Mixed-precision mode of the device (not a reference algorithm).
 *
 * Restates, in the device's exact operation order, the FP32 sweep phase and
 * the FP64 polish of the mixed mode, so that device results can be checked
 * bit for bit; accuracy against the FP64 restatement above is checked
 * separately, within the stated tolerances.
 *
 * Forward.  FP32 phase: the reference's update (:910-1079) with every operand
 * rounded to float, each sweep evaluating the candidates whose neighbours
 * are upwind in its direction, rewritten in the increment u = tau - tau_nb so that no
 * coefficient carries the (T0/dx)^2 scale: with tt = tX + u,
 *   Tx = gx + px*u,  gx = T0x*tX
 *   Ty = gy + py*u,  gy = T0y*tX + sy*(T0/dy)*(tX - tY)
 * so a*Tx^2 + b*Ty^2 - 2c*Tx*Ty = fun^2 becomes ea*u^2 + eb*u + ec = 0 with
 * coefficients of the size of the eikonal residual; roots by the
 * cancellation-free form q = -(eb + sign(eb)*sqrt(disc))/2, u = q/ea, ec/q.
 * One-sided candidates: u = (+-dis - tn*T0d)/p.  Causality tests, acceptance
 * rules, sweep order, tau init, frozen set and the stop rule (with the
 * residual |dtau|*T0 accumulated in double) are the reference's.  FP64
 * polish: the same upwind update in double on dtau = (double)dtau32, with
 * the period fields, T0x and T0y as the FP32 phase reads them (float,
 * widened) and T0 in FP64, and the same stop rule (at least one
 * iteration); T = (1 + dtau)*T0.
 *
 * Adjoint.  The face coefficients and d are computed in FP64 by the
 * reference's formulas (:1348-1370) on the metric rounded to float, with
 * products by the reciprocals of dx, dy, 4dx, 4dy, and the update
 * (:1392-1402) is taken in the form Ta = k0 + kW*TaW + kE*TaE + kS*TaS +
 * kN*TaN with k0 = delta/d, kW = a1p/dx/d, kE = -a2m/dx/d, kS = b1p/dy/d,
 * kN = -b2m/dy/d computed in double (as products with 1/d) and rounded to
 * float (all zero where |d| < eps).  FP32 phase: that
 * update in float, iterated until the reference's absolute rule holds or
 * the largest change is at most 2^-21 of the largest |Ta| (float
 * resolution).  FP64 polish: one iteration of the same update in double on
 * the same float coefficients, from Ta = (double)Ta32.
 *
 * The FP32 upwind cell and the FP32 adjoint update are written with
 * explicit fused multiply-adds (fmaf, correctly rounded), as the device
 * writes them; the FP64 polish evaluates the same expressions with separate
 * roundings (muladd), and everything else is compiled without contraction.
 * ====================================================================== */

static const float MIX_REL_STOP = 4.76837158203125e-07f; /* 2^-21 */

/* One cell of the upwind forward sweep of the mixed mode (see the note
 * above) in the sweep direction (xd, yd), in float (FP32 phase) or double
 * (FP64 polish): the two are one text, instantiated per type below.  Only
 * the candidates whose neighbours are upwind in this sweep are evaluated:
 * the quadrant of neighbours (i-xd, j) and (i, j-yd), and the one-sided
 * candidates from each of them.  Over the four sweeps of an iteration every
 * candidate of the reference's update is thus evaluated once, with its
 * neighbours' newest values; a fixed point is a fixed point of the full
 * update.  The field holds dtau = tau - 1: tau stays near 1 wherever T0 is a
 * good approximation, so the offset keeps 3-4 more bits of tau than tau
 * itself would in float.  The iterations follow; they return the iteration
 * count (FWD_MAXITER when the stop rule never held) and tell in *conv
 * whether it held. */
#define DEFINE_FWD_UP(SUF, V, SQRT, FMA, TOLSCALE)                                             \
static void fwd_cell_up_##SUF(int xd, int yd, int nx, int ny, int i, int j, V rdx,             \
                       V rdy, const V *a, const V *b, const V *c,                              \
                       const V *f, const V *disx, const V *disy,                               \
                       const V *T0, const V *T0x, const V *T0y,                                \
                       V *dtau) {                                                              \
  const size_t k = IDX(i, j);                                                                  \
  const V ak = a[k], bk = b[k], ck = c[k], fk = f[k];                                          \
  const V t0 = T0[k], t0x = T0x[k], t0y = T0y[k];                                              \
  const int hasX = xd > 0 ? i > 0 : i < nx - 1;                                                \
  const int hasY = yd > 0 ? j > 0 : j < ny - 1;                                                \
  const size_t kX = hasX ? IDX(i - xd, j) : k, kY = hasY ? IDX(i, j - yd) : k;                 \
  const V dX = dtau[kX], dY = dtau[kY], zX = T0[kX], zY = T0[kY];                              \
  /* px = T0x +- T0/dx, py = T0y +- T0/dy (upwind neighbour side) */                           \
  const V px = FMA(t0, xd > 0 ? rdx : -rdx, t0x);                                              \
  const V py = FMA(t0, yd > 0 ? rdy : -rdy, t0y);                                              \
  V best = dtau[k];                                                                            \
  if (hasX && hasY) {                                                                          \
    /* gx = T0x*tX, gy = T0y*tX +- (T0/dy)*(tX - tY) in offsets */                             \
    const V t0dy = t0 * rdy;                                                                   \
    const V gx = FMA(t0x, dX, t0x);                                                            \
    const V gy = FMA(yd > 0 ? t0dy : -t0dy, dX - dY, FMA(t0y, dX, t0y));                       \
    /* ea u^2 + eb u + ec = 0: ea = a px^2 + b py^2 - 2c px py,                                \
     * eb = 2(gx (a px - c py) + gy (b py - c px)),                                            \
     * ec = gx (a gx - 2c gy) + b gy^2 - f^2 */                                                \
    const V apx = ak * px, bpy = bk * py, cpx = ck * px, cpy = ck * py;                        \
    const V ea = FMA(apx, px, py * FMA((V)-2, cpx, bpy));                                      \
    const V eb = (V)2 * FMA(gx, apx - cpy, gy * (bpy - cpx));                                  \
    const V ec = FMA(gx, FMA(ak, gx, (V)-2 * (ck * gy)), FMA(bk * gy, gy, -(fk * fk)));        \
    const V disc = FMA(eb, eb, (V)-4 * ea * ec);                                               \
    if (disc >= 0) {                                                                           \
      const V sq = SQRT(disc);                                                                 \
      const V qq = (V)-0.5 * (eb >= 0 ? eb + sq : eb - sq);                                    \
      for (int isol = 1; isol <= 2; ++isol) {                                                  \
        const V u = isol == 1 ? qq / ea : ec / qq;                                             \
        const V dt = dX + u;                                                                   \
        const V Tx = FMA(px, u, gx), Ty = FMA(py, u, gy);                                      \
        const V chx = FMA(ak, Tx, -(ck * Ty)), chy = FMA(bk, Ty, -(ck * Tx));                  \
        const int ok = (xd > 0 ? chx >= 0 : chx <= 0) && (yd > 0 ? chy >= 0 : chy <= 0);       \
        if (ok && dt < best) best = dt;                                                        \
      }                                                                                        \
    }                                                                                          \
  }                                                                                            \
  /* one-sided candidates (:996-1075): tt*T0 >= tn*T0n and tt > tn*0.5 in                      \
   * offsets; p is the quadrant's px (x neighbour) or py (y neighbour) */                      \
  for (int side = 0; side < 2; ++side) {                                                       \
    if (side == 0 ? !hasX : !hasY) continue;                                                   \
    const V p = side == 0 ? px : py, dis = side == 0 ? disx[k] : disy[k];                      \
    const V t0d = side == 0 ? t0x : t0y;                                                       \
    const V dn = side == 0 ? dX : dY, zn = side == 0 ? zX : zY;                                \
    const V rp = (V)1 / p, base = FMA(t0d, dn, t0d);                                           \
    const V rhs = FMA(dn, zn, zn), half = FMA((V)0.5, dn, (V)-0.5);                            \
    V dt = FMA(dis - base, rp, dn);                                                            \
    if (FMA(dt, t0, t0) >= rhs && dt > half && dt < best) best = dt;                           \
    dt = FMA(-dis - base, rp, dn);                                                             \
    if (FMA(dt, t0, t0) >= rhs && dt > half && dt < best) best = dt;                           \
  }                                                                                            \
  dtau[k] = best;                                                                              \
}                                                                                              \
                                                                                               \
static int fwd_iterate_up_##SUF(int nx, int ny, V rdx, V rdy, const V *fld,                    \
                                const unsigned char *ischange, V *tau, V *old,                 \
                                int *conv) {                                                   \
  const size_t n = (size_t)nx * ny;                                                            \
  const V *a = fld, *b = fld + n, *c = fld + 2 * n, *f = fld + 3 * n,                          \
          *disx = fld + 4 * n, *disy = fld + 5 * n, *T0 = fld + 6 * n,                         \
          *T0x = fld + 7 * n, *T0y = fld + 8 * n;                                              \
  for (int iter = 1; iter <= FWD_MAXITER; ++iter) {                                            \
    memcpy(old, tau, n * sizeof(V));                                                           \
    for (int isw = 1; isw <= 4; ++isw) {                                                       \
      const int xd = isw <= 2 ? 1 : -1;                                                        \
      const int yd = (isw == 1 || isw == 3) ? 1 : -1;                                          \
      for (int ii = 0; ii < nx; ++ii) {                                                        \
        const int i = xd > 0 ? ii : nx - 1 - ii;                                               \
        for (int jj = 0; jj < ny; ++jj) {                                                      \
          const int j = yd > 0 ? jj : ny - 1 - jj;                                             \
          if (ischange[IDX(i, j)])                                                             \
            fwd_cell_up_##SUF(xd, yd, nx, ny, i, j, rdx, rdy, a, b, c, f, disx, disy, T0, T0x, \
                              T0y, tau);                                                       \
        }                                                                                      \
      }                                                                                        \
    }                                                                                          \
    double l1 = 0, linf = 0;                                                                   \
    for (int i = 0; i < nx; ++i)                                                               \
      for (int j = 0; j < ny; ++j) {                                                           \
        const size_t k = IDX(i, j);                                                            \
        const double d = fabs((double)tau[k] - (double)old[k]) * (double)T0[k];                \
        l1 = l1 + d;                                                                           \
        linf = linf > d ? linf : d;                                                            \
      }                                                                                        \
    l1 = l1 / (double)nx / (double)ny;                                                         \
    if (fabs(l1) < (TOLSCALE) * FWD_TOL && linf < (TOLSCALE) * FWD_TOL) {                      \
      *conv = 1;                                                                               \
      return iter;                                                                             \
    }                                                                                          \
  }                                                                                            \
  *conv = 0;                                                                                   \
  return FWD_MAXITER;                                                                          \
}                                                                                              \

/* a*b + c with two roundings, for the FP64 polish */
static double muladd(double a, double b, double c) { return a * b + c; }

/* The FP32 phase stops at 4 times the reference's tolerance (the device's
 * kMixFwdTolScale); the FP64 polish iterates to the reference's own. */
DEFINE_FWD_UP(f32, float, sqrtf, fmaf, 4.0)
DEFINE_FWD_UP(f64, double, sqrt, muladd, 1.0)

static int fwd_mixed_common(const double *xx_deg, const double *yy_deg, int nx,
                            int ny, const double *spha, const double *sphb,
                            const double *sphc, const double *fun, double x0_deg,
                            double y0_deg, const double *abc_in,
                            const double *T0_in, const double *T0x_in,
                            const double *T0y_in, double *T, int *iters) {
  if (nx < 2 || ny < 2) return ORC_E_ARG;
  const size_t n = (size_t)nx * ny;
  double *xx = malloc(sizeof(double) * nx), *yy = malloc(sizeof(double) * ny);
  double *buf = malloc(sizeof(double) * n * 8);
  float *f32 = malloc(sizeof(float) * n * 11);
  unsigned char *isc = malloc(n);
  if (!xx || !yy || !buf || !f32 || !isc) {
    free(xx); free(yy); free(buf); free(f32); free(isc);
    return ORC_E_NOMEM;
  }
  double *a = buf, *b = buf + n, *c = buf + 2 * n, *T0 = buf + 3 * n,
         *T0x = buf + 4 * n, *T0y = buf + 5 * n, *tau = buf + 6 * n,
         *tau_old = buf + 7 * n;
  float *tau32 = f32 + 9 * n, *old32 = f32 + 10 * n;
  for (int i = 0; i < nx; ++i) xx[i] = deg2rad_ref(xx_deg[i]);
  for (int j = 0; j < ny; ++j) yy[j] = deg2rad_ref(yy_deg[j]);
  const double x0 = deg2rad_ref(x0_deg), y0 = deg2rad_ref(y0_deg);
  const double dx = xx[1] - xx[0], dy = yy[1] - yy[0];
  if (abc_in) {
    memcpy(a, abc_in, n * sizeof(double));
    memcpy(b, abc_in + n, n * sizeof(double));
    memcpy(c, abc_in + 2 * n, n * sizeof(double));
  } else {
    metric_abc(nx, ny, yy, spha, sphb, sphc, a, b, c);
  }
  int st = ORC_OK;
  if (T0_in) {
    memcpy(T0, T0_in, n * sizeof(double));
    memcpy(T0x, T0x_in, n * sizeof(double));
    memcpy(T0y, T0y_in, n * sizeof(double));
  } else {
    double fun0;
    st = fwd_source(nx, ny, xx, yy, fun, x0, y0, dx, dy, &fun0);
    if (st == ORC_OK) fwd_t0_fields(nx, ny, xx, yy, x0, y0, fun0, T0, T0x, T0y);
  }
  if (st == ORC_OK) st = fwd_tau_init(nx, ny, xx, yy, x0, y0, dx, dy, tau, isc);
  int it32 = 0, it64 = 0;
  if (st == ORC_OK) {
    for (size_t k = 0; k < n; ++k) {
      const double ak = a[k], bk = b[k], ck = c[k], fk = fun[k];
      f32[k] = (float)ak;
      f32[n + k] = (float)bk;
      f32[2 * n + k] = (float)ck;
      f32[3 * n + k] = (float)fk;
      f32[4 * n + k] = (float)sqrt(fk * fk * (bk / (ak * bk - ck * ck)));
      f32[5 * n + k] = (float)sqrt(fk * fk * (ak / (ak * bk - ck * ck)));
      f32[6 * n + k] = (float)T0[k];
      f32[7 * n + k] = (float)T0x[k];
      f32[8 * n + k] = (float)T0y[k];
      tau32[k] = (float)(tau[k] - 1.0);
    }
    int c32 = 0, c64 = 0;
    it32 = fwd_iterate_up_f32(nx, ny, (float)(1.0 / dx), (float)(1.0 / dy), f32, isc, tau32,
                              old32, &c32);
    /* polish on dtau in double: the period fields and T0x, T0y as the FP32
     * phase has them (rounded to float once, widened), T0 in FP64 */
    double *f64 = malloc(sizeof(double) * n * 9);
    if (!f64) {
      st = ORC_E_NOMEM;
    } else {
      for (size_t k = 0; k < n; ++k) {
        for (int q = 0; q < 6; ++q) f64[q * n + k] = (double)f32[q * n + k];
        f64[6 * n + k] = T0[k];
        f64[7 * n + k] = (double)f32[7 * n + k];
        f64[8 * n + k] = (double)f32[8 * n + k];
        tau[k] = (double)tau32[k];
      }
      it64 = fwd_iterate_up_f64(nx, ny, 1.0 / dx, 1.0 / dy, f64, isc, tau, tau_old, &c64);
      for (size_t k = 0; k < n; ++k) T[k] = (1.0 + tau[k]) * T0[k];
      st = c64 ? ORC_OK : ORC_W_MAXITER;
      free(f64);
    }
  }
  if (iters) { iters[0] = it32; iters[1] = it64; }
  free(xx); free(yy); free(buf); free(f32); free(isc);
  return st;
}

int orc_fwd_mixed(const double *xx_deg, const double *yy_deg, int nx, int ny,
                  const double *spha, const double *sphb, const double *sphc,
                  const double *fun, double x0_deg, double y0_deg, double *T,
                  int *iters) {
  return fwd_mixed_common(xx_deg, yy_deg, nx, ny, spha, sphb, sphc, fun, x0_deg,
                          y0_deg, NULL, NULL, NULL, NULL, T, iters);
}

int orc_fwd_mixed_pre(const double *xx_deg, const double *yy_deg, int nx, int ny,
                      const double *a, const double *b, const double *c,
                      const double *fun, double x0_deg, double y0_deg,
                      const double *T0, const double *T0x, const double *T0y,
                      double *T, int *iters) {
  const size_t n = (size_t)nx * ny;
  double *abc = malloc(sizeof(double) * 3 * n);
  if (!abc) return ORC_E_NOMEM;
  memcpy(abc, a, n * sizeof(double));
  memcpy(abc + n, b, n * sizeof(double));
  memcpy(abc + 2 * n, c, n * sizeof(double));
  const int st = fwd_mixed_common(xx_deg, yy_deg, nx, ny, NULL, NULL, NULL, fun, x0_deg,
                                  y0_deg, abc, T0, T0x, T0y, T, iters);
  free(abc);
  return st;
}

/* The mixed adjoint's float coefficients k0, kW, kE, kS, kN (n each, in
 * kk; see the note above): the face coefficients of :1348-1370 on the
 * metric a, b, c rounded to float (the float period fields the device
 * reads), with the divisions by dx, dy, 4dx, 4dy taken as products with
 * their reciprocals and the halvings as products with 0.5, then
 * kW = a1p*rdx*(1/d) and so on, in the device's order. */
static void adj_kform(int nx, int ny, double dx, double dy, const double *T,
                      const double *a, const double *b, const double *c,
                      const double *delta, float *kk) {
  const size_t n = (size_t)nx * ny;
  float *k0 = kk, *kW = kk + n, *kE = kk + 2 * n, *kS = kk + 3 * n, *kN = kk + 4 * n;
  for (size_t k = 0; k < n; ++k) k0[k] = kW[k] = kE[k] = kS[k] = kN[k] = 0;
  const double rdx = 1.0 / dx, rdy = 1.0 / dy, rdx4 = 0.25 * rdx, rdy4 = 0.25 * rdy;
#define F(p, ii, jj) ((double)(float)p[IDX(ii, jj)])
  for (int i = 1; i < nx - 1; ++i)
    for (int j = 1; j < ny - 1; ++j) {
      const size_t k = IDX(i, j);
      const double Tk = T[k], TW = T[IDX(i - 1, j)], TE = T[IDX(i + 1, j)];
      const double TS = T[IDX(i, j - 1)], TN = T[IDX(i, j + 1)];
      const double TWS = T[IDX(i - 1, j - 1)], TWN = T[IDX(i - 1, j + 1)];
      const double TES = T[IDX(i + 1, j - 1)], TEN = T[IDX(i + 1, j + 1)];
      const double ak = F(a, i, j), aW = F(a, i - 1, j), aE = F(a, i + 1, j);
      const double bk = F(b, i, j), bS = F(b, i, j - 1), bN = F(b, i, j + 1);
      const double ck = F(c, i, j), cW = F(c, i - 1, j), cE = F(c, i + 1, j);
      const double cS = F(c, i, j - 1), cN = F(c, i, j + 1);
      const double a1 = -((Tk - TW) * rdx * (ak + aW) * 0.5) + (TN - TS + TWN - TWS) * rdy4 * (ck + cW) * 0.5;
      const double a2 = -((TE - Tk) * rdx * (aE + ak) * 0.5) + (TEN - TES + TN - TS) * rdy4 * (cE + ck) * 0.5;
      const double b1 = -((Tk - TS) * rdy * (bk + bS) * 0.5) + (TE - TW + TES - TWS) * rdx4 * (ck + cS) * 0.5;
      const double b2 = -((TN - Tk) * rdy * (bN + bk) * 0.5) + (TEN - TWN + TE - TW) * rdx4 * (cN + ck) * 0.5;
      const double a1p = (a1 + fabs(a1)) * 0.5, a1m = (a1 - fabs(a1)) * 0.5;
      const double a2p = (a2 + fabs(a2)) * 0.5, a2m = (a2 - fabs(a2)) * 0.5;
      const double b1p = (b1 + fabs(b1)) * 0.5, b1m = (b1 - fabs(b1)) * 0.5;
      const double b2p = (b2 + fabs(b2)) * 0.5, b2m = (b2 - fabs(b2)) * 0.5;
      const double d = (a2p - a1m) * rdx + (b2p - b1m) * rdy;
      if (fabs(d) < ADJ_EPS) continue;
      const double rd = 1.0 / d;
      k0[k] = (float)(delta[k] * rd);
      kW[k] = (float)(a1p * rdx * rd);
      kE[k] = (float)(-a2m * rdx * rd);
      kS[k] = (float)(b1p * rdy * rd);
      kN[k] = (float)(-b2m * rdy * rd);
    }
#undef F
}

/* FP32 phase of the mixed adjoint (see the note above) on the float
 * coefficients kk.  Ta holds the reference's initial field on entry and the
 * FP32 result, widened, on exit.  Returns the iteration count or a negative
 * status; *done tells whether the stop rule held. */
static int adj_phase32(int nx, int ny, const float *kk, double *Ta, int *done_out) {
  const size_t n = (size_t)nx * ny;
  float *tt = malloc(sizeof(float) * n * 2);
  if (!tt) return ORC_E_NOMEM;
  const float *k0 = kk, *kW = kk + n, *kE = kk + 2 * n, *kS = kk + 3 * n, *kN = kk + 4 * n;
  float *ta = tt, *old = tt + n;
  for (size_t k = 0; k < n; ++k) ta[k] = (float)Ta[k];
  int iter, done = 0;
  for (iter = 1; iter <= ADJ_MAXITER; ++iter) {
    memcpy(old, ta, n * sizeof(float));
    for (int xdir = -1; xdir <= 1; xdir += 2)
      for (int ydir = -1; ydir <= 1; ydir += 2)
        for (int ii = 1; ii < nx - 1; ++ii) {
          const int i = xdir < 0 ? ii : nx - 1 - ii;
          for (int jj = 1; jj < ny - 1; ++jj) {
            const int j = ydir < 0 ? jj : ny - 1 - jj;
            const size_t k = IDX(i, j);
            ta[k] = fmaf(kN[k], ta[IDX(i, j + 1)],
                         fmaf(kS[k], ta[IDX(i, j - 1)],
                              fmaf(kE[k], ta[IDX(i + 1, j)], fmaf(kW[k], ta[IDX(i - 1, j)], k0[k]))));
          }
        }
    double l1 = 0, linf = 0, amax = 0;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        const size_t k = IDX(i, j);
        const double dd = fabs((double)ta[k] - (double)old[k]);
        const double av = fabs((double)ta[k]);
        l1 = l1 + dd;
        linf = linf > dd ? linf : dd;
        amax = amax > av ? amax : av;
      }
    l1 = l1 / (double)(nx * ny);
    if ((fabs(l1) < ADJ_TOL && linf < ADJ_TOL) || linf <= (double)MIX_REL_STOP * amax) {
      done = 1;
      break;
    }
  }
  if (!done) iter = ADJ_MAXITER;
  for (size_t k = 0; k < n; ++k) Ta[k] = (double)ta[k];
  free(tt);
  *done_out = done;
  return iter;
}

int orc_adj_mixed(const double *xx_deg, const double *yy_deg, int nx, int ny,
                  const double *spha, const double *sphb, const double *sphc,
                  const double *T, const double *xrec_deg, const double *yrec_deg,
                  const double *src_adj, int nr, double *Ta, int *iters) {
  return adj_common(xx_deg, yy_deg, nx, ny, spha, sphb, sphc, T, xrec_deg,
                    yrec_deg, src_adj, nr, NULL, 1, Ta, iters);
}

int orc_adj_mixed_pre(const double *xx_deg, const double *yy_deg, int nx, int ny,
                      const double *a, const double *b, const double *c,
                      const double *T, const double *delta, double *Ta, int *iters) {
  const size_t n = (size_t)nx * ny;
  double *pre = malloc(sizeof(double) * 4 * n);
  if (!pre) return ORC_E_NOMEM;
  memcpy(pre, a, n * sizeof(double));
  memcpy(pre + n, b, n * sizeof(double));
  memcpy(pre + 2 * n, c, n * sizeof(double));
  memcpy(pre + 3 * n, delta, n * sizeof(double));
  const int st = adj_common(xx_deg, yy_deg, nx, ny, NULL, NULL, NULL, T, NULL,
                            NULL, NULL, 0, pre, 1, Ta, iters);
  free(pre);
  return st;
}
