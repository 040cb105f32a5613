/*
 * Oracle (test infrastructure only; never linked into the product): a CPU
 * restatement of the 3-D kernel steps downstream of the stacked 2-D kernels
 * in the reference's att_acqui:
 *
 *   orc_combine_kernels     combine_kernels        src/acqui.f90:197-245
 *   orc_dalpha_dbeta,
 *   orc_drho_dalpha         dalpha_dbeta_*, drho_dalpha_0
 *                                                  src/shared/utils.f90:99-145
 *   orc_locate_within_cube  locate_within_cube     src/eikon_solver_2d.f90:1470-1535
 *   orc_inv_grid_iso        inv_grid_iso           src/eikon_solver_2d.f90:1538-1610
 *   orc_inv2fwd_iso         inv2fwd_iso            src/eikon_solver_2d.f90:1613-1691
 *   orc_kernel_postproc     post_processing_for_kernel + regularize_ker_density
 *                                                  src/acqui.f90:247-304
 *   orc_interp1             interp1_0              src/shared/utils.f90:786-805
 *   orc_correct_depth       correct_depth          src/surfker/surfker.f90:260-283
 *
 * Statement order and evaluation order follow the reference (built with
 * -ffp-contract=off).  Array layouts are the reference's Fortran ones:
 * adj_s(nperiod,nx,ny), sen_*(nperiod,nx,ny,nz), ker(nker,nx,ny,nz),
 * vs3d(nx,ny,nz), xinv(nxinv,nset), gk(nset*nxinv*nyinv*nzinv).
 */
#include <math.h>
#include <stddef.h>

#include "satt_oracle.h"

/* shared_par.f90:16,23 */
static const double RHO_SCALING = 0.33;
static const double PRECOND_THRES = 1.0e-4;

/* The REAL(4) literals of utils.f90:99-145 (e.g. 2*0.4721 is a default-real
 * product), widened when multiplied by the REAL(8) argument; integer powers
 * as gfortran expands them (x**3 = x*x*x, x**4 = (x*x)*(x*x)). */
double orc_drho_dalpha(double vp) {
  const double c0 = (double)1.6612f, c1 = (double)(2.0f * 0.4721f);
  const double c2 = (double)(3.0f * 0.0671f), c3 = (double)(4.0f * 0.0043f);
  const double c4 = (double)(5.0f * 0.000106f);
  const double v2 = vp * vp, v3 = v2 * vp, v4 = v2 * v2;
  return c0 - c1 * vp + c2 * v2 - c3 * v3 + c4 * v4;
}

double orc_dalpha_dbeta(double vs) {
  const double c0 = (double)2.0947f, c1 = (double)(2.0f * 0.8206f);
  const double c2 = (double)(3.0f * 0.2683f), c3 = (double)(4.0f * 0.0251f);
  const double v2 = vs * vs, v3 = v2 * vs;
  return c0 - c1 * vs + c2 * v2 - c3 * v3;
}

/* combine_kernels (acqui.f90:197-245), rank 0's part: ker and ker_density
 * (nker = 3 with use_alpha_beta_rho, else 1) are accumulated into, as the
 * reference does after prepare_inv zeroed them. */
void orc_combine_kernels(int nperiod, int nx, int ny, int nz, const double *adj_s,
                         const double *adj_den, const double *sen_vs, const double *sen_vp,
                         const double *sen_rho, const double *vs3d, const double *vp3d,
                         int use_alpha_beta_rho, int rho_scaling, double *ker,
                         double *ker_den) {
  const size_t N2 = (size_t)nx * ny, N3 = N2 * nz;
  if (use_alpha_beta_rho) {
    for (int ip = 0; ip < nperiod; ++ip)
      for (int k = 0; k < nz; ++k)
        for (size_t c2 = 0; c2 < N2; ++c2) {
          const size_t c3 = c2 + N2 * k;
          const double a = adj_s[ip + (size_t)nperiod * c2];
          const double ad = adj_den[ip + (size_t)nperiod * c2];
          const size_t s = ip + (size_t)nperiod * c3;
          ker[0 + 3 * c3] = ker[0 + 3 * c3] - a * sen_vs[s];
          ker_den[0 + 3 * c3] = ker_den[0 + 3 * c3] - ad * sen_vs[s];
          ker[1 + 3 * c3] = ker[1 + 3 * c3] - a * sen_vp[s];
          ker_den[1 + 3 * c3] = ker_den[1 + 3 * c3] - ad * sen_vp[s];
          if (!rho_scaling) {
            ker[2 + 3 * c3] = ker[2 + 3 * c3] - a * sen_rho[s];
            ker_den[2 + 3 * c3] = ker_den[2 + 3 * c3] - ad * sen_rho[s];
          }
        }
    if (rho_scaling)
      for (size_t c3 = 0; c3 < N3; ++c3) {
        ker[2 + 3 * c3] = RHO_SCALING * ker[0 + 3 * c3];
        ker_den[2 + 3 * c3] = ker_den[0 + 3 * c3];
      }
  } else {
    for (int ip = 0; ip < nperiod; ++ip)
      for (int k = 0; k < nz; ++k)
        for (size_t c2 = 0; c2 < N2; ++c2) {
          const size_t c3 = c2 + N2 * k;
          const double a = adj_s[ip + (size_t)nperiod * c2];
          const double ad = adj_den[ip + (size_t)nperiod * c2];
          const size_t s = ip + (size_t)nperiod * c3;
          const double dab = orc_dalpha_dbeta(vs3d[c3]);
          const double dra = orc_drho_dalpha(vp3d[c3]);
          ker[c3] = ker[c3] - a * sen_vs[s] - a * sen_vp[s] * dab - a * sen_rho[s] * dra * dab;
          ker_den[c3] = ker_den[c3] - ad * sen_vs[s] - ad * sen_vp[s] * dab -
                        ad * sen_rho[s] * dra * dab;
        }
  }
}

/* locate_within_cube (eikon_solver_2d.f90:1470-1535): the walk from the
 * cube (*i1,*j1,*k1) (1-based, updated in place); returns 1 and the eight
 * weights when the point is inside, else 0. */
int orc_locate_within_cube(double xx, double yy, double zz, int nx, int ny, int nz,
                           const double *xv, const double *yv, const double *zv, int *i1,
                           int *j1, int *k1, double *wt) {
  double wx = 0, wy = 0, wz = 0;
  int fx = 0, fy = 0, fz = 0;
  while (!fx && *i1 > 0 && *i1 < nx) {
    if (xx < xv[*i1 - 1]) {
      *i1 -= 1;
    } else if (xx <= xv[*i1]) {
      fx = 1;
      wx = (xx - xv[*i1 - 1]) / (xv[*i1] - xv[*i1 - 1]);
    } else {
      *i1 += 1;
    }
  }
  while (!fy && *j1 > 0 && *j1 < ny) {
    if (yy < yv[*j1 - 1]) {
      *j1 -= 1;
    } else if (yy <= yv[*j1]) {
      fy = 1;
      wy = (yy - yv[*j1 - 1]) / (yv[*j1] - yv[*j1 - 1]);
    } else {
      *j1 += 1;
    }
  }
  while (!fz && *k1 > 0 && *k1 < nz) {
    if (zz < zv[*k1 - 1]) {
      *k1 -= 1;
    } else if (zz <= zv[*k1]) {
      fz = 1;
      wz = (zz - zv[*k1 - 1]) / (zv[*k1] - zv[*k1 - 1]);
    } else {
      *k1 += 1;
    }
  }
  if (!(fx && fy && fz)) return 0;
  wt[0] = (1.0 - wx) * (1.0 - wy) * (1.0 - wz);
  wt[1] = (1.0 - wx) * wy * (1.0 - wz);
  wt[2] = wx * wy * (1.0 - wz);
  wt[3] = wx * (1.0 - wy) * (1.0 - wz);
  wt[4] = (1.0 - wx) * (1.0 - wy) * wz;
  wt[5] = (1.0 - wx) * wy * wz;
  wt[6] = wx * wy * wz;
  wt[7] = wx * (1.0 - wy) * wz;
  return 1;
}

/* 0-based offset of corner n (0..7) of cube (i1,j1,k1) (1-based) in one
 * set's block (eikon_solver_2d.f90:1585-1602 less one). */
static size_t corner(int n, int i1, int j1, int k1, int nxi, int nyi) {
  static const int dx[8] = {0, 0, 1, 1, 0, 0, 1, 1};
  static const int dy[8] = {0, 1, 1, 0, 0, 1, 1, 0};
  static const int dz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  return (size_t)nxi * nyi * (k1 - 1 + dz[n]) + (size_t)nxi * (j1 - 1 + dy[n]) +
         (size_t)(i1 - 1 + dx[n]);
}

int orc_inv_grid_iso(const double *xinv, const double *yinv, const double *zinv, int nxinv,
                     int nyinv, int nzinv, int nset, const double *adj, int nx, int ny, int nz,
                     double *gk, const double *vx, const double *vy, const double *vz) {
  const double dx = vx[1] - vx[0], dy = vy[1] - vy[0], dz = vz[1] - vz[0];
  const size_t nb = (size_t)nxinv * nyinv * nzinv;
  for (size_t m = 0; m < nb * nset; ++m) gk[m] = 0;
  for (int s = 0; s < nset; ++s) {
    const double *xi = xinv + (size_t)nxinv * s, *yi = yinv + (size_t)nyinv * s;
    const double *zi = zinv + (size_t)nzinv * s;
    double *g = gk + nb * s;
    int i1 = 1, j1 = 1, k1 = 1;
    for (int j = 1; j < ny - 1; ++j)
      for (int i = 1; i < nx - 1; ++i)
        for (int k = 1; k < nz - 1; ++k) {
          double wt[8];
          if (!orc_locate_within_cube(vx[i], vy[j], vz[k], nxinv, nyinv, nzinv, xi, yi, zi, &i1,
                                      &j1, &k1, wt))
            return -1;
          const double a = adj[i + (size_t)nx * (j + (size_t)ny * k)];
          for (int n = 0; n < 8; ++n) {
            const size_t m = corner(n, i1, j1, k1, nxinv, nyinv);
            g[m] = g[m] + a * dx * dy * dz * wt[n];
          }
        }
  }
  return 0;
}

int orc_inv2fwd_iso(const double *xinv, const double *yinv, const double *zinv, int nxinv,
                    int nyinv, int nzinv, int nset, int nx, int ny, int nz, const double *gk,
                    const double *vx, const double *vy, const double *vz, double *update) {
  const size_t nb = (size_t)nxinv * nyinv * nzinv, N = (size_t)nx * ny * nz;
  for (size_t c = 0; c < N; ++c) update[c] = 0;
  for (int s = 0; s < nset; ++s) {
    const double *xi = xinv + (size_t)nxinv * s, *yi = yinv + (size_t)nyinv * s;
    const double *zi = zinv + (size_t)nzinv * s;
    const double *g = gk + nb * s;
    int i1 = 1, j1 = 1, k1 = 1;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nz; ++k) {
          double wt[8];
          if (!orc_locate_within_cube(vx[i], vy[j], vz[k], nxinv, nyinv, nzinv, xi, yi, zi, &i1,
                                      &j1, &k1, wt))
            return -1;
          double val = 0.0;
          for (int n = 0; n < 8; ++n) val = val + wt[n] * g[corner(n, i1, j1, k1, nxinv, nyinv)];
          const size_t c = i + (size_t)nx * (j + (size_t)ny * k);
          update[c] = update[c] + val;
        }
  }
  return 0;
}

/* post_processing_for_kernel (acqui.f90:247-285) with regularize_ker_density
 * (:287-304), rank 0's part: for each kernel component iker, the optional
 * density preconditioning of ker (in place), the projection to the
 * inversion grids and back, and gradient_s += update/nset/chi0*weight.
 * work holds nset*nxinv*nyinv*nzinv + 2*nx*ny*nz doubles. */
int orc_kernel_postproc(int nker, int nx, int ny, int nz, const double *vx, const double *vy,
                        const double *vz, int nset, int nxinv, int nyinv, int nzinv,
                        const double *xinv, const double *yinv, const double *zinv, double *ker,
                        const double *ker_den, double kdensity_coe, double chi0, double weight,
                        double *gradient_s, double *work) {
  const size_t N = (size_t)nx * ny * nz, nb = (size_t)nxinv * nyinv * nzinv * nset;
  double *gk = work, *one = work + nb, *update = one + N;
  for (int ik = 0; ik < nker; ++ik) {
    if (kdensity_coe > 0) {
      double mx = 0;
      for (size_t c = 0; c < N; ++c) {
        const double a = fabs(ker_den[ik + (size_t)nker * c]);
        if (a > mx) mx = a;
      }
      for (size_t c = 0; c < N; ++c) {
        double p = fabs(ker_den[ik + (size_t)nker * c]) / mx;
        p = p < PRECOND_THRES ? 0.0 : 1 / pow(p, kdensity_coe);
        ker[ik + (size_t)nker * c] = ker[ik + (size_t)nker * c] * p;
      }
    }
    for (size_t c = 0; c < N; ++c) one[c] = ker[ik + (size_t)nker * c];
    if (orc_inv_grid_iso(xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, one, nx, ny, nz, gk, vx,
                         vy, vz) != 0)
      return -1;
    if (orc_inv2fwd_iso(xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, nx, ny, nz, gk, vx, vy, vz,
                        update) != 0)
      return -1;
    for (size_t c = 0; c < N; ++c) {
      const double u = update[c] / (double)nset / chi0;
      gradient_s[ik + (size_t)nker * c] = gradient_s[ik + (size_t)nker * c] + u * weight;
    }
  }
  return 0;
}

/* interp1_0 (utils.f90:786-805): x ascending; x1 the last node <= xq, x2 the
 * first node > xq; past the last node the last value. */
double orc_interp1(const double *x, const double *v, int n, double xq) {
  int x1 = -1, x2 = -1;
  for (int i = 0; i < n; ++i) {
    if (xq >= x[i]) x1 = i;
    else if (x2 < 0) x2 = i;
  }
  if (x2 < 0) return v[n - 1];
  if (x1 < 0) return NAN;  /* below the first node: the reference reads x(0) */
  const double vn = fabs(x[x2] - x[x1]);
  const double vq = v[x1] * (x[x2] - xq) + v[x2] * (xq - x[x1]);
  return vq / vn;
}

/* correct_depth (surfker.f90:260-283): every (period, column) of the depth
 * kernels (nperiod,nx,ny,nz) resampled at depz/cosd(dip) from the dip
 * angles (nperiod,nx,ny); cosd(x) = cos(x*pi/180) with the true pi.
 * work: 2*nz doubles. */
void orc_correct_depth(int nperiod, int nx, int ny, int nz, double *ker_vs, double *ker_vp,
                       double *ker_rho, const double *dip, const double *depz, double *work) {
  static const double PI_TRUE = 3.14159265358979323846;
  double *col = work, *newz = work + nz;
  double *ks[3] = {ker_vs, ker_vp, ker_rho};
  const size_t n3 = (size_t)nperiod * nx * ny;
  for (size_t c = 0; c < n3; ++c) {
    const double cd = cos(dip[c] * (PI_TRUE / 180.0));
    for (int k = 0; k < nz; ++k) newz[k] = depz[k] / cd;
    for (int q = 0; q < 3; ++q) {
      for (int k = 0; k < nz; ++k) col[k] = ks[q][c + n3 * k];
      for (int k = 0; k < nz; ++k) ks[q][c + n3 * k] = orc_interp1(depz, col, nz, newz[k]);
    }
  }
}
