/*
 * CPU restatement of SurfATT's forward + adjoint hot path ("the oracle").
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libsatt.so) never links or calls it.
 *
 * Every function follows the reference's statement order and floating-point
 * evaluation order (Fortran left-to-right, no FMA contraction; compiled with
 * -ffp-contract=off), including the reference's single-precision quirks:
 *   - pi is a default-real literal widened to double
 *     (src/eikon_solver_2d.f90:785,1297);
 *   - the bilinear weights r1, r2 are implicitly REAL(4)
 *     (src/eikon_solver_2d.f90:828-841,1329-1335);
 *   - the tolerances are REAL(4) constant expressions
 *     (src/eikon_solver_2d.f90:775,1294).
 *
 * Array layout follows the reference: a field f(nx,ny) is column-major with
 * x fastest, f[i + nx*j]; period-major model arrays svel(nperiod,nx,ny) are
 * v[p + nperiod*(i + nx*j)].  All indices in this API are 0-based.
 */
#ifndef SATT_ORACLE_H
#define SATT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (same values as include/satt.h) */
#define ORC_OK 0
#define ORC_W_MAXITER 1
#define ORC_E_ARG -1
#define ORC_E_SRC_ON_BOUNDARY -2
#define ORC_E_REC_OUT_OF_GRID -3
#define ORC_E_NOMEM -4

/* src/eikon_solver_2d.f90:764-1126  FSM_UW_PS_lonlat_2d */
int orc_fwd_lonlat(const double *xx_deg, const double *yy_deg, int nx, int ny,
                   const double *spha, const double *sphb, const double *sphc,
                   const double *fun, double x0_deg, double y0_deg, double *T,
                   int *iters);

/* Same solver with the libm-dependent setup supplied by the caller: the
 * converted metric a, b, c (src/eikon_solver_2d.f90:817-821) and T0, T0x,
 * T0y (:845-857).  Used to compare a device solve bit-for-bit when the
 * device's libm differs from the host's in the last bits.  tau init and the
 * frozen set are still derived as :859-872. */
int orc_fwd_lonlat_pre(const double *xx_deg, const double *yy_deg, int nx,
                       int ny, const double *a, const double *b,
                       const double *c, const double *fun, double x0_deg,
                       double y0_deg, const double *T0, const double *T0x,
                       const double *T0y, double *T, int *iters);

/* The analytic factor alone (src/eikon_solver_2d.f90:801-857). */
int orc_fwd_t0(const double *xx_deg, const double *yy_deg, int nx, int ny,
               const double *fun, double x0_deg, double y0_deg, double *T0,
               double *T0x, double *T0y);

/* src/eikon_solver_2d.f90:1276-1446  FSM_O1_JSE_lonlat_2d */
int orc_adj_lonlat(const double *xx_deg, const double *yy_deg, int nx, int ny,
                   const double *spha, const double *sphb, const double *sphc,
                   const double *T, const double *xrec_deg,
                   const double *yrec_deg, const double *src_adj, int nr,
                   double *Ta, int *iters);

/* Adjoint with a, b, c (:1338-1344) and the delta source (:1320-1336)
 * supplied by the caller (same purpose as orc_fwd_lonlat_pre). */
int orc_adj_lonlat_pre(const double *xx_deg, const double *yy_deg, int nx,
                       int ny, const double *a, const double *b,
                       const double *c, const double *T, const double *delta,
                       double *Ta, int *iters);

/* The device's mixed-precision mode, restated (not a reference algorithm):
 * an FP32 sweep phase followed by FP64 polish iterations with the
 * reference's update and stop rule (satt_oracle.c describes both phases).
 * iters[0] = FP32 iterations, iters[1] = FP64 polish iterations. */
int orc_fwd_mixed(const double *xx_deg, const double *yy_deg, int nx, int ny,
                  const double *spha, const double *sphb, const double *sphc,
                  const double *fun, double x0_deg, double y0_deg, double *T,
                  int *iters);
int orc_fwd_mixed_pre(const double *xx_deg, const double *yy_deg, int nx, int ny,
                      const double *a, const double *b, const double *c,
                      const double *fun, double x0_deg, double y0_deg,
                      const double *T0, const double *T0x, const double *T0y,
                      double *T, int *iters);
int orc_adj_mixed(const double *xx_deg, const double *yy_deg, int nx, int ny,
                  const double *spha, const double *sphb, const double *sphc,
                  const double *T, const double *xrec_deg, const double *yrec_deg,
                  const double *src_adj, int nr, double *Ta, int *iters);
int orc_adj_mixed_pre(const double *xx_deg, const double *yy_deg, int nx, int ny,
                      const double *a, const double *b, const double *c,
                      const double *T, const double *delta, double *Ta, int *iters);

/* src/eikon_solver_2d.f90:1449-1467  mask */
void orc_mask(const double *xx, const double *yy, int nx, int ny,
              double *table, double x0, double y0);

/* src/shared/utils.f90:917-956  interp2_1_dp */
int orc_interp2(const double *x, const double *y, int nx, int ny,
                const double *v, const double *xq, const double *yq, int nq,
                double *vq);

/* The per-(source, period) loop: src/tomo2d.f90:158-201 (dim==2) and
 * src/acqui.f90:91-161 (dim==3), with att_measadj (src/measadj.f90:62-114)
 * and post_proc_eikokernel (src/tomo2d.f90:145-156, src/acqui.f90:307-316).
 *
 * items i = 0..nitems-1 in isrcs order; item_pidx[i] is the 0-based period
 * index; receivers of item i are rec_off[i]..rec_off[i+1]-1.
 * Outputs (any may be NULL when the matching flag is off):
 *   tt_fwd[npath]            overwritten (zeros where no receiver maps)
 *   adj_s[nperiod*nx*ny]      period-major, overwritten
 *   adj_density[...]          dim==3 only
 *   item_chi[nitems], iters[3*nitems] (fwd, adj, density-adj)
 * nthreads > 1 splits the item list into the reference's static contiguous
 * blocks (src/shared/parallel.f90:1877-1903) and sums the per-block kernels
 * in block order, as MPI_Reduce over ranks would.
 */
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
                         double *weight_out);

/* ---- model-side steps (satt_oracle_geo.c) */
/* gps2dist (src/shared/utils.f90:1250-1294), km */
double orc_gps2dist(double lat0, double lon0, double lat1, double lon1);
/* window half-widths of gaussian_smooth_geo_2 (src/shared/utils.f90:1204-1212) */
int orc_smooth_halfwidth_x(const double *lons, double sigma);
int orc_smooth_halfwidth_y(const double *lats, int j, double sigma);
/* gaussian_smooth_geo_2 (src/shared/utils.f90:1194-1223), one field, rows
 * j0..j1-1 (the others are left as they are) */
int orc_smooth_geo2_rows(const double *lons, const double *lats, int nx, int ny,
                         const double *data, double sigma, int j0, int j1, double *out);
int orc_smooth_geo2(const double *lons, const double *lats, int nx, int ny, const double *data,
                    double sigma, double *out);
/* gradient_2_geo (src/shared/utils.f90:1118-1160) */
int orc_gradient_geo2(const double *f, const double *lon, const double *lat, int nx, int ny,
                      double *tx, double *ty);
/* src/grid.f90:116 */
double orc_topo_sigma(double vel, double period, double wavelen_factor);
/* get_topo's metric and dip angle (src/grid.f90:114-123,137-141;
 * src/topo.f90:123-130), outputs period-major; angle may be NULL */
int orc_topo_metric(const double *lons, const double *lats, int nx, int ny, const double *z,
                    int nperiod, const double *sigma, double *m11, double *m12, double *m22,
                    double *angle);
/* optimize + steepest_descent (src/tomo2d.f90:203-251), period-major */
int orc_steepest_descent(const double *lons, const double *lats, int nx, int ny, int nperiod,
                         const double *adj_s, double sigma_2d, double updatemax, double *svel,
                         double *ker_s);

/* ---- 3-D kernel steps (satt_oracle_ker3d.c) */
/* src/shared/utils.f90:99-145 */
double orc_drho_dalpha(double vp);
double orc_dalpha_dbeta(double vs);
/* combine_kernels (src/acqui.f90:197-245), accumulating into ker/ker_den
 * (nker = 3 with use_alpha_beta_rho, else 1) */
void orc_combine_kernels(int nperiod, int nx, int ny, int nz, const double *adj_s,
                         const double *adj_den, const double *sen_vs, const double *sen_vp,
                         const double *sen_rho, const double *vs3d, const double *vp3d,
                         int use_alpha_beta_rho, int rho_scaling, double *ker,
                         double *ker_den);
/* src/eikon_solver_2d.f90:1470-1691; -1 when a point is outside the grid */
int orc_locate_within_cube(double xx, double yy, double zz, int nx, int ny, int nz,
                           const double *xv, const double *yv, const double *zv, int *i1,
                           int *j1, int *k1, double *wt);
int orc_inv_grid_iso(const double *xinv, const double *yinv, const double *zinv, int nxinv,
                     int nyinv, int nzinv, int nset, const double *adj, int nx, int ny, int nz,
                     double *gk, const double *vx, const double *vy, const double *vz);
int orc_inv2fwd_iso(const double *xinv, const double *yinv, const double *zinv, int nxinv,
                    int nyinv, int nzinv, int nset, int nx, int ny, int nz, const double *gk,
                    const double *vx, const double *vy, const double *vz, double *update);
/* interp1_0 (src/shared/utils.f90:786-805) */
double orc_interp1(const double *x, const double *v, int n, double xq);
/* correct_depth (src/surfker/surfker.f90:260-283); work: 2*nz doubles */
void orc_correct_depth(int nperiod, int nx, int ny, int nz, double *ker_vs, double *ker_vp,
                       double *ker_rho, const double *dip, const double *depz, double *work);
/* post_processing_for_kernel + regularize_ker_density (src/acqui.f90:247-304);
 * work: nset*nxinv*nyinv*nzinv + 2*nx*ny*nz doubles */
int orc_kernel_postproc(int nker, int nx, int ny, int nz, const double *vx, const double *vy,
                        const double *vz, int nset, int nxinv, int nyinv, int nzinv,
                        const double *xinv, const double *yinv, const double *zinv, double *ker,
                        const double *ker_den, double kdensity_coe, double chi0, double weight,
                        double *gradient_s, double *work);

#ifdef __cplusplus
}
#endif
#endif
