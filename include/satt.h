/*
 * satt.h -- C-ABI of the MI355X-native SurfATT forward + adjoint hot path.
 *
 * libsatt.so replaces, behind a plain C ABI (ISO_C_BINDING-callable: plain
 * pointers, C ints and doubles, no C++ types), the reference's per-solve
 * external subroutines and the body of its per-(source, period) loop:
 *
 *   satt_fwd_lonlat        FSM_UW_PS_lonlat_2d   src/eikon_solver_2d.f90:764
 *   satt_adj_lonlat        FSM_O1_JSE_lonlat_2d  src/eikon_solver_2d.f90:1276
 *   satt_mask              mask                  src/eikon_solver_2d.f90:1449
 *   satt_interp2           interp2_1_dp          src/shared/utils.f90:944
 *   satt_forward_simulate  loop body of att_tomo_2d%forward_simulate
 *                          (src/tomo2d.f90:168-195) and att_acqui%forward_simulate
 *                          (src/acqui.f90:121-151), with att_measadj
 *                          (src/measadj.f90:40-114) and post_proc_eikokernel
 *                          (src/tomo2d.f90:145-156, src/acqui.f90:307-316)
 *   satt_plan_*            the same loop with the model, gathers and fields
 *                          resident in device memory across calls
 *   satt_smooth_geo2       gaussian_smooth_geo_2  src/shared/utils.f90:1194
 *   satt_gradient_geo2     gradient_2_geo         src/shared/utils.f90:1118
 *   satt_topo_metric       the metric loop of att_grid%get_topo
 *                          (src/grid.f90:114-123,137-141)
 *   satt_steepest_descent  att_tomo_2d%optimize + steepest_descent
 *                          (src/tomo2d.f90:203-251)
 *   satt_plan_linesearch   att_tomo_2d%linesearch (src/tomo2d.f90:253-288)
 *   satt_fwdsurf3d         fwdsurf3d_decomp / fwdsurf1d with surfdisp96
 *                          (src/surfker/surfker.f90:200-254,
 *                          src/surfker/surfdisp96.f:57)
 *   satt_combine_kernels   att_acqui%combine_kernels (src/acqui.f90:197-245)
 *   satt_inv_grid_iso      inv_grid_iso           src/eikon_solver_2d.f90:1538
 *   satt_inv2fwd_iso       inv2fwd_iso            src/eikon_solver_2d.f90:1613
 *   satt_kernel_postproc   att_acqui%post_processing_for_kernel
 *                          (src/acqui.f90:247-304)
 *   satt_correct_depth     correct_depth          src/surfker/surfker.f90:260
 *   satt_depthkernel       depthkernel_decomp / depthkernel1d with sregn96
 *                          (src/surfker/surfker.f90:37-103)
 *
 * Conventions
 *   - A grid field f(nx,ny) is column-major with x fastest: f[i + nx*j], as
 *     the reference's Fortran arrays.  Period-major model arrays
 *     svel/m11/m12/m22(nperiod,nx,ny) are v[p + nperiod*(i + nx*j)], the
 *     reference's att_grid layout (src/grid.f90:25-34).
 *   - Coordinates are in degrees, like the reference's call sites
 *     (src/measadj.f90:68-71,85-89).  Indices passed in are 0-based.
 *   - All host arrays are caller-owned; outputs are fully overwritten.
 *   - Every entry point returns a status code.  Fatal codes are negative.
 *     SATT_W_MAXITER is a warning: the fields are returned, as the reference
 *     returns them silently (src/eikon_solver_2d.f90:1117-1119,1436-1438).
 *   - A context owns one HIP stream on one device and is not thread-safe;
 *     use one context per host thread.
 */
#ifndef SATT_H
#define SATT_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define SATT_OK 0
#define SATT_W_MAXITER 1          /* MaxIter reached without convergence     */
#define SATT_E_ARG (-1)           /* invalid sizes / pointers                */
#define SATT_E_SRC_ON_BOUNDARY (-2) /* src/eikon_solver_2d.f90:863-867 'stop' */
#define SATT_E_REC_OUT_OF_GRID (-3) /* receiver bracket leaves the grid; the
                                     reference reads out of bounds
                                     (:1329-1335, src/shared/utils.f90:927) */
#define SATT_E_NOMEM (-4)         /* device or host allocation failed        */
#define SATT_E_HIP (-5)           /* HIP runtime error                       */
#define SATT_E_NODEV (-6)         /* no usable GPU                           */
#define SATT_E_TEAM_WAIT (-7)     /* a multi-wave solve's wait for another wave
                                     ran past its bound; the item's fields are
                                     not valid                               */
#define SATT_E_DISPERSION (-8)    /* no root of the period equation for a
                                     column (the reference stops with
                                     'Dispersion calculation failed',
                                     src/surfker/surfdisp96.f:343)           */

/* precision modes for the batched path */
#define SATT_PREC_FP64 0  /* the reference's FP64 arithmetic, sweep for sweep */
#define SATT_PREC_MIXED 1 /* FP32 sweeps, then FP64 polish: forward to the
                             reference's stopping rule, adjoint one iteration;
                             the polish computes in FP64 on coefficients
                             stored in float; T0, residuals, interpolation
                             and kernel accumulation in FP64.  Batched form
                             only. */

/* launch forms for satt_plan_run */
#define SATT_LAUNCH_BATCHED 0   /* all items of a chunk in one launch per stage */
#define SATT_LAUNCH_PER_SOLVE 1 /* one solve per launch (the plain form)     */

typedef struct satt_ctx satt_ctx;
typedef struct satt_plan satt_plan;

/* The batched form, the plans and the server take grids of at most
 * SATT_MAX_NY rows (latitudes): a solve passes two columns of FP64 values per
 * row between its strips through the workgroup's 160 KB of LDS.  Larger
 * grids are SATT_E_ARG. */
#define SATT_MAX_NY 10240

/* Create a context on HIP device `device` (its own stream). */
int satt_create(satt_ctx **ctx, int device);
void satt_destroy(satt_ctx *ctx);
/* Message of the last failing call on this context (static storage). */
const char *satt_last_error(const satt_ctx *ctx);
/* Library version string. */
const char *satt_version(void);

/* ---------------------------------------------------- per-solve (plain) form
 * One solve per launch; host arrays in and out.  `iters` (may be NULL)
 * receives the iteration count at exit. */

/* FSM_UW_PS_lonlat_2d(xx_deg,yy_deg,nx,ny,spha,sphb,sphc,T,fun,x0_deg,y0_deg,u)
 * src/eikon_solver_2d.f90:764.  `u` only feeds an unused error norm in the
 * reference (:1099-1104) and is not taken. */
int satt_fwd_lonlat(satt_ctx *ctx, const double *xx_deg, const double *yy_deg,
                    int nx, int ny, const double *spha, const double *sphb,
                    const double *sphc, const double *fun, double x0_deg,
                    double y0_deg, double *T, int *iters);

/* FSM_O1_JSE_lonlat_2d(xx_deg,yy_deg,nx,ny,spha,sphb,sphc,T,Ta,xrec_deg,
 * yrec_deg,sourceAdj,nr)  src/eikon_solver_2d.f90:1276. */
int satt_adj_lonlat(satt_ctx *ctx, const double *xx_deg, const double *yy_deg,
                    int nx, int ny, const double *spha, const double *sphb,
                    const double *sphc, const double *T, const double *xrec_deg,
                    const double *yrec_deg, const double *src_adj, int nr,
                    double *Ta, int *iters);

/* satt_fwd_lonlat without any device allocation: the same kernels on the same
 * operands, hence the same T bit for bit.  No counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its work
 * space is one static device array per GPU (about 28 MB, held until the
 * process ends), so grids are limited to SATT_NOALLOC_MAX_CELLS cells and
 * SATT_NOALLOC_MAX_AXIS nodes per axis (larger ones are SATT_E_ARG).  All
 * calls on one GPU share that array: calls that overlap on one GPU (from
 * several threads or contexts) race and must be serialised by the caller. */
#define SATT_NOALLOC_MAX_CELLS 262144 /* 512 x 512 */
#define SATT_NOALLOC_MAX_AXIS 8192
int satt_fwd_lonlat_noalloc(satt_ctx *ctx, const double *xx_deg, const double *yy_deg,
                            int nx, int ny, const double *spha, const double *sphb,
                            const double *sphc, const double *fun, double x0_deg,
                            double y0_deg, double *T, int *iters);

/* mask(xx,yy,nx,ny,TableAdj,x0,y0)  src/eikon_solver_2d.f90:1449 (in place). */
int satt_mask(satt_ctx *ctx, const double *xx, const double *yy, int nx, int ny,
              double *table, double x0, double y0);

/* vq = interp2(x, y, v, xq, yq)  src/shared/utils.f90:944 (interp2_1_dp). */
int satt_interp2(satt_ctx *ctx, const double *x, const double *y, int nx,
                 int ny, const double *v, const double *xq, const double *yq,
                 int nq, double *vq);

/* ------------------------------------------- model-side steps (per call) */
/* Axes for these calls: lons(nx), lats(ny) in degrees, increasing, nx and ny
 * >= 2, nx <= SATT_GEO_MAX_NX, the longitude axis uniform (as arange makes it,
 * src/shared/utils.f90:645-653; to 1e-6 of its spacing). */
#define SATT_GEO_MAX_NX 4096

/* out = gaussian_smooth_geo_2(data, lons, lats, sigma) per field
 * (src/shared/utils.f90:1194-1223), nfield fields at once.  data holds ndata
 * fields, ndata = nfield or 1 (one field smoothed with every sigma); data and
 * out are period-major, (ndata, nx, ny) and (nfield, nx, ny); sigma[nfield]
 * in degrees, > 0 (the reference's tomo2d refuses sigma_2d = 0,
 * src/tomo2d.f90:48-49).  Within 1e-12 of the reference relative to max|data|:
 * a weight is evaluated per (row pair, column offset) rather than per node
 * pair, which differs from the reference by the rounding of lon(i')-lon(i). */
int satt_smooth_geo2(satt_ctx *ctx, const double *lons, const double *lats,
                     int nx, int ny, int nfield, int ndata, const double *data,
                     const double *sigma, double *out);

/* gradient_2_geo(f, lon, lat, tx, ty) (src/shared/utils.f90:1118-1160) per
 * field, nfield fields, period-major (nfield, nx, ny) in and out. */
int satt_gradient_geo2(satt_ctx *ctx, const double *lons, const double *lats,
                       int nx, int ny, int nfield, const double *f, double *tx,
                       double *ty);

/* The topographic metric of att_grid%get_topo (src/grid.f90:114-123,137-141)
 * for nperiod periods: z(nx,ny) (km, on the model grid, as after
 * att_topo%grid_topo) smoothed with sigma[p], then m11 = a, m22 = b,
 * m12 = -c from its gradient, and the dip angle (src/topo.f90:123-130) in
 * degrees unless angle is NULL.  Outputs period-major (nperiod, nx, ny). */
int satt_topo_metric(satt_ctx *ctx, const double *lons, const double *lats,
                     int nx, int ny, const double *z, int nperiod,
                     const double *sigma, double *m11, double *m12,
                     double *m22, double *angle);

/* sigma of one period for satt_topo_metric, src/grid.f90:116:
 * vel*period*wavelen_factor*km2deg/(2*pi), vel the period's velocity of the
 * 1-D model (fwdsurf1d), in degrees. */
double satt_topo_sigma(double vel, double period, double wavelen_factor);

/* att_tomo_2d%optimize followed by steepest_descent (src/tomo2d.f90:203-251):
 * ker_s(p,:,:) = gaussian_smooth_geo_2(adj_s(p,:,:), lons, lats, sigma_2d)
 * for every period, then svel = svel*(1 + updatemax*ker_s/maxval(abs(ker_s)))
 * in place.  The caller applies the misfit-increase shrink of updatemax
 * (:238-246) first.  adj_s, svel, ker_s period-major (nperiod, nx, ny). */
int satt_steepest_descent(satt_ctx *ctx, const double *lons, const double *lats,
                          int nx, int ny, int nperiod, const double *adj_s,
                          double sigma_2d, double updatemax, double *svel,
                          double *ker_s);

/* The surface-wave velocity maps of att_grid%fwdsurf (src/grid.f90:150-163):
 * for every column of vs3d(nx,ny,nz) (km/s, Fortran order) on depths depz(nz)
 * (km), fwdsurf1d (src/surfker/surfker.f90:224-254: Vp and density from Vs by
 * Brocher's relations, layers between the depths over a half-space) and
 * surfdisp96 (src/surfker/surfdisp96.f:57-377) with a spherical earth and the
 * fundamental mode: the phase (igr 0) or group (igr > 0) velocity of Love
 * (iwave 1) or Rayleigh (iwave 2) waves at periods(nperiod) (s, in the order
 * given; the search of each period starts from the previous one's root).
 * svel is period-major (nperiod, nx, ny), as att_grid%svel.  A column whose
 * search finds no root gets zeros from that period on and the call returns
 * SATT_E_DISPERSION (the reference stops the run there); col_status (nx*ny,
 * may be NULL) holds each column's code.  nz <= SATT_DISP_MAX_NZ. */
#define SATT_DISP_MAX_NZ 256
int satt_fwdsurf3d(satt_ctx *ctx, const double *vs3d, int nx, int ny, int nz,
                   const double *depz, int nperiod, const double *periods,
                   int iwave, int igr, double *svel, int *col_status);

/* The depth sensitivity kernels of att_acqui%depthkernel (src/acqui.f90:
 * 163-195): depthkernel_decomp / depthkernel1d (src/surfker/surfker.f90:
 * 37-103) for every column of vs3d(nx,ny,nz) on depths depz(nz): REAL(4) Vs,
 * Vp and density by Brocher's relations, refinegrid's layers
 * (:285-307: nz layers and a half-space copy of the last), surfdisp96's
 * phase velocity at each period and CPS sregn96 (src/surfker/sregn96.f90)
 * at that velocity with a spherical earth.  sen_vs, sen_vp, sen_rho are
 * (nperiod, nx, ny, nz), as sen_vsRc/sen_vpRc/sen_rhoRc: dc/dVs, dc/dVp and
 * dc/drho of each layer.  The Rayleigh phase-velocity branch (iwave 2,
 * igr 0) is built; Love waves (slegn96) and group velocities (sregnpu,
 * slegnpu) are SATT_E_ARG.  cphase (nperiod, nx, ny; may be NULL) receives
 * the phase velocities the kernels were evaluated at (depthkernel1d's cp).
 * A column whose dispersion search finds no root gets zero kernels and the
 * call returns SATT_E_DISPERSION; col_status (nx*ny, may be NULL) holds each
 * column's code.  nz + 1 <= SATT_DISP_MAX_NZ. */
int satt_depthkernel(satt_ctx *ctx, const double *vs3d, int nx, int ny, int nz,
                     const double *depz, int nperiod, const double *periods, int iwave,
                     int igr, double *sen_vs, double *sen_vp, double *sen_rho,
                     double *cphase, int *col_status);

/* ---- 3-D kernel steps of the att_acqui path (after the stacking) ---- */

/* combine_kernels (src/acqui.f90:197-245): the stacked period kernels
 * adj_s, adj_density (nperiod,nx,ny) times the depth sensitivities sen_vs,
 * sen_vp, sen_rho (nperiod,nx,ny,nz), summed over periods into ker and
 * ker_density (nker,nx,ny,nz), which are accumulated into (the reference
 * zeroes them first, prepare_inv :73-88).  use_alpha_beta_rho: nker = 3
 * (Vs, Vp, density components; with rho_scaling the density component is
 * 0.33 times the Vs one and sen_rho may be NULL); else nker = 1 through the
 * chain rule of Brocher's relations on vs3d, vp3d (nx,ny,nz)
 * (dalpha_dbeta, drho_dalpha, src/shared/utils.f90:99-145). */
int satt_combine_kernels(satt_ctx *ctx, int nperiod, int nx, int ny, int nz,
                         const double *adj_s, const double *adj_density,
                         const double *sen_vs, const double *sen_vp, const double *sen_rho,
                         const double *vs3d, const double *vp3d, int use_alpha_beta_rho,
                         int rho_scaling, double *ker, double *ker_density);

/* correct_depth (src/surfker/surfker.f90:260-283), the topography step of
 * att_acqui%depthkernel (src/acqui.f90:183-187): every (period, column)
 * profile of the depth kernels ker_vs, ker_vp, ker_rho (nperiod,nx,ny,nz),
 * in place, becomes its linear interpolation (interp1, src/shared/
 * utils.f90:786-805) at depz/cosd(dip) for the column's dip angle
 * dip_angle (nperiod,nx,ny, degrees in [0, 90); satt_topo_metric's angle).
 * depz (nz) increases from >= 0. */
int satt_correct_depth(satt_ctx *ctx, int nperiod, int nx, int ny, int nz,
                       double *ker_vs, double *ker_vp, double *ker_rho,
                       const double *dip_angle, const double *depz);

/* inv_grid_iso (src/eikon_solver_2d.f90:1538-1610), the reference's
 * arguments: gk(nset*nxinv*nyinv*nzinv) = the interior points of adj
 * (nx,ny,nz) spread trilinearly onto the inversion grids xinv(nxinv,nset),
 * yinv, zinv, times dx*dy*dz.  A point outside a set's grid is SATT_E_ARG
 * (the reference stops). */
int satt_inv_grid_iso(satt_ctx *ctx, const double *xinv, const double *yinv,
                      const double *zinv, int nxinv, int nyinv, int nzinv, int nset,
                      const double *adj, int nx, int ny, int nz, double *gk,
                      const double *vel_x, const double *vel_y, const double *vel_z);
/* inv2fwd_iso (src/eikon_solver_2d.f90:1613-1691): update(nx,ny,nz) = the
 * trilinear interpolation of gk at every model point, summed over sets. */
int satt_inv2fwd_iso(satt_ctx *ctx, const double *xinv, const double *yinv,
                     const double *zinv, int nxinv, int nyinv, int nzinv, int nset,
                     int nx, int ny, int nz, const double *gk, const double *vel_x,
                     const double *vel_y, const double *vel_z, double *update);
/* post_processing_for_kernel (src/acqui.f90:247-285) for every component
 * iker of ker (nker,nx,ny,nz): with kdensity_coe > 0 the component is first
 * preconditioned in place by regularize_ker_density (:287-304) from
 * ker_density (nker,nx,ny,nz; may be NULL otherwise); then projected to the
 * inversion grids and back, and gradient_s(iker,:,:,:) +=
 * update/nset/chi0*weight.  The projections stay on the device. */
int satt_kernel_postproc(satt_ctx *ctx, int nker, int nx, int ny, int nz,
                         const double *vel_x, const double *vel_y, const double *vel_z,
                         int nset, int nxinv, int nyinv, int nzinv, const double *xinv,
                         const double *yinv, const double *zinv, double *ker,
                         const double *ker_density, double kdensity_coe, double chi0,
                         double weight, double *gradient_s);

/* Achievable HBM bandwidth: the best of five runs of a streaming read and of a
 * streaming copy over nbytes (1 MiB to 64 GiB; allocates 2*nbytes), in GB/s
 * (the copy counts bytes read and written).  A measurement aid for the
 * roofline of the kernels above; no counterpart in the reference. */
int satt_hbm_probe(satt_ctx *ctx, double nbytes, double *read_gbs, double *copy_gbs);
/* Achievable LDS read bandwidth (best of five runs of 16-byte LDS reads by
 * every thread of four workgroups per CU), and the peak it is held against,
 * 256 bytes per clock per CU (gfx950) at the device's clock, both in GB/s.  A
 * measurement aid; no counterpart in the reference. */
int satt_lds_probe(satt_ctx *ctx, double *read_gbs, double *peak_gbs);

/* Device time in ms of the kernels of the last satt_smooth_geo2,
 * satt_gradient_geo2, satt_topo_metric, satt_steepest_descent,
 * satt_fwdsurf3d, satt_depthkernel, satt_combine_kernels, satt_inv_grid_iso,
 * satt_inv2fwd_iso, satt_kernel_postproc or satt_correct_depth call on the
 * context (HIP events on its stream; transfers excluded). */
double satt_last_device_ms(const satt_ctx *ctx);

/* ------------------------------------------------------------ batched form */

/* One forward_simulate call's inputs, in the reference's shapes.
 *   dim 2: att_tomo_2d path (src/tomo2d.f90:158): kernel adj/svel**2.
 *   dim 3: att_acqui path (src/acqui.f90:91): adds the density adjoint
 *          (RHS -weight, src/measadj.f90:92-104); kernels adj/svel**3 and
 *          kden/svel**3 (src/acqui.f90:314-315).
 * Items are the reference's isrcs list in order (src/src_rec.f90:338-369);
 * item_pidx is 0-based (isrcs(:,1)-1); the receiver gather of item i
 * (get_evt_gather, src/src_rec.f90:252-283) is rec_off[i]..rec_off[i+1]-1 of
 * recx/recy/tt/weight/idx_path (idx_path 0-based). */
typedef struct {
  int dim;
  int nx, ny, nperiod;
  const double *xgrids, *ygrids;               /* am%xgrids, am%ygrids (deg) */
  const double *svel, *m11, *m12, *m22;        /* (nperiod,nx,ny) */
  int nitems;
  const int *item_pidx;
  const double *item_srcx, *item_srcy;         /* source lon/lat (deg) */
  const int *rec_off;                          /* nitems+1 */
  const double *recx, *recy, *tt, *weight;     /* rec_off[nitems] each */
  const int *idx_path;
  int npath;
} satt_batch;

/* Outputs; any pointer may be NULL to skip it. */
typedef struct {
  double *tt_fwd;      /* npath: sum_all(local_tt) (src/tomo2d.f90:200)     */
  double *chi;         /* 1: sum of 0.5*sum(w*(t-tobs)**2) over items       */
  double *adj_s;       /* (nperiod,nx,ny) stacked kernel (isadj)            */
  double *adj_density; /* (nperiod,nx,ny) dim 3 only (isadj)                */
  double *item_chi;    /* nitems                                            */
  int *iters;          /* 3*nitems: forward, adjoint, density-adjoint       */
  int *item_status;    /* nitems: per-item status                           */
  double *weight_out;  /* rec_off[nitems]: weights after outlier rejection  */
  int *iters_fp32;     /* 3*nitems: FP32-phase iterations (mixed mode; 0 in
                          FP64 mode); `iters` then counts the FP64 polish  */
} satt_result;

/* One call of forward_simulate(chi, istotable, isadj): upload, run, download,
 * as satt_plan_create + satt_plan_run + satt_plan_results.  The plan stays
 * resident in the context: the reference calls forward_simulate once per
 * iteration and per line-search step with the same gathers and a new model
 * (src/tomo2d.f90:104,272), so a call whose batch equals the previous one in
 * everything but svel/m11/m12/m22 (and the precision) only uploads the model
 * (satt_plan_set_model) before running.  A new plan has the team size of
 * each sweep stage tuned on its own inputs once (satt_plan_tune_waves).  A
 * call with other gathers replaces the plan; its device memory is held until
 * then, satt_forward_simulate_release or satt_destroy. */
int satt_forward_simulate(satt_ctx *ctx, const satt_batch *batch,
                          int istotable, int isadj, int precision,
                          satt_result *result);
/* Frees satt_forward_simulate's resident plan (no-op without one). */
void satt_forward_simulate_release(satt_ctx *ctx);

/* Resident form: the plan keeps the grid, the model, the gathers and every
 * field in device memory.  satt_plan_run is asynchronous on the context's
 * stream; satt_plan_results waits for it and downloads. */
int satt_plan_create(satt_ctx *ctx, const satt_batch *batch, int precision,
                     satt_plan **plan);
/* Several grids in one plan (uneven batching): batch k is part k, with its
 * own grid, model, items and outputs; the parts' items share each launch
 * (the most expensive grid first).  All parts have the same dim.  A plan of
 * one part is satt_plan_create's.  No counterpart in the reference, whose
 * runs hold one grid. */
int satt_plan_create_multi(satt_ctx *ctx, int nbatch, const satt_batch *batches,
                           int precision, satt_plan **plan);
/* Replace the model (same shapes) for the next run, e.g. after a model
 * update or inside a line search (src/tomo2d.f90:272). */
int satt_plan_set_model(satt_plan *plan, const double *svel, const double *m11,
                        const double *m12, const double *m22);
int satt_plan_set_model_part(satt_plan *plan, int part, const double *svel,
                             const double *m11, const double *m12, const double *m22);
int satt_plan_run(satt_plan *plan, int istotable, int isadj, int launch_form);
/* Replay form of the batched run: with `on` set, the first batched run of
 * each (istotable, isadj) pair captures the whole launch chain -- for every
 * launch group the forward, the receivers (misfit and adjoint sources), the
 * adjoint(s) and the kernel stacking -- into one HIP graph, and later runs
 * with that pair replay it as one launch.  The graph holds the plan's device
 * buffers, which satt_plan_set_model refills in place, so model updates need
 * no new capture.  Stage times are not split in this form (ms_total only).
 * No counterpart in the reference. */
int satt_plan_set_graph(satt_plan *plan, int on);
/* Waves per item of the batched sweep kernels (default 1).  With W > 1 an
 * item's strips are shared by W waves of one workgroup, each strip of a
 * sweep following the previous one at least a strip width behind, which
 * cuts an item's latency on wide grids; one wave per item is the throughput
 * form when the items fill the device.  W is capped per launch group (at
 * most 8, at most the strip count, and the LDS of the edge channels).  The
 * results do not depend on W.  No counterpart in the reference. */
int satt_plan_set_waves(satt_plan *plan, int waves);
/* waves = 0: per launch group and sweep stage, the team sizes the autotuner
 * measured for its grid shape and size (tuned on the spot when the context
 * has none yet). */
/* Team sizes per sweep stage, waves[4] >= 1 in satt_tune_result.stage_waves
 * order (forward FP32, forward FP64, adjoint FP32, adjoint FP64), each capped
 * per launch group as above: pins a configuration the autotuner found.  No
 * counterpart in the reference. */
int satt_plan_set_stage_waves(satt_plan *plan, const int *waves);
/* How many steps ahead each sweep stage fetches its operands, depth[4] in
 * satt_tune_result.stage_depth order, each 2 (the default) or 4.  Fetching
 * further ahead hides more memory latency where few waves share a SIMD, at
 * the price of registers.  The results do not depend on it.  No counterpart
 * in the reference. */
int satt_plan_set_stage_depth(satt_plan *plan, const int *depth);

/* Autotuner of the launch shape: times a synthetic batch of `nitems`
 * fwd+adj solves on an nx x ny grid for each team size W in 1..8 (as far as
 * the grid's strips allow) with the two-step prefetch, then with the deep
 * (four-step) prefetch at the W each stage chose and at W = 1, and keeps, for
 * each of the four sweep stages (forward FP32, forward FP64, adjoint FP32,
 * adjoint FP64), the fastest (W, depth) in the context's table, keyed by
 * strips, rows, item count and precision, for plans
 * (satt_plan_set_waves(plan, 0)) and servers (waves 0).  waves is the W of
 * the fastest whole two-step run and ms[W-1] that run's time;
 * stage_ms[W-1][k] is stage k's two-step time at W (0: not tried or not run
 * in this precision), stage_ms_deep its deep-prefetch times, and
 * stage_waves[k], stage_depth[k] its choice.  No counterpart in the
 * reference. */
typedef struct {
  int waves;
  double ms[8];
  int stage_waves[4];
  double stage_ms[8][4];
  int stage_depth[4];          /* prefetch depth chosen per stage: 2 or 4   */
  double stage_ms_deep[2][4];  /* each stage's time with the deep prefetch:
                                  [0] at the team sizes the two-step sweep
                                  chose, [1] at one wave per item (0: not
                                  tried)                                     */
} satt_tune_result;
int satt_tune_waves(satt_ctx *ctx, int nx, int ny, int nitems, int precision,
                    satt_tune_result *result);

/* Autotunes the plan on its own inputs: runs it (istotable and isadj set) with
 * every W in 1..8 and the two prefetch depths as satt_tune_waves does, keeps
 * each sweep stage's fastest (W, depth) and sets those (as
 * satt_plan_set_stage_waves and satt_plan_set_stage_depth); result fields as
 * satt_tune_waves.  The plan's
 * results afterwards are those of a run (they do not depend on W).  No
 * counterpart in the reference. */
int satt_plan_tune_waves(satt_plan *plan, satt_tune_result *result);
/* Wait for the plan's last run (no download). */
int satt_plan_wait(satt_plan *plan);
int satt_plan_results(satt_plan *plan, satt_result *result);
/* The results of part `part` (satt_plan_results: part 0). */
int satt_plan_results_part(satt_plan *plan, int part, satt_result *result);
/* att_tomo_2d%linesearch (src/tomo2d.f90:253-288) on a one-grid plan, with
 * prepare_fwd_linesearch (src/acqui_2d.f90:173-183).  For sit = 1 ..
 * max_sub_niter: the plan's velocity becomes
 * svel*(1 + (*updatemax)*ker_s/maxval(abs(ker_s))) (formed on the device),
 * the plan runs forward only (forward_simulate(chi, .false., .false.)) and
 * if its misfit exceeds chi0 the step shrinks (*updatemax *= maxshrink) and
 * the search goes on, else it stops.  svel_out gets the last velocity tried
 * (acqui%svel = acqui%ag%svel, also when every sub-iteration raised the
 * misfit), *chi_out the last misfit, *nsub the sub-iterations run; the plan
 * keeps that velocity as its model (m11, m12, m22 unchanged).  svel, ker_s,
 * svel_out are (nperiod,nx,ny); with max_sub_niter < 1 svel_out = svel and
 * *chi_out = chi0.  ker_s must have a nonzero finite maximum. */
int satt_plan_linesearch(satt_plan *plan, const double *svel, const double *ker_s,
                         double chi0, double maxshrink, int max_sub_niter,
                         double *updatemax, double *svel_out, double *chi_out,
                         int *nsub);
void satt_plan_destroy(satt_plan *plan);

/* Sizing and timing hooks of a plan (for measurement harnesses). */
typedef struct {
  int chunk;                /* items per launch                              */
  int launches;             /* launch groups of the last run (per stage)     */
  long long device_bytes;   /* device memory held by the plan                */
  double ms_forward;        /* device time of the last run's forward kernels */
  double ms_adjoint;        /* ... adjoint kernels                           */
  double ms_total;          /* ... whole run                                 */
  long long fwd_sweeps;     /* sum over items of 4*forward iterations (FP64) */
  long long adj_sweeps;     /* sum over items of 4*adjoint iterations (both) */
  long long fwd_sweeps32;   /* the same for the FP32 phase of the mixed mode */
  long long adj_sweeps32;
  double ms_fwd_fp64;       /* forward FP64 sweep kernels                    */
  double ms_fwd_fp32;       /* forward FP32-phase kernels (mixed mode)       */
  double ms_adj_fp64;       /* adjoint FP64 kernels (both adjoints)          */
  double ms_adj_fp32;       /* adjoint FP32-phase kernels (mixed mode)       */
  double ms_fwd_setup;      /* forward setup (T0 fields, tau init)           */
  double ms_adj_setup;      /* adjoint setup (delta source, coefficients)    */
} satt_plan_stats;
int satt_plan_get_stats(satt_plan *plan, satt_plan_stats *stats);

/* A field of one item of part 0 in the plan's last run, in natural layout
 * (nx,ny): SATT_FIELD_T travel times, SATT_FIELD_TA adjoint field (masked),
 * SATT_FIELD_TD density adjoint field (dim 3).  Only items of the run's last
 * launch group are resident; others give SATT_E_ARG. */
#define SATT_FIELD_T 0
#define SATT_FIELD_TA 1
#define SATT_FIELD_TD 2
int satt_plan_item_field(satt_plan *plan, int item, int which, double *out);

/* ------------------------------------------------------------ serving
 * A request loop for independent fwd+adj solves (no counterpart in the
 * reference, whose loop runs a fixed list).  Each request carries its own
 * grid, phase-velocity map (one period) and receivers; the server copies the
 * inputs at submit.  A scheduler thread admits waiting requests while device
 * memory (a pool reserved at creation) and launch streams are free, runs each
 * admitted group as one batched launch sequence on its own stream -- groups
 * overlap on the device -- and fills the responses as groups complete.  A
 * request's results do not depend on what it was batched with. */
typedef struct satt_server satt_server;

typedef struct {
  int precision;          /* SATT_PREC_FP64 or SATT_PREC_MIXED                 */
  int max_batch;          /* requests per launch group (0: 256)                */
  int streams;            /* launch groups in flight (0: 6)                    */
  long long device_bytes; /* pool for requests in flight (0: half the free memory) */
  int waves;              /* waves per request in the sweep kernels, capped by
                             the grid's strips (0: the context's size tuned
                             for one request of the group's widest grid, see
                             satt_tune_waves, else 4); a team of waves cuts a
                             request's latency, see satt_plan_set_waves      */
} satt_server_config;

typedef struct {
  int nx, ny;
  const double *xgrids, *ygrids;  /* degrees */
  const double *svel;             /* (nx,ny) phase velocity of the request's period */
  const double *m11, *m12, *m22;  /* (nx,ny) metric, or all NULL: 1, 0, 1       */
  double srcx, srcy;              /* source, degrees                             */
  int nrec;
  const double *recx, *recy, *tt, *weight;  /* receivers (degrees), tobs, weight */
} satt_request;

typedef struct {
  double *tt_rec;        /* nrec: travel times at the receivers (may be NULL)   */
  double *kernel;        /* (nx,ny): masked adjoint field / svel**2 (may be NULL) */
  double *T;             /* (nx,ny): travel-time field (may be NULL)           */
  double chi;            /* 0.5*sum(w*(t-tobs)**2)                              */
  int status;            /* the request's status, once done                     */
  int iters[2];          /* forward, adjoint iterations (FP64; polish if mixed)  */
  int iters_fp32[2];     /* FP32-phase iterations (mixed)                        */
  double queue_ms;       /* submit to launch                                    */
  double latency_ms;     /* submit to results in this struct                    */
} satt_response;

int satt_server_create(satt_ctx *ctx, const satt_server_config *cfg,
                       satt_server **server);
/* Queue one request; the response must stay valid until the request is done.
 * Returns SATT_E_ARG for a malformed request or one larger than the pool. */
int satt_server_submit(satt_server *server, const satt_request *req,
                       satt_response *resp, long long *ticket);
/* 1 if the request is done, 0 if not. */
int satt_server_test(satt_server *server, long long ticket);
/* Block until the request is done; returns its status. */
int satt_server_wait(satt_server *server, long long ticket);
/* Block until every submitted request is done. */
int satt_server_drain(satt_server *server);
typedef struct {
  long long submitted, completed, launches;
  long long max_in_flight;  /* most requests on the device at once */
  double device_busy_ms;    /* sum of launch-group device times   */
} satt_server_stats;
int satt_server_get_stats(satt_server *server, satt_server_stats *stats);
/* Drains, stops the scheduler and frees the pool. */
void satt_server_destroy(satt_server *server);

/* ------------------------------------------------------------ inspection
 * The per-solve setup fields as the device computes them, for bit-level
 * comparison against a CPU restatement that takes them as inputs (the
 * device's libm differs from the host's in the last bits).
 *   forward: a, b, c (metric, :817-821), T0, T0x, T0y (:845-857)
 *   adjoint: a, b, c, delta (:1320-1344) */
int satt_fwd_setup_fields(satt_ctx *ctx, const double *xx_deg,
                          const double *yy_deg, int nx, int ny,
                          const double *spha, const double *sphb,
                          const double *sphc, const double *fun, double x0_deg,
                          double y0_deg, double *a, double *b, double *c,
                          double *T0, double *T0x, double *T0y);
int satt_adj_setup_fields(satt_ctx *ctx, const double *xx_deg,
                          const double *yy_deg, int nx, int ny,
                          const double *spha, const double *sphb,
                          const double *sphc, const double *xrec_deg,
                          const double *yrec_deg, const double *src_adj, int nr,
                          double *a, double *b, double *c, double *delta);

#ifdef __cplusplus
}
#endif
#endif /* SATT_H */
