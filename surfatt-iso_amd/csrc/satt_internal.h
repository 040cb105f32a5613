// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

struct satt_ctx;

namespace satt {

// The device and the stream of a context, and recording a failure (its
// message and code) on it (satt_api.hip).
int ctx_device(const satt_ctx *ctx);
hipStream_t ctx_stream(const satt_ctx *ctx);
int ctx_fail(satt_ctx *ctx, int code, const char *fmt, ...);

// Looks up the allocation-free entry point's scratch address on `device`,
// which must be current (satt_noalloc.hip); once per device, thread-safe.
int noalloc_init(int device);

// Key of the context's tuning table (satt_tune.hip): strips, rows rounded
// up to 32, the launch group's item count rounded up to a power of two,
// precision.
struct TuneKey {
  int ns, ny32, count2, precision;
  static TuneKey of(int nx, int ny, int count, int precision) {
    int c2 = 1;
    while (c2 < count) c2 <<= 1;
    return TuneKey{(nx + 63) / 64, (ny + 31) / 32 * 32, c2, precision};
  }
  bool operator<(const TuneKey &o) const {
    if (ns != o.ns) return ns < o.ns;
    if (ny32 != o.ny32) return ny32 < o.ny32;
    if (count2 != o.count2) return count2 < o.count2;
    return precision < o.precision;
  }
};
// Team sizes and prefetch depths of the four sweep stages (forward FP32,
// forward FP64, adjoint FP32, adjoint FP64).
struct TuneWaves {
  int w[4];
  int d[4] = {2, 2, 2, 2};
};
// The tuned team sizes of a key (false: not tuned), and storing them; both
// thread-safe.
bool ctx_tune_lookup(satt_ctx *ctx, const TuneKey &k, TuneWaves &out);
void ctx_tune_store(satt_ctx *ctx, const TuneKey &k, const TuneWaves &waves);

// Constants of the reference solvers, with their REAL(4) rounding kept:
// src/eikon_solver_2d.f90:774-776,785 (forward) and :1293-1297 (adjoint).
constexpr double kPiF = (double)3.14159265358979323846264338327950288f;
constexpr double kREarth = 6371.0;
constexpr double kFwdTol = (double)1.0e-4f;
constexpr double kAdjTol = (double)1.0e-9f;
constexpr double kAdjEps = (double)1.0e-14f;
constexpr int kFwdMaxIter = 2000;
constexpr int kAdjMaxIter = 1000;

// status codes (include/satt.h)
constexpr int kOk = 0, kWMaxIter = 1, kEArg = -1, kESrcOnBoundary = -2,
              kERecOutOfGrid = -3, kENoMem = -4, kEHip = -5, kETeamWait = -7,
              kEDispersion = -8;

// The lon/lat grid shared by every item of a batch.  Radian axes are made on
// the host with the reference's expression xx_deg/180.0*pi.
struct Grid {
  int nx, ny;
  double dx, dy;          // radians, xx(2)-xx(1), yy(2)-yy(1)
  const double *xx, *yy;  // radians
  const double *xd, *yd;  // degrees (interp2 and mask work in degrees)
  const double *sy, *cy;  // sin(yy), cos(yy), computed on the device
};

// Skewed layout of one field: ns strips of 64 columns, R = ny+63 rows of 64
// lanes.  Plus skew: cell (64c+l, j) at row j+l; minus skew: row j-l+63.
struct SkewGeom {
  int ns, R;
  __host__ __device__ size_t skew_elems() const { return (size_t)ns * R * 64; }
};

inline SkewGeom make_skew_geom(int nx, int ny) {
  SkewGeom s;
  s.ns = (nx + 63) / 64;
  s.R = ny + 63;
  return s;
}

// Everything the batched kernels need about one item, one descriptor per
// slot of a launch group: the item's grid (items of different grids share a
// launch), its period's fields, its scratch and its inputs and outputs.
struct SlotDesc {
  Grid g;
  SkewGeom sg;
  const double *pskew;    // the period's skewed fields, [2 skews][kPF][SE]
  const float *pskew32;   // the same in float (mixed mode), or null: per skew
                          // a, b, c and fun, dis_x, dis_y as two sets of
                          // three stored element by element
  const double *fun;      // the period's 1/svel, natural layout
  const double *vel;      // the period's svel, natural layout
  double *kern;           // the item's own kernel Ta/svel**2 (serving), or null
  double *skew;           // scratch: slot_skews skewed fields
  double *T, *Ta, *Td, *delta, *aux;  // natural fields and forward tables
  int slot_skews, aux_len;
  double x0, y0;          // source (degrees)
  int nr;                 // receivers; the pointers below start at the item's first
  const double *recx, *recy, *tt, *weight;
  const int *idx_path;
  double *tekio, *weight_out, *rhs_adj, *rhs_den;
  double *item_chi;       // &item_chi[item]
  double *tt_fwd;         // the part's tt_fwd (indexed by idx_path)
  int *status;            // &status[item]
  int *iters;             // &iters[item * 6]: fwd, adj, density (FP64), then FP32
};

// Per-period fields in natural layout, period p at offset p*nx*ny.
struct PeriodFields {
  const double *a, *b, *c;  // converted metric (src/eikon_solver_2d.f90:817-821)
  const double *fun;        // 1/svel (src/measadj.f90:67)
  const double *vel;        // svel (kernel stacking)
};

// One chunk of items: item k of the chunk is global item item0+k and uses
// scratch slot k.
struct Items {
  int item0, count;
  const int *pidx;                  // per global item
  const double *srcx, *srcy;        // degrees, per global item
  const int *rec_off;               // per global item (+1)
  const double *recx, *recy;        // degrees, per receiver (global index)
  const double *tt, *weight;        // per receiver
  const int *idx_path;              // per receiver
};

struct FwdWork {
  double *T;                        // out [slot][N]
  double *tau, *tau_old, *T0, *T0x, *T0y;  // scratch [slot][N]
};

struct AdjWork {
  const double *T;                  // forward travel times [slot][N]
  double *Ta;                       // out [slot][N]
  double *coef;                     // scratch [slot][5][N]: a1,a2,b1,b2,d
  double *delta, *old;              // scratch [slot][N]
  const double *rhs;                // per receiver (global index)
};


// ---- launches (satt_kernels.hip); all asynchronous on `s`
// sin(yy), cos(yy) of a grid into sy, cy (satt_strip.hip)
void launch_grid_trig(hipStream_t s, const Grid &g, double *sy, double *cy);
void launch_period_fields(hipStream_t s, const Grid &g, int nperiod,
                          const double *svel_pm, const double *m11_pm,
                          const double *m12_pm, const double *m22_pm,
                          double *a, double *b, double *c, double *fun,
                          double *vel);
void launch_metric_fields(hipStream_t s, const Grid &g, const double *spha,
                          const double *sphb, const double *sphc, double *a,
                          double *b, double *c);
void launch_forward(hipStream_t s, const Grid &g, const PeriodFields &pf,
                    const Items &it, const FwdWork &w, int *iters, int *status,
                    int iter_stride);
// Receivers of every slot of a launch group (descriptor-driven); tt_fwd
// is written when istotable is set.
void launch_receivers(hipStream_t s, const SlotDesc *sd, int count, int istotable);
void launch_adjoint(hipStream_t s, const Grid &g, const PeriodFields &pf,
                    const Items &it, const AdjWork &w, int *iters, int *status,
                    int iter_stride, int apply_mask);
// acc(p) += field/vel(p)**power over the slots per_slot[per_off[p]..) of a
// launch group, in list order; field = Ta (which 0) or Td (which 1).
void launch_stack(hipStream_t s, const Grid &g, const PeriodFields &pf,
                  int nperiod, const int *per_off, const int *per_slot,
                  const SlotDesc *sd, int which, double *acc, int power);
void launch_stack_item(hipStream_t s, const Grid &g, const PeriodFields &pf,
                       const int *pidx, int item, const double *field,
                       double *acc, int power);
void launch_to_period_major(hipStream_t s, int N, int nperiod,
                            const double *nat, double *pm);
// Zero `bytes` (a multiple of 4) of device memory with a kernel.
void launch_zero(hipStream_t s, void *p, size_t bytes);
void launch_mask(hipStream_t s, const Grid &g, double *table, double x0_deg,
                 double y0_deg);
void launch_interp2(hipStream_t s, const Grid &g, const double *v,
                    const double *xq, const double *yq, int nq, double *vq,
                    int *status);
void launch_fwd_setup_fields(hipStream_t s, const Grid &g, const double *a,
                             const double *fun, double x0_deg, double y0_deg,
                             double *T0, double *T0x, double *T0y, int *status);
void launch_adj_delta(hipStream_t s, const Grid &g, const double *xr,
                      const double *yr, const double *rhs, int nr,
                      double *delta, int *status);

// ---- geographic smoothing, gradient, metric (satt_geo.hip); all on `s`
// Smoothing of nfield fields (natural layout [f][j][i], or one shared field
// when ndata is 1) with per-field sigma (degrees); window half-widths hx[f]
// and hy[f*ny + j] from the host; s2 from launch_geo_s2.
struct GeoSmooth {
  int nx, ny, ndata;
  const double *lat;            // degrees
  const double *s2;             // sin((lon(k)-lon(0))*deg2rad/2)**2, k < nx
  const double *data, *sigma;
  const int *hx, *hy;
  double *out;                  // [f][j][i]
};
struct GeoGrad {
  int nx, ny;
  const double *lon, *lat;      // degrees
  const double *cosd_lat;       // cosd(lat(j)), host-computed
};
size_t smooth_lds_bytes(int nx);
void launch_geo_s2(hipStream_t s, const double *lon, int nx, double *s2);
void launch_smooth_geo2(hipStream_t s, const GeoSmooth &P, int nfield);
// tx, ty period-major (nf, nx, ny) of nf natural fields
void launch_gradient_geo2(hipStream_t s, const GeoGrad &G, int nf, const double *f_nat,
                          double *tx_pm, double *ty_pm);
// m11, m12, m22 (and angle unless null) period-major from nf smoothed fields
void launch_topo_metric(hipStream_t s, const GeoGrad &G, int nf, const double *sm_nat,
                        double *m11, double *m12, double *m22, double *angle);
void launch_pm_to_nat(hipStream_t s, size_t N, int nf, const double *pm, double *nat);
// scratch `part` holds absmax_parts(n) + 1 doubles
int absmax_parts(size_t n);
void launch_sd_update(hipStream_t s, size_t n, const double *ker, double updatemax, double *svel,
                      double *part);
// max |v| into part[absmax_parts(n)]; returns that address
const double *launch_absmax(hipStream_t s, size_t n, const double *v, double *part);
// out = base*(1 + updatemax*ker/(*mx)), the line search's trial model
void launch_ls_model(hipStream_t s, size_t n, const double *base, const double *ker,
                     const double *mx, double updatemax, double *out);

// ---- 3-D kernel combination and inversion-grid projection (satt_ker3d.hip)
struct Combine {
  int nperiod, abr, rho_scaling;
  size_t n2, n3;  // nx*ny, nx*ny*nz
  const double *adj_s, *adj_den;              // (nperiod, nx, ny)
  const double *sen_vs, *sen_vp, *sen_rho;    // (nperiod, nx, ny, nz)
  const double *vs3d, *vp3d;                  // (nx, ny, nz)
  double *ker, *ker_den;                      // (3 or 1, nx, ny, nz)
};
struct InvGrid {
  int nx, ny, nz, nxi, nyi, nzi, nset;
  const double *vx, *vy, *vz;           // model axes
  const double *xinv, *yinv, *zinv;     // (n?i, nset)
};
void launch_combine(hipStream_t s, const Combine &C);
// scratch: inv_grid_scratch_bytes(G) bytes (cube and terms of every point)
size_t inv_grid_scratch_bytes(const InvGrid &G);
void launch_inv_grid(hipStream_t s, const InvGrid &G, const double *adj, double *gk,
                     void *scratch);
void launch_inv2fwd(hipStream_t s, const InvGrid &G, const double *gk, double *update);
void launch_component(hipStream_t s, size_t n, int nker, int q, const double *ker,
                      const double *kd, double *kc, double *kdc);
void launch_precond(hipStream_t s, size_t n, int nker, int q, const double *kdc,
                    const double *mx, double coe, double *ker, double *kc);
// the three kernels (vs, vp, rho) consecutive in `in` and `out`, each
// (n3 = nperiod*nx*ny, nz)
void launch_correct_depth(hipStream_t s, size_t n3, int nz, const double *in, double *out,
                          const double *dip, const double *depz);
void launch_grad_update(hipStream_t s, size_t n, int nker, int q, const double *update,
                        int nset, double chi0, double weight, double *grad);

// ---- surface-wave dispersion per column (satt_disp.hip)
struct DispJob {
  int ncol, nz, kmax, iwave, igr;
  int refine;              // 0: fwdsurf1d's nz layers; 1: refinegrid's nz + 1
                           // (depthkernel1d, src/surfker/surfker.f90:285-307)
  const double *vs;        // [nz][ncol]: the 3-D model vs3d(nx,ny,nz)
  const double *depz;      // [nz] km
  const double *periods;   // [kmax] s
  double *out;             // [col][kmax]: period-major (kmax, nx, ny)
  int *status;             // [col]
};
// Streaming read and copy bandwidth of the device (satt_probe.hip), GB/s
int stream_probe(hipStream_t s, int cus, size_t nbytes, int reps, double &read_gbs,
                 double &copy_gbs);
int lds_probe(hipStream_t s, int cus, int reps, double &read_gbs);
// LDS of one column's wave (the layered model and the per-layer terms) for
// n layers
size_t fwdsurf_lds_bytes(int n);
void launch_fwdsurf(hipStream_t s, const DispJob &J);
// Rayleigh-wave depth kernels (satt_depth.hip): one thread per (column,
// period), sregn96 at the column's phase velocity cp [col][kmax] on the
// refinegrid model; sen_vs/vp/rho (kmax, ncol, nz) period-major, status per
// column (kOk or the dispersion search's failure, which zeroes its kernels).
struct DepthJob {
  int ncol, nz, kmax;
  const double *vs, *depz, *periods, *cp;
  const int *col_status;
  double *sen_vs, *sen_vp, *sen_rho;
  double *work;            // depth_work_doubles(ncol*kmax, nz + 1)
};
size_t depth_work_doubles(size_t nthreads, int mmax);
void launch_depthkernel(hipStream_t s, const DepthJob &J);

}  // namespace satt
