// HIP kernels of the SurfATT forward + adjoint hot path for gfx950.
//
// Exactness: the sweeps keep the reference's Gauss-Seidel order.  A sweep in
// direction (xd, yd) updates cell (i,j) after (i-xd,j) and (i,j-yd) and before
// (i+xd,j) and (i,j+yd) (src/eikon_solver_2d.f90:906-907, :1389-1390).  With
// ii = xd>0 ? i : nx-1-i and jj = yd>0 ? j : ny-1-j, every cell of the
// anti-diagonal ii+jj = k reads the same neighbour values as in the
// sequential loop: the k-1 diagonal is already updated, the k+1 diagonal not
// yet.  So one workgroup steps k = 0 .. nx+ny-2 with a barrier between
// diagonals and updates a diagonal's cells in parallel, giving the sequential
// result bit for bit.  All arithmetic is FP64 in the reference's evaluation
// order; the library is built with -ffp-contract=off so no product is fused
// into an add.
#include <climits>

#include <algorithm>

#include "satt_internal.h"

namespace satt {
namespace {

constexpr int kBS = 256;  // threads per workgroup for the per-item kernels

__device__ __forceinline__ double deg2rad(double d) { return d / 180.0 * kPiF; }

// r = min(1.0, v) assigned to an implicit REAL(4) (src/eikon_solver_2d.f90:829)
__device__ __forceinline__ float weight_r4(double v) {
  return (float)(v < 1.0 ? v : 1.0);
}

__device__ __forceinline__ void block_sum_max(double &s, double &m, double *sh) {
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off, 64);
    const double o = __shfl_down(m, off, 64);
    m = o > m ? o : m;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    sh[2 * wid] = s;
    sh[2 * wid + 1] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = 0, M = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      S += sh[2 * w];
      M = sh[2 * w + 1] > M ? sh[2 * w + 1] : M;
    }
    sh[2 * 32] = S;
    sh[2 * 32 + 1] = M;
  }
  __syncthreads();
  s = sh[2 * 32];
  m = sh[2 * 32 + 1];
}

// Source slowness by bilinear weights in REAL(4) (src/eikon_solver_2d.f90:825-841).
__device__ int fwd_fun0(const Grid &g, const double *fun, double x0, double y0,
                        double *fun0) {
  const int nx = g.nx, ny = g.ny;
  const int idx0 = (int)floor((x0 - g.xx[0]) / g.dx + 1);
  const int idy0 = (int)floor((y0 - g.yy[0]) / g.dy + 1);
  if (idx0 < 1 || idy0 < 1 || idx0 + 1 > nx || idy0 + 1 > ny) return kESrcOnBoundary;
  const int i0 = idx0 - 1, j0 = idy0 - 1;
  const float r1 = weight_r4((x0 - g.xx[i0]) / g.dx);
  const float r2 = weight_r4((y0 - g.yy[j0]) / g.dy);
  const float w00 = (1.0f - r1) * (1.0f - r2);
  const float w01 = (1.0f - r1) * r2;
  const float w10 = r1 * (1.0f - r2);
  const float w11 = r1 * r2;
  *fun0 = (double)w00 * fun[i0 + (size_t)nx * j0] +
          (double)w01 * fun[i0 + (size_t)nx * (j0 + 1)] +
          (double)w10 * fun[i0 + 1 + (size_t)nx * j0] +
          (double)w11 * fun[i0 + 1 + (size_t)nx * (j0 + 1)];
  return kOk;
}

// T0, T0x, T0y of one cell (src/eikon_solver_2d.f90:845-857).
__device__ __forceinline__ void t0_cell(double xxi, double yyj, double x0,
                                        double sy0, double cy0, double fun0,
                                        double &T0, double &T0x, double &T0y) {
  const double sy = sin(yyj), cy = cos(yyj);
  const double sdx = sin(xxi - x0), cdx = cos(xxi - x0);
  const double t = sy * sy0 + cy * cy0 * cdx;
  double tc = t < 1.0 ? t : 1.0;
  tc = tc > -1.0 ? tc : -1.0;
  T0 = acos(tc) * kREarth * fun0;
  if (T0 == 0) {
    T0x = 0;
    T0y = 0;
  } else {
    const double s = sqrt(1 - t * t);
    T0x = -(1.0 / s * cy * cy0 * (-1.0) * sdx * kREarth * fun0);
    T0y = -(1.0 / s * (cy * sy0 - sy * cy0 * cdx) * kREarth * fun0);
  }
}

// One cell of the forward sweep body (src/eikon_solver_2d.f90:910-1079):
// four two-sided quadrant solves with the causality test, four one-sided
// candidates, then tau = min(tau, candidates).  Returns the new tau.
__device__ double fwd_cell(int i, int j, int nx, int ny, double dx, double dy,
                           double ak, double bk, double ck, double fk,
                           double t0, double t0x, double t0y,
                           const double *__restrict__ tau,
                           const double *__restrict__ T0) {
  const size_t k = i + (size_t)nx * j;
  double tnew = tau[k];
  double px, qx, py, qy;
#pragma unroll
  for (int icase = 1; icase <= 4; ++icase) {
    if (icase == 1) {
      if (i == 0 || j == 0) continue;
      px = (t0x + t0 / dx); qx = -(t0 / dx * tau[k - 1]);
      py = (t0y + t0 / dy); qy = -(t0 / dy * tau[k - nx]);
    } else if (icase == 2) {
      if (i == 0 || j == ny - 1) continue;
      px = (t0x + t0 / dx); qx = -(t0 / dx * tau[k - 1]);
      py = (t0y - t0 / dy); qy = t0 / dy * tau[k + nx];
    } else if (icase == 3) {
      if (i == nx - 1 || j == 0) continue;
      px = (t0x - t0 / dx); qx = +(t0 / dx * tau[k + 1]);
      py = (t0y + t0 / dy); qy = -(t0 / dy * tau[k - nx]);
    } else {
      if (i == nx - 1 || j == ny - 1) continue;
      px = (t0x - t0 / dx); qx = t0 / dx * tau[k + 1];
      py = (t0y - t0 / dy); qy = t0 / dy * tau[k + nx];
    }
    const double ea = ak * (px * px) + bk * (py * py) - 2 * ck * px * py;
    const double eb = 2 * ak * px * qx + 2 * bk * py * qy - 2 * ck * (px * qy + py * qx);
    const double ec = ak * (qx * qx) + bk * (qy * qy) - 2 * ck * qx * qy - fk * fk;
    const double delta = eb * eb - 4 * ea * ec;
    if (delta >= 0) {
      const double sq = sqrt(delta);
#pragma unroll
      for (int isol = 1; isol <= 2; ++isol) {
        const double tt = isol == 1 ? (-eb + sq) / (2 * ea) : (-eb - sq) / (2 * ea);
        const double Tx = px * tt + qx;
        const double Ty = py * tt + qy;
        const double chx = ak * Tx - ck * Ty;
        const double chy = bk * Ty - ck * Tx;
        bool ok;
        if (icase == 1) ok = (chx >= 0 && chy >= 0);
        else if (icase == 2) ok = (chx >= 0 && chy <= 0);
        else if (icase == 3) ok = (chx <= 0 && chy >= 0);
        else ok = (chx <= 0 && chy <= 0);
        if (ok && tt < tnew) tnew = tt;
      }
    }
  }
#pragma unroll
  for (int icase = 1; icase <= 4; ++icase) {
    size_t kn;
    double p, q, dis;
    if (icase == 1) {
      if (i == 0) continue;
      kn = k - 1;
      p = (t0x + t0 / dx); q = -(t0 / dx * tau[kn]);
      dis = sqrt(fk * fk * (bk / (ak * bk - ck * ck)));
    } else if (icase == 2) {
      if (i == nx - 1) continue;
      kn = k + 1;
      p = (t0x - t0 / dx); q = t0 / dx * tau[kn];
      dis = sqrt(fk * fk * (bk / (ak * bk - ck * ck)));
    } else if (icase == 3) {
      if (j == 0) continue;
      kn = k - nx;
      p = (t0y + t0 / dy); q = -(t0 / dy * tau[kn]);
      dis = sqrt(fk * fk * (ak / (ak * bk - ck * ck)));
    } else {
      if (j == ny - 1) continue;
      kn = k + nx;
      p = (t0y - t0 / dy); q = t0 / dy * tau[kn];
      dis = sqrt(fk * fk * (ak / (ak * bk - ck * ck)));
    }
    const double tn = tau[kn], t0n = T0[kn];
    double tt = (dis - q) / p;
    if (tt * t0 >= tn * t0n && tt > tn * 0.5 && tt < tnew) tnew = tt;
    tt = (-dis - q) / p;
    if (tt * t0 >= tn * t0n && tt > tn * 0.5 && tt < tnew) tnew = tt;
  }
  return tnew;
}

// ------------------------------------------------------------------ forward

__global__ __launch_bounds__(kBS) void k_forward(Grid g, PeriodFields pf, Items it,
                                                 FwdWork w, int *iters, int *status,
                                                 int iter_stride) {
  const int slot = blockIdx.x;
  const int item = it.item0 + slot;
  const int nx = g.nx, ny = g.ny, tid = threadIdx.x;
  const size_t N = (size_t)nx * ny;
  const int p = it.pidx[item];
  const double *__restrict__ a = pf.a + p * N;
  const double *__restrict__ b = pf.b + p * N;
  const double *__restrict__ c = pf.c + p * N;
  const double *__restrict__ fun = pf.fun + p * N;
  double *tau = w.tau + slot * N, *tau_old = w.tau_old + slot * N;
  double *T0 = w.T0 + slot * N, *T0x = w.T0x + slot * N, *T0y = w.T0y + slot * N;
  double *T = w.T + slot * N;
  __shared__ int s_fr[4];
  __shared__ int s_done;
  __shared__ double s_red[2 * 33];

  const double x0 = deg2rad(it.srcx[item]), y0 = deg2rad(it.srcy[item]);
  if (tid == 0) {
    s_fr[0] = INT_MAX; s_fr[1] = -1; s_fr[2] = INT_MAX; s_fr[3] = -1;
  }
  __syncthreads();
  double fun0 = 0;
  const int st0 = fwd_fun0(g, fun, x0, y0, &fun0);
  if (st0 != kOk) {
    if (tid == 0) { status[item] = st0; iters[item * iter_stride] = 0; }
    return;
  }
  // frozen set (src/eikon_solver_2d.f90:861): the columns with
  // abs((xx-x0)/dx)<=1 form one range because xx is monotone; same for rows.
  for (int i = tid; i < nx; i += kBS)
    if (fabs((g.xx[i] - x0) / g.dx) <= 1) { atomicMin(&s_fr[0], i); atomicMax(&s_fr[1], i); }
  for (int j = tid; j < ny; j += kBS)
    if (fabs((g.yy[j] - y0) / g.dy) <= 1) { atomicMin(&s_fr[2], j); atomicMax(&s_fr[3], j); }
  __syncthreads();
  const int fi0 = s_fr[0], fi1 = s_fr[1], fj0 = s_fr[2], fj1 = s_fr[3];
  if (fi0 <= fi1 && fj0 <= fj1 && (fi0 == 0 || fi1 == nx - 1 || fj0 == 0 || fj1 == ny - 1)) {
    if (tid == 0) { status[item] = kESrcOnBoundary; iters[item * iter_stride] = 0; }
    return;
  }
  const double sy0 = sin(y0), cy0 = cos(y0);
  for (size_t k = tid; k < N; k += kBS) {
    const int i = (int)(k % nx), j = (int)(k / nx);
    double t0, t0x, t0y;
    t0_cell(g.xx[i], g.yy[j], x0, sy0, cy0, fun0, t0, t0x, t0y);
    T0[k] = t0; T0x[k] = t0x; T0y[k] = t0y;
    const bool frozen = i >= fi0 && i <= fi1 && j >= fj0 && j <= fj1;
    tau[k] = frozen ? 1.0 : 10.0;
  }
  __syncthreads();

  const double dx = g.dx, dy = g.dy;
  int iter = 1;
  bool conv = false;
  for (; iter <= kFwdMaxIter; ++iter) {
    for (size_t k = tid; k < N; k += kBS) tau_old[k] = tau[k];
    __syncthreads();
    for (int sw = 0; sw < 4; ++sw) {
      const int xd = sw < 2 ? 1 : -1;
      const int yd = (sw == 0 || sw == 2) ? 1 : -1;
      for (int d = 0; d <= nx + ny - 2; ++d) {
        const int ilo = d - (ny - 1) > 0 ? d - (ny - 1) : 0;
        const int ihi = d < nx - 1 ? d : nx - 1;
        for (int ii = ilo + tid; ii <= ihi; ii += kBS) {
          const int jj = d - ii;
          const int i = xd > 0 ? ii : nx - 1 - ii;
          const int j = yd > 0 ? jj : ny - 1 - jj;
          if (i >= fi0 && i <= fi1 && j >= fj0 && j <= fj1) continue;
          const size_t k = i + (size_t)nx * j;
          tau[k] = fwd_cell(i, j, nx, ny, dx, dy, a[k], b[k], c[k], fun[k], T0[k],
                            T0x[k], T0y[k], tau, T0);
        }
        __syncthreads();
      }
    }
    double s = 0, m = 0;
    for (size_t k = tid; k < N; k += kBS) {
      const double dd = fabs(tau[k] - tau_old[k]) * T0[k];
      s += dd;
      m = dd > m ? dd : m;
    }
    block_sum_max(s, m, s_red);
    if (tid == 0) {
      const double l1 = s / (double)nx / (double)ny;
      s_done = (fabs(l1) < kFwdTol && m < kFwdTol);
    }
    __syncthreads();
    if (s_done) { conv = true; break; }
  }
  if (!conv) iter = kFwdMaxIter;
  for (size_t k = tid; k < N; k += kBS) T[k] = tau[k] * T0[k];
  if (tid == 0) {
    status[item] = conv ? kOk : kWMaxIter;
    iters[item * iter_stride] = iter;
  }
}

// ---------------------------------------------------------------- receivers

// Masked minloc / maxloc of src/shared/utils.f90:925-928, first index on ties.
__device__ int minloc_ge(const double *x, int n, double q) {
  int best = -1;
  double bv = 0;
  for (int k = 0; k < n; ++k)
    if (q >= x[k]) {
      const double v = q - x[k];
      if (best < 0 || v < bv) { best = k; bv = v; }
    }
  return best;
}
__device__ int maxloc_lt(const double *x, int n, double q) {
  int best = -1;
  double bv = 0;
  for (int k = 0; k < n; ++k)
    if (q < x[k]) {
      const double v = q - x[k];
      if (best < 0 || v > bv) { best = k; bv = v; }
    }
  return best;
}

// interp2_0_dp (src/shared/utils.f90:917-942); returns false out of the grid.
__device__ bool interp2_point(const double *x, const double *y, int nx, int ny,
                              const double *v, double xq, double yq, double *out) {
  const int x1 = minloc_ge(x, nx, xq), y1 = minloc_ge(y, ny, yq);
  const int x2 = maxloc_lt(x, nx, xq), y2 = maxloc_lt(y, ny, yq);
  if (x1 < 0 || y1 < 0 || x2 < 0 || y2 < 0) return false;
  const double vn = fabs((x[x2] - x[x1]) * (y[y2] - y[y1]));
  const double xr0 = x[x1], xr1 = x[x2], yr0 = y[y1], yr1 = y[y2];
  const double N0 = fabs((xr1 - xq) * (yr1 - yq));
  const double N1 = fabs((xr0 - xq) * (yr1 - yq));
  const double N2 = fabs((xr1 - xq) * (yr0 - yq));
  const double N3 = fabs((xr0 - xq) * (yr0 - yq));
  const size_t sx = (size_t)nx;
  double s = 0;
  s = s + v[x1 + sx * y1] * (N0 / vn);
  s = s + v[x2 + sx * y1] * (N1 / vn);
  s = s + v[x1 + sx * y2] * (N2 / vn);
  s = s + v[x2 + sx * y2] * (N3 / vn);
  *out = s;
  return true;
}

// Receiver interpolation, misfit, outlier weights and adjoint sources of one
// item (src/measadj.f90:71-72, :82-88, :102, :111-113); one block per slot.
__global__ __launch_bounds__(kBS) void k_receivers(const SlotDesc *sd, int istotable) {
  const SlotDesc D = sd[blockIdx.x];
  const Grid &g = D.g;
  const int nx = g.nx, ny = g.ny, tid = threadIdx.x;
  const int nr = D.nr;
  __shared__ int s_bad;
  if (tid == 0) s_bad = 0;
  __syncthreads();
  if (*D.status < 0) return;  // forward failed: nothing to measure
  for (int r = tid; r < nr; r += kBS) {
    double v = 0;
    if (!interp2_point(g.xd, g.yd, nx, ny, D.T, D.recx[r], D.recy[r], &v)) s_bad = 1;
    D.tekio[r] = v;
  }
  __syncthreads();
  if (s_bad) {
    if (tid == 0) *D.status = kERecOutOfGrid;
    return;
  }
  if (tid == 0) {
    double s = 0;
    for (int r = 0; r < nr; ++r) {
      const double d = D.tekio[r] - D.tt[r];
      s = s + D.weight[r] * (d * d);
    }
    *D.item_chi = 0.5 * s;
  }
  for (int r = tid; r < nr; r += kBS) {
    const double tek = D.tekio[r], tobs = D.tt[r];
    double wt = D.weight[r];
    if (fabs(tek - tobs) / tobs > 0.5) wt = 0.;
    D.weight_out[r] = wt;
    D.rhs_adj[r] = wt * (tek - tobs);
    D.rhs_den[r] = -1.0 * wt;
    if (istotable) D.tt_fwd[D.idx_path[r]] = tek;
  }
}

// ------------------------------------------------------------------ adjoint

// Delta source of src/eikon_solver_2d.f90:1320-1336, receivers in order.
__device__ int adj_splat(const Grid &g, const double *xr_deg, const double *yr_deg,
                         const double *rhs, int nr, double *delta) {
  const int nx = g.nx, ny = g.ny;
  for (int r = 0; r < nr; ++r) {
    const double xr = deg2rad(xr_deg[r]), yr = deg2rad(yr_deg[r]);
    const int idx0 = (int)floor((xr - g.xx[0]) / g.dx + 1);
    const int idy0 = (int)floor((yr - g.yy[0]) / g.dy + 1);
    if (idx0 < 1 || idy0 < 1 || idx0 + 1 > nx || idy0 + 1 > ny) return kERecOutOfGrid;
    const int i0 = idx0 - 1, j0 = idy0 - 1;
    const float r1 = weight_r4((xr - g.xx[i0]) / g.dx);
    const float r2 = weight_r4((yr - g.yy[j0]) / g.dy);
    const double sa = rhs[r];
    const double den = g.dx * kREarth * cos(yr) * g.dy * kREarth;
    const size_t k00 = i0 + (size_t)nx * j0, k01 = k00 + nx, k10 = k00 + 1, k11 = k00 + nx + 1;
    delta[k00] = delta[k00] + sa * (double)(1.0f - r1) * (double)(1.0f - r2) / den;
    delta[k01] = delta[k01] + sa * (double)(1.0f - r1) * (double)r2 / den;
    delta[k10] = delta[k10] + sa * (double)r1 * (double)(1.0f - r2) / den;
    delta[k11] = delta[k11] + sa * (double)r1 * (double)r2 / den;
  }
  return kOk;
}

__global__ __launch_bounds__(kBS) void k_adjoint(Grid g, PeriodFields pf, Items it,
                                                 AdjWork w, int *iters, int *status,
                                                 int iter_stride, int apply_mask) {
  const int slot = blockIdx.x;
  const int item = it.item0 + slot;
  const int nx = g.nx, ny = g.ny, tid = threadIdx.x;
  const size_t N = (size_t)nx * ny;
  if (status[item] < 0) return;  // an earlier stage failed
  const int p = it.pidx[item];
  const double *__restrict__ a = pf.a + p * N;
  const double *__restrict__ b = pf.b + p * N;
  const double *__restrict__ c = pf.c + p * N;
  const double *__restrict__ T = w.T + slot * N;
  double *Ta = w.Ta + slot * N;
  double *A1 = w.coef + slot * 5 * N, *A2 = A1 + N, *B1 = A1 + 2 * N, *B2 = A1 + 3 * N,
         *D = A1 + 4 * N;
  double *delta = w.delta + slot * N, *old = w.old + slot * N;
  __shared__ int s_err, s_done;
  __shared__ double s_red[2 * 33];
  const double dx = g.dx, dy = g.dy;

  for (size_t k = tid; k < N; k += kBS) delta[k] = 0;
  __syncthreads();
  if (tid == 0) {
    const int r0 = it.rec_off[item], nr = it.rec_off[item + 1] - r0;
    s_err = adj_splat(g, it.recx + r0, it.recy + r0, w.rhs + r0, nr, delta);
  }
  __syncthreads();
  if (s_err != kOk) {
    if (tid == 0) { status[item] = s_err; iters[item * iter_stride] = 0; }
    return;
  }
  // face coefficients (src/eikon_solver_2d.f90:1348-1370); Ta=100 inside,
  // 0 on the boundary (:1373-1380).  d is precomputed with the sweep's own
  // expression (:1392) and stored as 0 where abs(d)<eps (:1394-1395).
  for (size_t k = tid; k < N; k += kBS) {
    const int i = (int)(k % nx), j = (int)(k / nx);
    if (i == 0 || j == 0 || i == nx - 1 || j == ny - 1) {
      Ta[k] = 0;
      continue;
    }
    const size_t W = k - 1, E = k + 1, S = k - nx, Nn = k + nx;
    const double a1 = -((T[k] - T[W]) / dx * (a[k] + a[W]) / 2) +
                      (T[Nn] - T[S] + T[W + nx] - T[W - nx]) / (4 * dy) * (c[k] + c[W]) / 2;
    const double a2 = -((T[E] - T[k]) / dx * (a[E] + a[k]) / 2) +
                      (T[E + nx] - T[E - nx] + T[Nn] - T[S]) / (4 * dy) * (c[E] + c[k]) / 2;
    const double b1 = -((T[k] - T[S]) / dy * (b[k] + b[S]) / 2) +
                      (T[E] - T[W] + T[E - nx] - T[W - nx]) / (4 * dx) * (c[k] + c[S]) / 2;
    const double b2 = -((T[Nn] - T[k]) / dy * (b[Nn] + b[k]) / 2) +
                      (T[E + nx] - T[W + nx] + T[E] - T[W]) / (4 * dx) * (c[Nn] + c[k]) / 2;
    A1[k] = a1; A2[k] = a2; B1[k] = b1; B2[k] = b2;
    const double a1m = (a1 - fabs(a1)) / 2, a2p = (a2 + fabs(a2)) / 2;
    const double b1m = (b1 - fabs(b1)) / 2, b2p = (b2 + fabs(b2)) / 2;
    const double d = (a2p - a1m) / dx + (b2p - b1m) / dy;
    D[k] = fabs(d) < kAdjEps ? 0.0 : d;
    Ta[k] = 100;
  }
  __syncthreads();

  int iter = 1;
  bool conv = false;
  for (; iter <= kAdjMaxIter; ++iter) {
    for (size_t k = tid; k < N; k += kBS) old[k] = Ta[k];
    __syncthreads();
    // (xdirec, ydirec) = (-1,-1), (-1,1), (1,-1), (1,1); xdirec=-1 runs x
    // ascending over 2..nx-1 (src/eikon_solver_2d.f90:1385-1390)
    for (int sw = 0; sw < 4; ++sw) {
      const int xasc = sw < 2, yasc = (sw == 0 || sw == 2);
      for (int dsum = 2; dsum <= nx + ny - 4; ++dsum) {
        const int ilo = dsum - (ny - 2) > 1 ? dsum - (ny - 2) : 1;
        const int ihi = dsum - 1 < nx - 2 ? dsum - 1 : nx - 2;
        for (int ii = ilo + tid; ii <= ihi; ii += kBS) {
          const int jj = dsum - ii;
          const int i = xasc ? ii : nx - 1 - ii;
          const int j = yasc ? jj : ny - 1 - jj;
          const size_t k = i + (size_t)nx * j;
          const double dk = D[k];
          if (dk == 0.0) {
            Ta[k] = 0;
          } else {
            const double a1 = A1[k], a2 = A2[k], b1 = B1[k], b2 = B2[k];
            const double a1p = (a1 + fabs(a1)) / 2, a2m = (a2 - fabs(a2)) / 2;
            const double b1p = (b1 + fabs(b1)) / 2, b2m = (b2 - fabs(b2)) / 2;
            const double e = (Ta[k - 1] * a1p - Ta[k + 1] * a2m) / dx +
                             (Ta[k - nx] * b1p - Ta[k + nx] * b2m) / dy;
            Ta[k] = (delta[k] + e) / dk;
          }
        }
        __syncthreads();
      }
    }
    double s = 0, m = 0;
    for (size_t k = tid; k < N; k += kBS) {
      const double dd = fabs(Ta[k] - old[k]);
      s += dd;
      m = dd > m ? dd : m;
    }
    block_sum_max(s, m, s_red);
    if (tid == 0) {
      const double l1 = s / (double)(nx * ny);
      s_done = (fabs(l1) < kAdjTol && m < kAdjTol);
    }
    __syncthreads();
    if (s_done) { conv = true; break; }
  }
  if (!conv) iter = kAdjMaxIter;
  if (apply_mask) {
    // mask (src/eikon_solver_2d.f90:1449-1467), in degrees
    const double x0 = it.srcx[item], y0 = it.srcy[item];
    const double dxd = g.xd[1] - g.xd[0], dyd = g.yd[1] - g.yd[0];
    for (size_t k = tid; k < N; k += kBS) {
      const int i = (int)(k % nx), j = (int)(k / nx);
      if (fabs(g.xd[i] - x0) < 1.5 * dxd && fabs(g.yd[j] - y0) < 1.5 * dyd) Ta[k] = 0;
    }
  }
  if (tid == 0) {
    if (!conv && status[item] == kOk) status[item] = kWMaxIter;
    iters[item * iter_stride] = iter;
  }
}

// ------------------------------------------------------------ small kernels

// Per-period fields from the period-major model (src/grid.f90:25-34):
// vel, fun = 1./vel (src/measadj.f90:67), and the metric of the call site
// (m11, m22, -m12) converted as src/eikon_solver_2d.f90:817-821.
__global__ void k_period_fields(Grid g, int np, const double *svel, const double *m11,
                                const double *m12, const double *m22, double *a,
                                double *b, double *c, double *fun, double *vel) {
  const size_t N = (size_t)g.nx * g.ny;
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (k >= N) return;
  const size_t src = (size_t)p + (size_t)np * k, dst = (size_t)p * N + k;
  const int j = (int)(k / g.nx);
  const double cy = cos(g.yy[j]);
  const double r2 = kREarth * kREarth;
  const double v = svel[src];
  vel[dst] = v;
  fun[dst] = 1.0 / v;
  a[dst] = m11[src] / (r2 * (cy * cy));
  b[dst] = m22[src] / r2;
  c[dst] = (-m12[src]) / (r2 * cy);
}

__global__ void k_metric_fields(Grid g, const double *spha, const double *sphb,
                                const double *sphc, double *a, double *b, double *c) {
  const size_t N = (size_t)g.nx * g.ny;
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int j = (int)(k / g.nx);
  const double cy = cos(g.yy[j]);
  const double r2 = kREarth * kREarth;
  a[k] = spha[k] / (r2 * (cy * cy));
  b[k] = sphb[k] / r2;
  c[k] = sphc[k] / (r2 * cy);
}

// adj_s(pidx,:,:) += adj / svel**power, items of each period in item order
// (src/tomo2d.f90:155, src/acqui.f90:314-315).
__global__ void k_stack(Grid g, PeriodFields pf, const int *per_off, const int *per_slot,
                        const SlotDesc *sd, int which, double *acc, int power) {
  const size_t N = (size_t)g.nx * g.ny;
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (k >= N) return;
  const double v = pf.vel[p * N + k];
  const double den = power == 3 ? v * v * v : v * v;
  double s = acc[p * N + k];
  for (int q = per_off[p]; q < per_off[p + 1]; ++q) {
    const SlotDesc D = sd[per_slot[q]];
    s = s + (which ? D.Td : D.Ta)[k] / den;
  }
  acc[p * N + k] = s;
}

// The same accumulation for one item (one solve per launch).
__global__ void k_stack_item(Grid g, PeriodFields pf, const int *pidx, int item,
                             const double *field, double *acc, int power) {
  const size_t N = (size_t)g.nx * g.ny;
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int p = pidx[item];
  const double v = pf.vel[p * N + k];
  const double den = power == 3 ? v * v * v : v * v;
  acc[p * N + k] = acc[p * N + k] + field[k] / den;
}

// Zero n 32-bit words (a kernel rather than a memset, so that a captured
// launch chain holds kernels only).
__global__ void k_zero_words(unsigned int *p, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    p[k] = 0u;
}

__global__ void k_to_period_major(int N, int np, const double *nat, double *pm) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (k >= (size_t)N) return;
  pm[(size_t)p + (size_t)np * k] = nat[(size_t)p * N + k];
}

__global__ void k_mask(Grid g, double *table, double x0, double y0) {
  const size_t N = (size_t)g.nx * g.ny;
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int i = (int)(k % g.nx), j = (int)(k / g.nx);
  const double dx = g.xd[1] - g.xd[0], dy = g.yd[1] - g.yd[0];
  if (fabs(g.xd[i] - x0) < 1.5 * dx && fabs(g.yd[j] - y0) < 1.5 * dy) table[k] = 0;
}

__global__ void k_interp2(Grid g, const double *v, const double *xq, const double *yq,
                          int nq, double *vq, int *status) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  double out = 0;
  if (!interp2_point(g.xd, g.yd, g.nx, g.ny, v, xq[q], yq[q], &out))
    atomicMin(status, kERecOutOfGrid);
  vq[q] = out;
}

__global__ void k_fwd_setup_fields(Grid g, const double *fun, double x0d, double y0d,
                                   double *T0, double *T0x, double *T0y, int *status) {
  const size_t N = (size_t)g.nx * g.ny;
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const double x0 = deg2rad(x0d), y0 = deg2rad(y0d);
  double fun0 = 0;
  const int st = fwd_fun0(g, fun, x0, y0, &fun0);
  if (k >= N) return;
  if (st != kOk) {
    if (k == 0) *status = st;
    return;
  }
  const int i = (int)(k % g.nx), j = (int)(k / g.nx);
  double t0, t0x, t0y;
  t0_cell(g.xx[i], g.yy[j], x0, sin(y0), cos(y0), fun0, t0, t0x, t0y);
  T0[k] = t0; T0x[k] = t0x; T0y[k] = t0y;
}

__global__ void k_adj_delta(Grid g, const double *xr, const double *yr, const double *rhs,
                            int nr, double *delta, int *status) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *status = adj_splat(g, xr, yr, rhs, nr, delta);
}

inline dim3 cells_grid(size_t N, int np = 1) {
  return dim3((unsigned)((N + 255) / 256), (unsigned)np);
}

}  // namespace

// ------------------------------------------------------------------ launches

void launch_period_fields(hipStream_t s, const Grid &g, int np, const double *svel,
                          const double *m11, const double *m12, const double *m22,
                          double *a, double *b, double *c, double *fun, double *vel) {
  hipLaunchKernelGGL(k_period_fields, cells_grid((size_t)g.nx * g.ny, np), dim3(256), 0, s,
                     g, np, svel, m11, m12, m22, a, b, c, fun, vel);
}

void launch_metric_fields(hipStream_t s, const Grid &g, const double *spha,
                          const double *sphb, const double *sphc, double *a, double *b,
                          double *c) {
  hipLaunchKernelGGL(k_metric_fields, cells_grid((size_t)g.nx * g.ny), dim3(256), 0, s, g,
                     spha, sphb, sphc, a, b, c);
}

void launch_forward(hipStream_t s, const Grid &g, const PeriodFields &pf, const Items &it,
                    const FwdWork &w, int *iters, int *status, int iter_stride) {
  hipLaunchKernelGGL(k_forward, dim3(it.count), dim3(kBS), 0, s, g, pf, it, w, iters, status,
                     iter_stride);
}

void launch_receivers(hipStream_t s, const SlotDesc *sd, int count, int istotable) {
  hipLaunchKernelGGL(k_receivers, dim3(count), dim3(kBS), 0, s, sd, istotable);
}

void launch_adjoint(hipStream_t s, const Grid &g, const PeriodFields &pf, const Items &it,
                    const AdjWork &w, int *iters, int *status, int iter_stride,
                    int apply_mask) {
  hipLaunchKernelGGL(k_adjoint, dim3(it.count), dim3(kBS), 0, s, g, pf, it, w, iters, status,
                     iter_stride, apply_mask);
}

void launch_stack(hipStream_t s, const Grid &g, const PeriodFields &pf, int np,
                  const int *per_off, const int *per_slot, const SlotDesc *sd, int which,
                  double *acc, int power) {
  const size_t N = (size_t)g.nx * g.ny;
  hipLaunchKernelGGL(k_stack, dim3((unsigned)((N + 255) / 256), np), dim3(256), 0, s, g, pf,
                     per_off, per_slot, sd, which, acc, power);
}

void launch_stack_item(hipStream_t s, const Grid &g, const PeriodFields &pf, const int *pidx,
                       int item, const double *field, double *acc, int power) {
  hipLaunchKernelGGL(k_stack_item, cells_grid((size_t)g.nx * g.ny), dim3(256), 0, s, g, pf, pidx,
                     item, field, acc, power);
}

void launch_zero(hipStream_t s, void *p, size_t bytes) {
  const size_t n = bytes / 4;
  if (n == 0) return;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_zero_words, dim3(blocks), dim3(256), 0, s,
                     static_cast<unsigned int *>(p), n);
}

void launch_to_period_major(hipStream_t s, int N, int np, const double *nat, double *pm) {
  hipLaunchKernelGGL(k_to_period_major, cells_grid((size_t)N, np), dim3(256), 0, s, N, np,
                     nat, pm);
}

void launch_mask(hipStream_t s, const Grid &g, double *table, double x0, double y0) {
  hipLaunchKernelGGL(k_mask, cells_grid((size_t)g.nx * g.ny), dim3(256), 0, s, g, table, x0,
                     y0);
}

void launch_interp2(hipStream_t s, const Grid &g, const double *v, const double *xq,
                    const double *yq, int nq, double *vq, int *status) {
  hipLaunchKernelGGL(k_interp2, dim3((nq + 255) / 256), dim3(256), 0, s, g, v, xq, yq, nq, vq,
                     status);
}

void launch_fwd_setup_fields(hipStream_t s, const Grid &g, const double *a, const double *fun,
                             double x0d, double y0d, double *T0, double *T0x, double *T0y,
                             int *status) {
  (void)a;
  hipLaunchKernelGGL(k_fwd_setup_fields, cells_grid((size_t)g.nx * g.ny), dim3(256), 0, s, g,
                     fun, x0d, y0d, T0, T0x, T0y, status);
}

void launch_adj_delta(hipStream_t s, const Grid &g, const double *xr, const double *yr,
                      const double *rhs, int nr, double *delta, int *status) {
  hipLaunchKernelGGL(k_adj_delta, dim3(1), dim3(64), 0, s, g, xr, yr, rhs, nr, delta, status);
}

}  // namespace satt
