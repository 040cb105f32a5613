// Batched FP64 sweeps for gfx950: one wavefront per (source, period) item.
//
// Strip pipeline.  The grid is cut into strips of 64 columns; lane l of the
// wave owns column i = 64c+l of strip c.  For a sweep in direction (xd, yd),
// lane l processes its column's rows in order yd, lagging one step behind its
// upstream neighbour lane l-xd.  At step t the wave therefore works on cells
// (64c+l, j) with j+l (or j-l) constant: 64 mutually independent cells of one
// anti-diagonal.  Each cell reads
//   (i-xd, j)  new value : lane l-xd's result of step t-1 (register shuffle)
//   (i, j-yd)  new value : the lane's own result of step t-1
//   (i+xd, j)  old value : lane l+xd's prefetched next cell
//   (i, j+yd)  old value : the lane's own prefetched next cell
// which are exactly the values the sequential loop of
// src/eikon_solver_2d.f90:906-1079 (forward) and :1385-1405 (adjoint) reads,
// so the result is the sequential one bit for bit.  Strips run one after the
// other in direction xd; the strip edge lanes read the neighbour strip's
// column from memory (final values upstream, old values downstream).
// There is no barrier: one wave, in-order lanes.
//
// Layout.  tau / Ta live in natural layout (i + nx*j).  Read-only per-cell
// coefficients are stored twice, skewed so that the 64 cells of one step are
// one contiguous 512-byte row:
//   plus skew  (sweeps with xd == yd):  row r = j + l
//   minus skew (sweeps with xd != yd):  row r = j - l + 63
// each as [strip][row 0..ny+62][lane 0..63], zero padded.
#include <climits>

#include "satt_internal.h"
#include "satt_strip.h"

namespace satt {
namespace {

__device__ __forceinline__ double deg2rad_s(double d) { return d / 180.0 * kPiF; }
__device__ __forceinline__ float weight_r4_s(double v) { return (float)(v < 1.0 ? v : 1.0); }

__device__ __forceinline__ double shfl(double v, int src) { return __shfl(v, src, 64); }

__device__ __forceinline__ double wave_sum(double s) {
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  return s;
}
__device__ __forceinline__ double wave_max(double m) {
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  return m;
}
__device__ __forceinline__ int wave_min_i(int v) {
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Orders the wave's memory operations before the following ones (one
// wavefront per workgroup, so workgroup scope is the wave).
__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

__device__ __forceinline__ size_t skidx(int c, int r, int l, int R) {
  return ((size_t)c * R + r) * 64 + l;
}

// One cell of the forward sweep body (src/eikon_solver_2d.f90:910-1079) with
// every operand in registers.  The arithmetic and its order are those of the
// reference; dis_x = sqrt(fun**2*(b/(a*b-c**2))) and dis_y = sqrt(fun**2*
// (a/(a*b-c**2))) are per-cell constants precomputed by the same expression.
__device__ __forceinline__ double fwd_cell_r(int i, int j, int nx, int ny, double dx, double dy,
                                             double ak, double bk, double ck, double fk,
                                             double disx, double disy, double t0, double t0x,
                                             double t0y, double tc, double tW, double tE,
                                             double tS, double tN, double zW, double zE,
                                             double zS, double zN) {
  double tnew = tc;
  const double t0dx = t0 / dx, t0dy = t0 / dy;
#pragma unroll
  for (int icase = 1; icase <= 4; ++icase) {
    double px, qx, py, qy;
    if (icase == 1) {
      if (i == 0 || j == 0) continue;
      px = (t0x + t0dx); qx = -(t0dx * tW);
      py = (t0y + t0dy); qy = -(t0dy * tS);
    } else if (icase == 2) {
      if (i == 0 || j == ny - 1) continue;
      px = (t0x + t0dx); qx = -(t0dx * tW);
      py = (t0y - t0dy); qy = t0dy * tN;
    } else if (icase == 3) {
      if (i == nx - 1 || j == 0) continue;
      px = (t0x - t0dx); qx = +(t0dx * tE);
      py = (t0y + t0dy); qy = -(t0dy * tS);
    } else {
      if (i == nx - 1 || j == ny - 1) continue;
      px = (t0x - t0dx); qx = t0dx * tE;
      py = (t0y - t0dy); qy = t0dy * tN;
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
    double p, q, dis, tn, t0n;
    if (icase == 1) {
      if (i == 0) continue;
      p = (t0x + t0dx); q = -(t0dx * tW); dis = disx; tn = tW; t0n = zW;
    } else if (icase == 2) {
      if (i == nx - 1) continue;
      p = (t0x - t0dx); q = t0dx * tE; dis = disx; tn = tE; t0n = zE;
    } else if (icase == 3) {
      if (j == 0) continue;
      p = (t0y + t0dy); q = -(t0dy * tS); dis = disy; tn = tS; t0n = zS;
    } else {
      if (j == ny - 1) continue;
      p = (t0y - t0dy); q = t0dy * tN; dis = disy; tn = tN; t0n = zN;
    }
    double tt = (dis - q) / p;
    if (tt * t0 >= tn * t0n && tt > tn * 0.5 && tt < tnew) tnew = tt;
    tt = (-dis - q) / p;
    if (tt * t0 >= tn * t0n && tt > tn * 0.5 && tt < tnew) tnew = tt;
  }
  return tnew;
}

// ---------------------------------------------------------- period skews

// Per-period skewed fields: a, b, c, fun, dis_x, dis_y (kPF fields), both
// skews.  Thread = (strip, row, lane) of one skew of one period.
__global__ void k_period_skew(Grid g, SkewGeom sg, int np, const double *a, const double *b,
                              const double *c, const double *fun, double *out) {
  const size_t per_skew = sg.skew_elems();
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int sk = blockIdx.y, p = blockIdx.z;
  if (e >= per_skew) return;
  const int l = (int)(e & 63);
  const int r = (int)((e >> 6) % sg.R);
  const int cc = (int)((e >> 6) / sg.R);
  const int i = 64 * cc + l;
  const int j = sk == 0 ? r - l : r + l - 63;
  double v[kPF] = {0, 0, 0, 0, 0, 0};
  if (i < g.nx && j >= 0 && j < g.ny) {
    const size_t k = (size_t)p * g.nx * g.ny + i + (size_t)g.nx * j;
    const double ak = a[k], bk = b[k], ck = c[k], fk = fun[k];
    v[0] = ak; v[1] = bk; v[2] = ck; v[3] = fk;
    v[4] = sqrt(fk * fk * (bk / (ak * bk - ck * ck)));
    v[5] = sqrt(fk * fk * (ak / (ak * bk - ck * ck)));
  }
  double *base = out + ((size_t)p * 2 + sk) * kPF * per_skew;
#pragma unroll
  for (int f = 0; f < kPF; ++f) base[f * per_skew + e] = v[f];
}

// ------------------------------------------------------------- forward

__device__ int fwd_fun0_s(const Grid &g, const double *fun, double x0, double y0, double *fun0) {
  const int nx = g.nx, ny = g.ny;
  const int idx0 = (int)floor((x0 - g.xx[0]) / g.dx + 1);
  const int idy0 = (int)floor((y0 - g.yy[0]) / g.dy + 1);
  if (idx0 < 1 || idy0 < 1 || idx0 + 1 > nx || idy0 + 1 > ny) return kESrcOnBoundary;
  const int i0 = idx0 - 1, j0 = idy0 - 1;
  const float r1 = weight_r4_s((x0 - g.xx[i0]) / g.dx);
  const float r2 = weight_r4_s((y0 - g.yy[j0]) / g.dy);
  const float w00 = (1.0f - r1) * (1.0f - r2);
  const float w01 = (1.0f - r1) * r2;
  const float w10 = r1 * (1.0f - r2);
  const float w11 = r1 * r2;
  *fun0 = (double)w00 * fun[i0 + (size_t)nx * j0] + (double)w01 * fun[i0 + (size_t)nx * (j0 + 1)] +
          (double)w10 * fun[i0 + 1 + (size_t)nx * j0] +
          (double)w11 * fun[i0 + 1 + (size_t)nx * (j0 + 1)];
  return kOk;
}

// T0, T0x, T0y of one cell (src/eikon_solver_2d.f90:845-857).
__device__ __forceinline__ void t0_cell_s(double xxi, double yyj, double x0, double sy0,
                                          double cy0, double fun0, double &T0, double &T0x,
                                          double &T0y) {
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

struct FCoef {  // one cell's forward operands
  double a, b, c, f, dx_, dy_, t0, t0x, t0y, tau;
};

__global__ __launch_bounds__(64) void k_fwd_strip(Grid g, SkewGeom sg, const double *pskew,
                                                  const double *fun_nat, Items it,
                                                  StripFwdWork w, int *iters, int *status,
                                                  int iter_stride) {
  const int slot = blockIdx.x;
  const int item = it.item0 + slot;
  const int lane = threadIdx.x;
  const int nx = g.nx, ny = g.ny, R = sg.R, ns = sg.ns;
  const size_t N = (size_t)nx * ny, SE = sg.skew_elems();
  const int p = it.pidx[item];
  const double *PS[2] = {pskew + ((size_t)p * 2 + 0) * kPF * SE,
                         pskew + ((size_t)p * 2 + 1) * kPF * SE};
  double *IS[2] = {w.iskew + (size_t)slot * 2 * 3 * SE, w.iskew + (size_t)slot * 2 * 3 * SE + 3 * SE};
  double *tau = w.tau + (size_t)slot * N, *snap = w.snap + (size_t)slot * N;
  double *T = w.T + (size_t)slot * N;
  const double x0 = deg2rad_s(it.srcx[item]), y0 = deg2rad_s(it.srcy[item]);
  const double dx = g.dx, dy = g.dy;

  double fun0 = 0;
  const int st0 = fwd_fun0_s(g, fun_nat + (size_t)p * N, x0, y0, &fun0);
  if (st0 != kOk) {
    if (lane == 0) { status[item] = st0; iters[item * iter_stride] = 0; }
    return;
  }
  // frozen set (src/eikon_solver_2d.f90:861-872): a rectangle of columns and rows
  int fi0 = INT_MAX, fi1 = -1, fj0 = INT_MAX, fj1 = -1;
  for (int i = lane; i < nx; i += 64)
    if (fabs((g.xx[i] - x0) / dx) <= 1) { fi0 = min(fi0, i); fi1 = max(fi1, i); }
  for (int j = lane; j < ny; j += 64)
    if (fabs((g.yy[j] - y0) / dy) <= 1) { fj0 = min(fj0, j); fj1 = max(fj1, j); }
  fi0 = wave_min_i(fi0); fi1 = wave_max_i(fi1); fj0 = wave_min_i(fj0); fj1 = wave_max_i(fj1);
  if (fi0 <= fi1 && fj0 <= fj1 && (fi0 == 0 || fi1 == nx - 1 || fj0 == 0 || fj1 == ny - 1)) {
    if (lane == 0) { status[item] = kESrcOnBoundary; iters[item * iter_stride] = 0; }
    return;
  }
  // T0 fields in both skews, tau init
  const double sy0 = sin(y0), cy0 = cos(y0);
  for (int sk = 0; sk < 2; ++sk) {
    double *b0 = IS[sk];
    for (int cc = 0; cc < ns; ++cc) {
      const int i = 64 * cc + lane;
      for (int r = 0; r < R; ++r) {
        const int j = sk == 0 ? r - lane : r + lane - 63;
        double t0 = 0, t0x = 0, t0y = 0;
        if (i < nx && j >= 0 && j < ny) t0_cell_s(g.xx[i], g.yy[j], x0, sy0, cy0, fun0, t0, t0x, t0y);
        const size_t e = skidx(cc, r, lane, R);
        b0[e] = t0; b0[SE + e] = t0x; b0[2 * SE + e] = t0y;
      }
    }
  }
  for (size_t k = lane; k < N; k += 64) {
    const int i = (int)(k % nx), j = (int)(k / nx);
    tau[k] = (i >= fi0 && i <= fi1 && j >= fj0 && j <= fj1) ? 1.0 : 10.0;
  }
  wave_fence();

  int iter = 1;
  bool conv = false;
  for (; iter <= kFwdMaxIter; ++iter) {
    double acc_s = 0, acc_m = 0;
    for (int sw = 0; sw < 4; ++sw) {
      const int xd = sw < 2 ? 1 : -1;
      const int yd = (sw == 0 || sw == 2) ? 1 : -1;
      const int sk = xd == yd ? 0 : 1;
      const double *ps = PS[sk];
      const double *is = IS[sk];
      for (int cq = 0; cq < ns; ++cq) {
        const int c = xd > 0 ? cq : ns - 1 - cq;
        const int L = min(64, nx - 64 * c);
        const int i = 64 * c + lane;
        const bool lv = lane < L;
        const int lag = xd > 0 ? lane : L - 1 - lane;
        const int nsteps = ny + L - 1;
        // the uniform skew row of step t
        auto row_of = [&](int t) -> int {
          if (sk == 0) return yd > 0 ? t : ny + L - 2 - t;
          return yd > 0 ? t - L + 64 : ny + 62 - t;
        };
        auto load = [&](int t, FCoef &q, int &jj, bool &act) {
          const int u = t - lag;
          jj = yd > 0 ? u : ny - 1 - u;
          act = lv && u >= 0 && u < ny;
          const int r = row_of(t);
          if (r >= 0 && r < R) {
            const size_t e = skidx(c, r, lane, R);
            q.a = ps[e]; q.b = ps[SE + e]; q.c = ps[2 * SE + e]; q.f = ps[3 * SE + e];
            q.dx_ = ps[4 * SE + e]; q.dy_ = ps[5 * SE + e];
            q.t0 = is[e]; q.t0x = is[SE + e]; q.t0y = is[2 * SE + e];
          } else {
            q.a = q.b = q.c = q.f = q.dx_ = q.dy_ = q.t0 = q.t0x = q.t0y = 0;
          }
          q.tau = act ? tau[i + (size_t)nx * jj] : 0.0;
        };
        FCoef cur, nxt, nx2;
        int jc, jn, j2;
        bool ac, an, a2;
        load(0, cur, jc, ac);
        load(1, nxt, jn, an);
        double tprev = 0, zprev = 0;  // own result and T0 of the previous step
        // strip edges: the upstream neighbour column is final, the downstream
        // one still holds the previous sweep's values
        const bool up_edge = xd > 0 ? lane == 0 : lane == L - 1;
        const bool dn_edge = xd > 0 ? lane == L - 1 : lane == 0;
        const int iu = i - xd, id = i + xd;
        const int cu = c - xd, cdn = c + xd;   // neighbour strips
        const int lu = xd > 0 ? 63 : 0, ldn = xd > 0 ? 0 : 63;
        for (int t = 0; t < nsteps; ++t) {
          load(t + 2, nx2, j2, a2);
          const int src_up = lane - xd, src_dn = lane + xd;
          double tUx = shfl(tprev, src_up), zUx = shfl(zprev, src_up);
          double tDx = shfl(nxt.tau, src_dn), zDx = shfl(nxt.t0, src_dn);
          if (ac && up_edge && iu >= 0 && iu < nx) {
            tUx = tau[iu + (size_t)nx * jc];
            const int r = sk == 0 ? jc + lu : jc - lu + 63;
            zUx = is[skidx(cu, r, lu, R)];
          }
          if (ac && dn_edge && id >= 0 && id < nx) {
            tDx = tau[id + (size_t)nx * jc];
            const int r = sk == 0 ? jc + ldn : jc - ldn + 63;
            zDx = is[skidx(cdn, r, ldn, R)];
          }
          const double tW = xd > 0 ? tUx : tDx, tE = xd > 0 ? tDx : tUx;
          const double zW = xd > 0 ? zUx : zDx, zE = xd > 0 ? zDx : zUx;
          const double tS = yd > 0 ? tprev : nxt.tau, tN = yd > 0 ? nxt.tau : tprev;
          const double zS = yd > 0 ? zprev : nxt.t0, zN = yd > 0 ? nxt.t0 : zprev;
          double tnew = cur.tau;
          if (ac) {
            const bool frozen = i >= fi0 && i <= fi1 && jc >= fj0 && jc <= fj1;
            if (!frozen)
              tnew = fwd_cell_r(i, jc, nx, ny, dx, dy, cur.a, cur.b, cur.c, cur.f, cur.dx_,
                                cur.dy_, cur.t0, cur.t0x, cur.t0y, cur.tau, tW, tE, tS, tN, zW,
                                zE, zS, zN);
            const size_t k = i + (size_t)nx * jc;
            tau[k] = tnew;
            if (sw == 0) snap[k] = cur.tau;
            if (sw == 3) {
              const double d = fabs(tnew - snap[k]) * cur.t0;
              acc_s += d;
              acc_m = d > acc_m ? d : acc_m;
            }
          }
          tprev = tnew;
          zprev = cur.t0;
          cur = nxt; jc = jn; ac = an;
          nxt = nx2; jn = j2; an = a2;
        }
        wave_fence();
      }
    }
    // convergence (src/eikon_solver_2d.f90:1091-1115)
    const double s = wave_sum(acc_s), m = wave_max(acc_m);
    const double l1 = s / (double)nx / (double)ny;
    if (fabs(l1) < kFwdTol && m < kFwdTol) { conv = true; break; }
  }
  if (!conv) iter = kFwdMaxIter;
  // T = tau*T0 (src/eikon_solver_2d.f90:1124)
  for (size_t k = lane; k < N; k += 64) {
    const int i = (int)(k % nx), j = (int)(k / nx);
    const int l = i & 63;
    T[k] = tau[k] * IS[0][skidx(i >> 6, j + l, l, R)];
  }
  if (lane == 0) {
    status[item] = conv ? kOk : kWMaxIter;
    iters[item * iter_stride] = iter;
  }
}

// ------------------------------------------------------------- adjoint

struct ACoef {  // one cell's adjoint operands
  double a1, a2, b1, b2, d, dl, ta;
};

// Face coefficients of one interior cell (src/eikon_solver_2d.f90:1348-1370)
// and the sweep denominator d = (a2p-a1m)/dx + (b2p-b1m)/dy (:1392), stored
// as 0 where abs(d) < eps (:1394-1395).
__device__ __forceinline__ void adj_coef(const double *T, const double *a, const double *b,
                                         const double *c, int nx, size_t k, double dx,
                                         double dy, double *out) {
  const size_t W = k - 1, E = k + 1, S = k - nx, Nn = k + nx;
  const double a1 = -((T[k] - T[W]) / dx * (a[k] + a[W]) / 2) +
                    (T[Nn] - T[S] + T[W + nx] - T[W - nx]) / (4 * dy) * (c[k] + c[W]) / 2;
  const double a2 = -((T[E] - T[k]) / dx * (a[E] + a[k]) / 2) +
                    (T[E + nx] - T[E - nx] + T[Nn] - T[S]) / (4 * dy) * (c[E] + c[k]) / 2;
  const double b1 = -((T[k] - T[S]) / dy * (b[k] + b[S]) / 2) +
                    (T[E] - T[W] + T[E - nx] - T[W - nx]) / (4 * dx) * (c[k] + c[S]) / 2;
  const double b2 = -((T[Nn] - T[k]) / dy * (b[Nn] + b[k]) / 2) +
                    (T[E + nx] - T[W + nx] + T[E] - T[W]) / (4 * dx) * (c[Nn] + c[k]) / 2;
  const double a1m = (a1 - fabs(a1)) / 2, a2p = (a2 + fabs(a2)) / 2;
  const double b1m = (b1 - fabs(b1)) / 2, b2p = (b2 + fabs(b2)) / 2;
  const double d = (a2p - a1m) / dx + (b2p - b1m) / dy;
  out[0] = a1; out[1] = a2; out[2] = b1; out[3] = b2;
  out[4] = fabs(d) < kAdjEps ? 0.0 : d;
}

// Delta source of src/eikon_solver_2d.f90:1320-1336, receivers in order.
__device__ int adj_splat_s(const Grid &g, const double *xr_deg, const double *yr_deg,
                           const double *rhs, int nr, double *delta) {
  const int nx = g.nx, ny = g.ny;
  for (int r = 0; r < nr; ++r) {
    const double xr = deg2rad_s(xr_deg[r]), yr = deg2rad_s(yr_deg[r]);
    const int idx0 = (int)floor((xr - g.xx[0]) / g.dx + 1);
    const int idy0 = (int)floor((yr - g.yy[0]) / g.dy + 1);
    if (idx0 < 1 || idy0 < 1 || idx0 + 1 > nx || idy0 + 1 > ny) return kERecOutOfGrid;
    const int i0 = idx0 - 1, j0 = idy0 - 1;
    const float r1 = weight_r4_s((xr - g.xx[i0]) / g.dx);
    const float r2 = weight_r4_s((yr - g.yy[j0]) / g.dy);
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

__global__ __launch_bounds__(64) void k_adj_strip(Grid g, SkewGeom sg, PeriodFields pf, Items it,
                                                  StripAdjWork w, int *iters, int *status,
                                                  int iter_stride, int apply_mask) {
  const int slot = blockIdx.x;
  const int item = it.item0 + slot;
  const int lane = threadIdx.x;
  const int nx = g.nx, ny = g.ny, R = sg.R, ns = sg.ns;
  const size_t N = (size_t)nx * ny, SE = sg.skew_elems();
  if (status[item] < 0) return;  // an earlier stage failed
  const int p = it.pidx[item];
  const double *a = pf.a + p * N, *b = pf.b + p * N, *c = pf.c + p * N;
  const double *T = w.T + (size_t)slot * N;
  double *IS[2] = {w.iskew + (size_t)slot * 2 * kAF * SE,
                   w.iskew + (size_t)slot * 2 * kAF * SE + kAF * SE};
  double *Ta = w.Ta + (size_t)slot * N, *snap = w.snap + (size_t)slot * N;
  double *delta = w.delta + (size_t)slot * N;
  const double dx = g.dx, dy = g.dy;

  for (size_t k = lane; k < N; k += 64) delta[k] = 0;
  wave_fence();
  int err = kOk;
  if (lane == 0) {
    const int r0 = it.rec_off[item], nr = it.rec_off[item + 1] - r0;
    err = adj_splat_s(g, it.recx + r0, it.recy + r0, w.rhs + r0, nr, delta);
  }
  err = __shfl(err, 0, 64);
  if (err != kOk) {
    if (lane == 0) { status[item] = err; iters[item * iter_stride] = 0; }
    return;
  }
  wave_fence();
  for (int sk = 0; sk < 2; ++sk) {
    double *b0 = IS[sk];
    for (int cc = 0; cc < ns; ++cc) {
      const int i = 64 * cc + lane;
      for (int r = 0; r < R; ++r) {
        const int j = sk == 0 ? r - lane : r + lane - 63;
        double v[kAF] = {0, 0, 0, 0, 0, 0};
        if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) {
          const size_t k = i + (size_t)nx * j;
          adj_coef(T, a, b, c, nx, k, dx, dy, v);
          v[5] = delta[k];
        }
        const size_t e = skidx(cc, r, lane, R);
#pragma unroll
        for (int f = 0; f < kAF; ++f) b0[f * SE + e] = v[f];
      }
    }
  }
  for (size_t k = lane; k < N; k += 64) {
    const int i = (int)(k % nx), j = (int)(k / nx);
    Ta[k] = (i == 0 || j == 0 || i == nx - 1 || j == ny - 1) ? 0.0 : 100.0;
  }
  wave_fence();

  int iter = 1;
  bool conv = false;
  for (; iter <= kAdjMaxIter; ++iter) {
    double acc_s = 0, acc_m = 0;
    // (xdirec, ydirec) = (-1,-1), (-1,1), (1,-1), (1,1): x ascending first
    // (src/eikon_solver_2d.f90:1383-1390)
    for (int sw = 0; sw < 4; ++sw) {
      const int xd = sw < 2 ? 1 : -1;
      const int yd = (sw == 0 || sw == 2) ? 1 : -1;
      const int sk = xd == yd ? 0 : 1;
      const double *is = IS[sk];
      for (int cq = 0; cq < ns; ++cq) {
        const int cs = xd > 0 ? cq : ns - 1 - cq;
        const int L = min(64, nx - 64 * cs);
        const int i = 64 * cs + lane;
        const bool lv = lane < L;
        const int lag = xd > 0 ? lane : L - 1 - lane;
        const int nsteps = ny + L - 1;
        auto row_of = [&](int t) -> int {
          if (sk == 0) return yd > 0 ? t : ny + L - 2 - t;
          return yd > 0 ? t - L + 64 : ny + 62 - t;
        };
        auto load = [&](int t, ACoef &q, int &jj, bool &act) {
          const int u = t - lag;
          jj = yd > 0 ? u : ny - 1 - u;
          act = lv && u >= 0 && u < ny;
          const int r = row_of(t);
          if (r >= 0 && r < R) {
            const size_t e = skidx(cs, r, lane, R);
            q.a1 = is[e]; q.a2 = is[SE + e]; q.b1 = is[2 * SE + e]; q.b2 = is[3 * SE + e];
            q.d = is[4 * SE + e]; q.dl = is[5 * SE + e];
          } else {
            q.a1 = q.a2 = q.b1 = q.b2 = q.d = q.dl = 0;
          }
          q.ta = act ? Ta[i + (size_t)nx * jj] : 0.0;
        };
        ACoef cur, nxt, nx2;
        int jc, jn, j2;
        bool ac, an, a2;
        load(0, cur, jc, ac);
        load(1, nxt, jn, an);
        double tprev = 0;
        const bool up_edge = xd > 0 ? lane == 0 : lane == L - 1;
        const bool dn_edge = xd > 0 ? lane == L - 1 : lane == 0;
        const int iu = i - xd, id = i + xd;
        for (int t = 0; t < nsteps; ++t) {
          load(t + 2, nx2, j2, a2);
          double tUx = shfl(tprev, lane - xd);
          double tDx = shfl(nxt.ta, lane + xd);
          if (ac && up_edge && iu >= 0 && iu < nx) tUx = Ta[iu + (size_t)nx * jc];
          if (ac && dn_edge && id >= 0 && id < nx) tDx = Ta[id + (size_t)nx * jc];
          const double tW = xd > 0 ? tUx : tDx, tE = xd > 0 ? tDx : tUx;
          const double tS = yd > 0 ? tprev : nxt.ta, tN = yd > 0 ? nxt.ta : tprev;
          double tnew = cur.ta;
          if (ac) {
            const bool interior = i >= 1 && i <= nx - 2 && jc >= 1 && jc <= ny - 2;
            if (interior) {
              if (cur.d == 0.0) {
                tnew = 0;
              } else {
                const double a1p = (cur.a1 + fabs(cur.a1)) / 2, a2m = (cur.a2 - fabs(cur.a2)) / 2;
                const double b1p = (cur.b1 + fabs(cur.b1)) / 2, b2m = (cur.b2 - fabs(cur.b2)) / 2;
                const double e = (tW * a1p - tE * a2m) / dx + (tS * b1p - tN * b2m) / dy;
                tnew = (cur.dl + e) / cur.d;
              }
            }
            const size_t k = i + (size_t)nx * jc;
            Ta[k] = tnew;
            if (sw == 0) snap[k] = cur.ta;
            if (sw == 3) {
              const double dd = fabs(tnew - snap[k]);
              acc_s += dd;
              acc_m = dd > acc_m ? dd : acc_m;
            }
          }
          tprev = tnew;
          cur = nxt; jc = jn; ac = an;
          nxt = nx2; jn = j2; an = a2;
        }
        wave_fence();
      }
    }
    // convergence (src/eikon_solver_2d.f90:1412-1434)
    const double s = wave_sum(acc_s), m = wave_max(acc_m);
    const double l1 = s / (double)(nx * ny);
    if (fabs(l1) < kAdjTol && m < kAdjTol) { conv = true; break; }
  }
  if (!conv) iter = kAdjMaxIter;
  if (apply_mask) {
    // mask (src/eikon_solver_2d.f90:1449-1467), in degrees
    const double x0 = it.srcx[item], y0 = it.srcy[item];
    const double dxd = g.xd[1] - g.xd[0], dyd = g.yd[1] - g.yd[0];
    for (size_t k = lane; k < N; k += 64) {
      const int i = (int)(k % nx), j = (int)(k / nx);
      if (fabs(g.xd[i] - x0) < 1.5 * dxd && fabs(g.yd[j] - y0) < 1.5 * dyd) Ta[k] = 0;
    }
  }
  if (lane == 0) {
    if (!conv && status[item] == kOk) status[item] = kWMaxIter;
    iters[item * iter_stride] = iter;
  }
}

}  // namespace

void launch_adj_strip(hipStream_t s, const Grid &g, const SkewGeom &sg, const PeriodFields &pf,
                      const Items &it, const StripAdjWork &w, int *iters, int *status,
                      int iter_stride, int apply_mask) {
  hipLaunchKernelGGL(k_adj_strip, dim3(it.count), dim3(64), 0, s, g, sg, pf, it, w, iters, status,
                     iter_stride, apply_mask);
}

void launch_period_skew(hipStream_t s, const Grid &g, const SkewGeom &sg, int np,
                        const PeriodFields &pf, double *out) {
  const size_t SE = sg.skew_elems();
  hipLaunchKernelGGL(k_period_skew, dim3((unsigned)((SE + 255) / 256), 2, np), dim3(256), 0, s,
                     g, sg, np, pf.a, pf.b, pf.c, pf.fun, out);
}

void launch_fwd_strip(hipStream_t s, const Grid &g, const SkewGeom &sg, const double *pskew,
                      const PeriodFields &pf, const Items &it, const StripFwdWork &w, int *iters,
                      int *status, int iter_stride) {
  hipLaunchKernelGGL(k_fwd_strip, dim3(it.count), dim3(64), 0, s, g, sg, pskew, pf.fun, it, w,
                     iters, status, iter_stride);
}

}  // namespace satt
