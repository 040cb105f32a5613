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
// Layout.  Every per-cell field a sweep touches is stored skewed so that the
// 64 cells of one step are one contiguous 512-byte row:
//   plus skew  (sweeps with xd == yd):  row r = j + l
//   minus skew (sweeps with xd != yd):  row r = j - l + 63
// each as [strip][row 0..ny+62][lane 0..63], zero padded.  Read-only
// coefficients exist in both skews.  tau / Ta also exist in both: a sweep of
// the same family as the previous one updates its skew in place; the first
// sweep after a family change reads one skew and writes the other.  Natural
// layout (i + nx*j) is used only for the solve's inputs and outputs.
#include <climits>

#include "satt_internal.h"
#include "satt_strip.h"

namespace satt {
namespace {

__device__ __forceinline__ double deg2rad_s(double d) { return d / 180.0 * kPiF; }
__device__ __forceinline__ float weight_r4_s(double v) { return (float)(v < 1.0 ? v : 1.0); }

__device__ __forceinline__ double shfl(double v, int src) { return __shfl(v, src, 64); }

// Whole-wave lane shifts by one (DPP wave_shr:1 / wave_shl:1 of the GFX9
// family): lane l receives lane l-1 (SHR) or lane l+1 (SHL); the lane that
// has no source receives 0.  One VALU op per 32-bit half instead of an LDS
// permute, which keeps the step-to-step dependency off the LDS latency.
template <int FROM>  // FROM = -1: take lane l-1; +1: take lane l+1
__device__ __forceinline__ double lane_shift(double v) {
  constexpr int ctrl = FROM < 0 ? 0x138 : 0x130;  // wave_shr:1 : wave_shl:1
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), ctrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), ctrl, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

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


// Position of cell (i, j) in skew sk (0 plus, 1 minus).  For 64 lanes of one
// strip and a common offset (di, dj) the positions of (i+di, j+dj) are one
// contiguous run, except for the lane that crosses into the next strip.
__device__ __forceinline__ size_t skpos(int sk, int i, int j, int R) {
  const int c = i >> 6, l = i & 63;
  const int r = sk == 0 ? j + l : j - l + 63;
  return skidx(c, r, l, R);
}

// Skew row of step t (uniform across lanes), see the layout note above.
template <int XD, int YD>
__device__ __forceinline__ int skew_row(int t, int ny, int L) {
  constexpr int SK = XD == YD ? 0 : 1;
  if (SK == 0) return YD > 0 ? t : ny + L - 2 - t;
  return YD > 0 ? t - L + 64 : ny + 62 - t;
}

// The operands a lane needs for its cell of step t.  "Far" operands are the
// ones other lanes read one step early (tau, T0) plus the strip-edge and
// convergence values; they are fetched two steps ahead.  "Near" operands are
// the cell's own coefficients, fetched one step ahead.  Loads use clamped
// (always valid) addresses so that no branch separates a load from its use.
struct FFar {
  double tau, t0;    // old tau and T0 of the lane's cell
  double tup, zup;   // upstream neighbour column (edge lane): final tau, T0
  double tdn, zdn;   // downstream neighbour column (edge lane): old tau, T0
  double snap;       // tau at the start of the iteration (last sweep only)
  int j;
  bool act;
};
struct FNear {
  double a, b, c, f, dx_, dy_, t0x, t0y;
};

// Everything one forward solve's strip passes need.
struct FwdCtx {
  int nx, ny, R, lane;
  size_t SE;
  double dx, dy;
  int fi0, fi1, fj0, fj1;  // frozen rectangle
  const double *ps[2];     // period skews (plus, minus), kPF fields each
  const double *is[2];     // item T0 skews (plus, minus), kFF fields each
  double *snap;            // plus skew
};

// One sweep (XD, YD) over strip c, reading tau from `tin` (the sweep's own
// skew) and writing it to `tout`: in place, or (CROSS) into the other skew,
// which is where the next sweep family reads it.  MODE 0: plain; 1: record
// the snapshot (first sweep of an iteration); 2: accumulate |tau-snap|*T0
// (last sweep).
template <int XD, int YD, int MODE, bool CROSS>
__device__ void fwd_strip_pass(const FwdCtx &C, int c, const double *tin, double *tout,
                               double &acc_s, double &acc_m) {
  constexpr int SK = XD == YD ? 0 : 1;
  constexpr int SO = CROSS ? 1 - SK : SK;
  const int nx = C.nx, ny = C.ny, R = C.R, lane = C.lane;
  const size_t SE = C.SE;
  const double *ps = C.ps[SK], *is = C.is[SK];
  const int L = min(64, nx - 64 * c);
  const int i = 64 * c + lane;
  const bool lv = lane < L;
  const int lag = XD > 0 ? lane : L - 1 - lane;
  const int nsteps = ny + L - 1;
  // strip edges: the upstream neighbour column is final (its strip ran
  // before this one, writing `tout`), the downstream one still holds the
  // previous sweep's values (`tin`)
  const bool up_edge = XD > 0 ? lane == 0 : lane == L - 1;
  const bool dn_edge = XD > 0 ? lane == L - 1 : lane == 0;
  const int iu = i - XD, id = i + XD;
  const bool up_ok = lv && up_edge && iu >= 0 && iu < nx;
  const bool dn_ok = lv && dn_edge && id >= 0 && id < nx;

  auto load_far = [&](int t, FFar &q) {
    const int u = t - lag;
    const int j = YD > 0 ? u : ny - 1 - u;
    q.j = j;
    q.act = lv && u >= 0 && u < ny;
    const int jc = min(max(j, 0), ny - 1);
    const int r = min(max(skew_row<XD, YD>(t, ny, L), 0), R - 1);
    const size_t e = skidx(c, r, lane, R);
    q.t0 = is[e];
    q.tau = tin[e];
    if (up_ok) {
      q.tup = tout[skpos(SO, iu, jc, R)];
      q.zup = is[skpos(SK, iu, jc, R)];
    }
    if (dn_ok) {
      q.tdn = tin[skpos(SK, id, jc, R)];
      q.zdn = is[skpos(SK, id, jc, R)];
    }
    if (MODE == 2) q.snap = C.snap[e];
  };
  auto load_near = [&](int t, FNear &q) {
    const int r = min(max(skew_row<XD, YD>(t, ny, L), 0), R - 1);
    const size_t e = skidx(c, r, lane, R);
    q.a = ps[e]; q.b = ps[SE + e]; q.c = ps[2 * SE + e]; q.f = ps[3 * SE + e];
    q.dx_ = ps[4 * SE + e]; q.dy_ = ps[5 * SE + e];
    q.t0x = is[SE + e]; q.t0y = is[2 * SE + e];
  };

  FFar fc, fn, f2;
  FNear nc, nn;
  load_far(0, fc);
  load_far(1, fn);
  load_near(0, nc);
  double tprev = 0, zprev = 0;  // own result and T0 of the previous step
  for (int t = 0; t < nsteps; ++t) {
    load_far(t + 2, f2);
    load_near(t + 1, nn);
    double tUx = lane_shift<-XD>(tprev), zUx = lane_shift<-XD>(zprev);
    double tDx = lane_shift<XD>(fn.tau), zDx = lane_shift<XD>(fn.t0);
    if (up_ok) { tUx = fc.tup; zUx = fc.zup; }
    if (dn_ok) { tDx = fc.tdn; zDx = fc.zdn; }
    const double tW = XD > 0 ? tUx : tDx, tE = XD > 0 ? tDx : tUx;
    const double zW = XD > 0 ? zUx : zDx, zE = XD > 0 ? zDx : zUx;
    const double tS = YD > 0 ? tprev : fn.tau, tN = YD > 0 ? fn.tau : tprev;
    const double zS = YD > 0 ? zprev : fn.t0, zN = YD > 0 ? fn.t0 : zprev;
    const int jc = fc.j;
    double tnew = fc.tau;
    if (fc.act) {
      const bool frozen = i >= C.fi0 && i <= C.fi1 && jc >= C.fj0 && jc <= C.fj1;
      if (!frozen)
        tnew = fwd_cell_r(i, jc, nx, ny, C.dx, C.dy, nc.a, nc.b, nc.c, nc.f, nc.dx_, nc.dy_,
                          fc.t0, nc.t0x, nc.t0y, fc.tau, tW, tE, tS, tN, zW, zE, zS, zN);
      const size_t e = skidx(c, skew_row<XD, YD>(t, ny, L), lane, R);
      tout[CROSS ? skpos(SO, i, jc, R) : e] = tnew;
      if (MODE == 1) C.snap[e] = fc.tau;
      if (MODE == 2) {
        const double d = fabs(tnew - fc.snap) * fc.t0;
        acc_s += d;
        acc_m = d > acc_m ? d : acc_m;
      }
    }
    tprev = tnew;
    zprev = fc.t0;
    fc = fn;
    fn = f2;
    nc = nn;
  }
  wave_fence();
}

template <int XD, int YD, int MODE, bool CROSS>
__device__ void fwd_sweep(const FwdCtx &C, int ns, const double *tin, double *tout,
                          double &acc_s, double &acc_m) {
  for (int cq = 0; cq < ns; ++cq)
    fwd_strip_pass<XD, YD, MODE, CROSS>(C, XD > 0 ? cq : ns - 1 - cq, tin, tout, acc_s, acc_m);
}

__global__ __launch_bounds__(64) void k_fwd_strip(Grid g, SkewGeom sg, const double *pskew,
                                                  const double *fun_nat, Items it, StripWork w,
                                                  int *iters, int *status, int iter_stride) {
  const int slot = blockIdx.x;
  const int item = it.item0 + slot;
  const int lane = threadIdx.x;
  const int nx = g.nx, ny = g.ny, R = sg.R, ns = sg.ns;
  const size_t N = (size_t)nx * ny, SE = sg.skew_elems();
  const int p = it.pidx[item];
  double *sk0 = w.skew + (size_t)slot * kSlotSkews * SE;
  double *tauP = sk0 + kSkTauP * SE, *tauM = sk0 + kSkTauM * SE;
  double *TP = sk0 + kSkTP * SE, *TM = sk0 + kSkTM * SE;
  double *Tnat = w.T + (size_t)slot * N;
  FwdCtx C;
  C.nx = nx; C.ny = ny; C.R = R; C.lane = lane; C.SE = SE; C.dx = g.dx; C.dy = g.dy;
  C.ps[0] = pskew + ((size_t)p * 2 + 0) * kPF * SE;
  C.ps[1] = pskew + ((size_t)p * 2 + 1) * kPF * SE;
  C.is[0] = sk0;
  C.is[1] = sk0 + kFF * SE;
  C.snap = sk0 + kSkSnapF * SE;
  const double x0 = deg2rad_s(it.srcx[item]), y0 = deg2rad_s(it.srcy[item]);

  double fun0 = 0;
  const int st0 = fwd_fun0_s(g, fun_nat + (size_t)p * N, x0, y0, &fun0);
  if (st0 != kOk) {
    if (lane == 0) { status[item] = st0; iters[item * iter_stride] = 0; }
    return;
  }
  // frozen set (src/eikon_solver_2d.f90:861-872): a rectangle of columns and rows
  int fi0 = INT_MAX, fi1 = -1, fj0 = INT_MAX, fj1 = -1;
  for (int i = lane; i < nx; i += 64)
    if (fabs((g.xx[i] - x0) / g.dx) <= 1) { fi0 = min(fi0, i); fi1 = max(fi1, i); }
  for (int j = lane; j < ny; j += 64)
    if (fabs((g.yy[j] - y0) / g.dy) <= 1) { fj0 = min(fj0, j); fj1 = max(fj1, j); }
  fi0 = wave_min_i(fi0); fi1 = wave_max_i(fi1); fj0 = wave_min_i(fj0); fj1 = wave_max_i(fj1);
  if (fi0 <= fi1 && fj0 <= fj1 && (fi0 == 0 || fi1 == nx - 1 || fj0 == 0 || fj1 == ny - 1)) {
    if (lane == 0) { status[item] = kESrcOnBoundary; iters[item * iter_stride] = 0; }
    return;
  }
  C.fi0 = fi0; C.fi1 = fi1; C.fj0 = fj0; C.fj1 = fj1;
  // T0 fields in both skews; tau init in the plus skew
  const double sy0 = sin(y0), cy0 = cos(y0);
  for (int sk = 0; sk < 2; ++sk) {
    double *b0 = sk0 + (size_t)sk * kFF * SE;
    for (int cc = 0; cc < ns; ++cc) {
      const int i = 64 * cc + lane;
      for (int r = 0; r < R; ++r) {
        const int j = sk == 0 ? r - lane : r + lane - 63;
        const bool valid = i < nx && j >= 0 && j < ny;
        double t0 = 0, t0x = 0, t0y = 0;
        if (valid) t0_cell_s(g.xx[i], g.yy[j], x0, sy0, cy0, fun0, t0, t0x, t0y);
        const size_t e = skidx(cc, r, lane, R);
        b0[e] = t0; b0[SE + e] = t0x; b0[2 * SE + e] = t0y;
        if (sk == 0)
          tauP[e] = (valid && i >= fi0 && i <= fi1 && j >= fj0 && j <= fj1) ? 1.0 : 10.0;
      }
    }
  }
  wave_fence();

  int iter = 1;
  bool conv = false;
  for (; iter <= kFwdMaxIter; ++iter) {
    double acc_s = 0, acc_m = 0;
    // sweeps (x+,y+), (x+,y-), (x-,y+), (x-,y-)  (src/eikon_solver_2d.f90:884-898)
    fwd_sweep<1, 1, 1, true>(C, ns, tauP, tauM, acc_s, acc_m);
    fwd_sweep<1, -1, 0, false>(C, ns, tauM, tauM, acc_s, acc_m);
    fwd_sweep<-1, 1, 0, true>(C, ns, tauM, tauP, acc_s, acc_m);
    fwd_sweep<-1, -1, 2, false>(C, ns, tauP, tauP, acc_s, acc_m);
    // convergence (src/eikon_solver_2d.f90:1091-1115)
    const double s = wave_sum(acc_s), m = wave_max(acc_m);
    const double l1 = s / (double)nx / (double)ny;
    if (fabs(l1) < kFwdTol && m < kFwdTol) { conv = true; break; }
  }
  if (!conv) iter = kFwdMaxIter;
  // T = tau*T0 (src/eikon_solver_2d.f90:1124) in both skews and natural layout
  for (int cc = 0; cc < ns; ++cc) {
    const int i = 64 * cc + lane;
    for (int r = 0; r < R; ++r) {
      const size_t e = skidx(cc, r, lane, R);
      TP[e] = tauP[e] * C.is[0][e];
      const int j = r + lane - 63;
      double tm = 0;
      if (i < nx && j >= 0 && j < ny) tm = tauP[skpos(0, i, j, R)] * C.is[1][e];
      TM[e] = tm;
    }
  }
  wave_fence();
  for (size_t k = lane; k < N; k += 64) {
    const int i = (int)(k % nx), j = (int)(k / nx);
    Tnat[k] = TP[skpos(0, i, j, R)];
  }
  if (lane == 0) {
    status[item] = conv ? kOk : kWMaxIter;
    iters[item * iter_stride] = iter;
  }
}

// ------------------------------------------------------------- adjoint

// Face coefficients of one interior cell (src/eikon_solver_2d.f90:1348-1370)
// from T at the 3x3 neighbourhood and the metric at the faces, and the sweep
// denominator d = (a2p-a1m)/dx + (b2p-b1m)/dy (:1392), stored as 0 where
// abs(d) < eps (:1394-1395).
__device__ __forceinline__ void adj_coef_v(double Tk, double TW, double TE, double TS, double TN,
                                           double TWS, double TWN, double TES, double TEN,
                                           double ak, double aW, double aE, double bk, double bS,
                                           double bN, double ck, double cW, double cE, double cS,
                                           double cN, double dx, double dy, double *out) {
  const double a1 = -((Tk - TW) / dx * (ak + aW) / 2) + (TN - TS + TWN - TWS) / (4 * dy) * (ck + cW) / 2;
  const double a2 = -((TE - Tk) / dx * (aE + ak) / 2) + (TEN - TES + TN - TS) / (4 * dy) * (cE + ck) / 2;
  const double b1 = -((Tk - TS) / dy * (bk + bS) / 2) + (TE - TW + TES - TWS) / (4 * dx) * (ck + cS) / 2;
  const double b2 = -((TN - Tk) / dy * (bN + bk) / 2) + (TEN - TWN + TE - TW) / (4 * dx) * (cN + ck) / 2;
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

struct AFar {
  double ta, tup, tdn, snap;
  int j;
  bool act;
};
struct ANear {
  double a1, a2, b1, b2, d, dl;
};

struct AdjCtx {
  int nx, ny, R, lane;
  size_t SE;
  double dx, dy;
  const double *is[2];  // item coefficient skews, kAF fields each
  double *snap;         // plus skew
};

template <int XD, int YD, int MODE, bool CROSS>
__device__ void adj_strip_pass(const AdjCtx &C, int c, const double *tin, double *tout,
                               double &acc_s, double &acc_m) {
  constexpr int SK = XD == YD ? 0 : 1;
  constexpr int SO = CROSS ? 1 - SK : SK;
  const int nx = C.nx, ny = C.ny, R = C.R, lane = C.lane;
  const size_t SE = C.SE;
  const double *is = C.is[SK];
  const int L = min(64, nx - 64 * c);
  const int i = 64 * c + lane;
  const bool lv = lane < L;
  const int lag = XD > 0 ? lane : L - 1 - lane;
  const int nsteps = ny + L - 1;
  const bool up_edge = XD > 0 ? lane == 0 : lane == L - 1;
  const bool dn_edge = XD > 0 ? lane == L - 1 : lane == 0;
  const int iu = i - XD, id = i + XD;
  const bool up_ok = lv && up_edge && iu >= 0 && iu < nx;
  const bool dn_ok = lv && dn_edge && id >= 0 && id < nx;
  const bool interior_col = i >= 1 && i <= nx - 2;

  auto load_far = [&](int t, AFar &q) {
    const int u = t - lag;
    const int j = YD > 0 ? u : ny - 1 - u;
    q.j = j;
    q.act = lv && u >= 0 && u < ny;
    const int jc = min(max(j, 0), ny - 1);
    const int r = min(max(skew_row<XD, YD>(t, ny, L), 0), R - 1);
    const size_t e = skidx(c, r, lane, R);
    q.ta = tin[e];
    if (up_ok) q.tup = tout[skpos(SO, iu, jc, R)];
    if (dn_ok) q.tdn = tin[skpos(SK, id, jc, R)];
    if (MODE == 2) q.snap = C.snap[e];
  };
  auto load_near = [&](int t, ANear &q) {
    const int r = min(max(skew_row<XD, YD>(t, ny, L), 0), R - 1);
    const size_t e = skidx(c, r, lane, R);
    q.a1 = is[e]; q.a2 = is[SE + e]; q.b1 = is[2 * SE + e]; q.b2 = is[3 * SE + e];
    q.d = is[4 * SE + e]; q.dl = is[5 * SE + e];
  };
  AFar fc, fn, f2;
  ANear nc, nn;
  load_far(0, fc);
  load_far(1, fn);
  load_near(0, nc);
  double tprev = 0;
  for (int t = 0; t < nsteps; ++t) {
    load_far(t + 2, f2);
    load_near(t + 1, nn);
    double tUx = lane_shift<-XD>(tprev);
    double tDx = lane_shift<XD>(fn.ta);
    if (up_ok) tUx = fc.tup;
    if (dn_ok) tDx = fc.tdn;
    const double tW = XD > 0 ? tUx : tDx, tE = XD > 0 ? tDx : tUx;
    const double tS = YD > 0 ? tprev : fn.ta, tN = YD > 0 ? fn.ta : tprev;
    const int jc = fc.j;
    double tnew = fc.ta;
    if (fc.act) {
      if (interior_col && jc >= 1 && jc <= ny - 2) {
        // src/eikon_solver_2d.f90:1392-1402
        if (nc.d == 0.0) {
          tnew = 0;
        } else {
          const double a1p = (nc.a1 + fabs(nc.a1)) / 2, a2m = (nc.a2 - fabs(nc.a2)) / 2;
          const double b1p = (nc.b1 + fabs(nc.b1)) / 2, b2m = (nc.b2 - fabs(nc.b2)) / 2;
          const double e = (tW * a1p - tE * a2m) / C.dx + (tS * b1p - tN * b2m) / C.dy;
          tnew = (nc.dl + e) / nc.d;
        }
      }
      const size_t e = skidx(c, skew_row<XD, YD>(t, ny, L), lane, R);
      tout[CROSS ? skpos(SO, i, jc, R) : e] = tnew;
      if (MODE == 1) C.snap[e] = fc.ta;
      if (MODE == 2) {
        const double dd = fabs(tnew - fc.snap);
        acc_s += dd;
        acc_m = dd > acc_m ? dd : acc_m;
      }
    }
    tprev = tnew;
    fc = fn;
    fn = f2;
    nc = nn;
  }
  wave_fence();
}

template <int XD, int YD, int MODE, bool CROSS>
__device__ void adj_sweep(const AdjCtx &C, int ns, const double *tin, double *tout,
                          double &acc_s, double &acc_m) {
  for (int cq = 0; cq < ns; ++cq)
    adj_strip_pass<XD, YD, MODE, CROSS>(C, XD > 0 ? cq : ns - 1 - cq, tin, tout, acc_s, acc_m);
}

__global__ __launch_bounds__(64) void k_adj_strip(Grid g, SkewGeom sg, const double *pskew,
                                                  Items it, StripWork w, const double *rhs,
                                                  int out_nat, int *iters, int *status,
                                                  int iter_stride, int apply_mask) {
  const int slot = blockIdx.x;
  const int item = it.item0 + slot;
  const int lane = threadIdx.x;
  const int nx = g.nx, ny = g.ny, R = sg.R, ns = sg.ns;
  const size_t N = (size_t)nx * ny, SE = sg.skew_elems();
  if (status[item] < 0) return;  // an earlier stage failed
  const int p = it.pidx[item];
  double *sk0 = w.skew + (size_t)slot * kSlotSkews * SE;
  double *TaP = sk0 + kSkTaP * SE, *TaM = sk0 + kSkTaM * SE;
  const double *Tsk[2] = {sk0 + kSkTP * SE, sk0 + kSkTM * SE};
  double *delta = w.delta + (size_t)slot * N;
  double *Tout = (out_nat == kNatTd ? w.Td : w.Ta) + (size_t)slot * N;
  AdjCtx C;
  C.nx = nx; C.ny = ny; C.R = R; C.lane = lane; C.SE = SE; C.dx = g.dx; C.dy = g.dy;
  C.is[0] = sk0;
  C.is[1] = sk0 + kAF * SE;
  C.snap = sk0 + kSkSnapA * SE;
  const double dx = g.dx, dy = g.dy;

  for (size_t k = lane; k < N; k += 64) delta[k] = 0;
  wave_fence();
  int err = kOk;
  if (lane == 0) {
    const int r0 = it.rec_off[item], nr = it.rec_off[item + 1] - r0;
    err = adj_splat_s(g, it.recx + r0, it.recy + r0, rhs + r0, nr, delta);
  }
  err = __shfl(err, 0, 64);
  if (err != kOk) {
    if (lane == 0) { status[item] = err; iters[item * iter_stride] = 0; }
    return;
  }
  wave_fence();
  // coefficients in both skews; neighbours read with shifted loads from the
  // same skew (one contiguous run per (di, dj) offset)
  for (int sk = 0; sk < 2; ++sk) {
    double *b0 = sk0 + (size_t)sk * kAF * SE;
    const double *Ts = Tsk[sk];
    const double *pa = pskew + ((size_t)p * 2 + sk) * kPF * SE;
    const double *pb = pa + SE, *pc = pa + 2 * SE;
    for (int cc = 0; cc < ns; ++cc) {
      const int i = 64 * cc + lane;
      for (int r = 0; r < R; ++r) {
        const int j = sk == 0 ? r - lane : r + lane - 63;
        double v[kAF] = {0, 0, 0, 0, 0, 0};
        if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) {
          auto at = [&](int di, int dj) { return skpos(sk, i + di, j + dj, R); };
          const size_t k0 = at(0, 0), kW = at(-1, 0), kE = at(1, 0), kS = at(0, -1), kN = at(0, 1);
          adj_coef_v(Ts[k0], Ts[kW], Ts[kE], Ts[kS], Ts[kN], Ts[at(-1, -1)], Ts[at(-1, 1)],
                     Ts[at(1, -1)], Ts[at(1, 1)], pa[k0], pa[kW], pa[kE], pb[k0], pb[kS], pb[kN],
                     pc[k0], pc[kW], pc[kE], pc[kS], pc[kN], dx, dy, v);
          v[5] = delta[i + (size_t)nx * j];
        }
        const size_t e = skidx(cc, r, lane, R);
#pragma unroll
        for (int f = 0; f < kAF; ++f) b0[f * SE + e] = v[f];
        if (sk == 0) {
          const bool in_grid = i < nx && j >= 0 && j < ny;
          TaP[e] = (in_grid && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) ? 100.0 : 0.0;
        }
      }
    }
  }
  wave_fence();

  int iter = 1;
  bool conv = false;
  for (; iter <= kAdjMaxIter; ++iter) {
    double acc_s = 0, acc_m = 0;
    // (xdirec, ydirec) = (-1,-1), (-1,1), (1,-1), (1,1): x ascending first
    // (src/eikon_solver_2d.f90:1383-1390)
    adj_sweep<1, 1, 1, true>(C, ns, TaP, TaM, acc_s, acc_m);
    adj_sweep<1, -1, 0, false>(C, ns, TaM, TaM, acc_s, acc_m);
    adj_sweep<-1, 1, 0, true>(C, ns, TaM, TaP, acc_s, acc_m);
    adj_sweep<-1, -1, 2, false>(C, ns, TaP, TaP, acc_s, acc_m);
    // convergence (src/eikon_solver_2d.f90:1412-1434)
    const double s = wave_sum(acc_s), m = wave_max(acc_m);
    const double l1 = s / (double)(nx * ny);
    if (fabs(l1) < kAdjTol && m < kAdjTol) { conv = true; break; }
  }
  if (!conv) iter = kAdjMaxIter;
  // natural-layout output with the mask (src/eikon_solver_2d.f90:1449-1467, degrees)
  const double x0 = it.srcx[item], y0 = it.srcy[item];
  const double dxd = g.xd[1] - g.xd[0], dyd = g.yd[1] - g.yd[0];
  for (size_t k = lane; k < N; k += 64) {
    const int i = (int)(k % nx), j = (int)(k / nx);
    double v = TaP[skpos(0, i, j, R)];
    if (apply_mask && fabs(g.xd[i] - x0) < 1.5 * dxd && fabs(g.yd[j] - y0) < 1.5 * dyd) v = 0;
    Tout[k] = v;
  }
  if (lane == 0) {
    if (!conv && status[item] == kOk) status[item] = kWMaxIter;
    iters[item * iter_stride] = iter;
  }
}

}  // namespace

void launch_adj_strip(hipStream_t s, const Grid &g, const SkewGeom &sg, const double *pskew,
                      const Items &it, const StripWork &w, const double *rhs, int out_nat,
                      int *iters, int *status, int iter_stride, int apply_mask) {
  hipLaunchKernelGGL(k_adj_strip, dim3(it.count), dim3(64), 0, s, g, sg, pskew, it, w, rhs, out_nat,
                     iters, status, iter_stride, apply_mask);
}

void launch_period_skew(hipStream_t s, const Grid &g, const SkewGeom &sg, int np,
                        const PeriodFields &pf, double *out) {
  const size_t SE = sg.skew_elems();
  hipLaunchKernelGGL(k_period_skew, dim3((unsigned)((SE + 255) / 256), 2, np), dim3(256), 0, s,
                     g, sg, np, pf.a, pf.b, pf.c, pf.fun, out);
}

void launch_fwd_strip(hipStream_t s, const Grid &g, const SkewGeom &sg, const double *pskew,
                      const PeriodFields &pf, const Items &it, const StripWork &w, int *iters,
                      int *status, int iter_stride) {
  hipLaunchKernelGGL(k_fwd_strip, dim3(it.count), dim3(64), 0, s, g, sg, pskew, pf.fun, it, w,
                     iters, status, iter_stride);
}

}  // namespace satt
