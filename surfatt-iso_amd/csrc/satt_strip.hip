// Batched sweeps for gfx950: one wavefront per (source, period) item.
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
// 64 cells of one step are one contiguous row (512 bytes in FP64):
//   plus skew  (sweeps with xd == yd):  row r = j + l
//   minus skew (sweeps with xd != yd):  row r = j - l + 63
// each as [strip][row 0..ny+62][lane 0..63], zero padded.  Read-only
// coefficients exist in both skews.  tau / Ta also exist in both: every sweep
// updates its own family's skew in place, and at a family change the field
// is copied into the other skew row by row (coalesced stores, gathered
// loads that stay in cache).  Natural layout (i + nx*j) is used only for
// the solve's inputs and outputs.
//
// Teams.  A strip kernel runs one item per workgroup of W waves: the k-th
// strip of a sweep goes to wave k mod W, which follows the previous strip's
// wave at least a strip width behind, taking that strip's edge column from
// LDS (see Team).  W = 1 is one wave doing every strip in order.
//
// Precision.  The FP64 kernels (k_fwd_strip, k_adj_strip) run the
// reference's arithmetic.  The mixed mode runs an FP32 phase
// (k_fwd_mixed32, k_adj_mixed32) and then an FP64 polish (k_fwd_polish,
// k_adj_polish) in separate launches, so that the FP32 phase runs at the
// occupancy of its own (smaller) register footprint.
//   Forward FP32 phase: the reference's update with float operands, each
//   sweep evaluating only the candidates whose neighbours are upwind in its
//   direction, written in the increment u = tau - tau_nb: with tt = tX + u, Tx = gx + px*u and
//   Ty = gy + py*u where gx = T0x*tX, gy = T0y*tX + sy*(T0/dy)*(tX - tY), so
//   the quadratic a*Tx^2 + b*Ty^2 - 2c*Tx*Ty = fun^2 has coefficients of the
//   size of the eikonal residual instead of (T0/dx)^2 (whose cancellation
//   would leave about two significant digits in float).  Roots by
//   q = -(eb + sign(eb)*sqrt(disc))/2, u = q/ea and ec/q.  The field holds
//   dtau = tau - 1, which keeps 3-4 more bits than tau near 1.  Causality
//   tests, acceptance, sweep order and the stop rule (residual accumulated in
//   double) are the reference's.  Polish: the same upwind update in FP64 on
//   dtau widened to double, iterated to the reference's stop rule (at least
//   once); its fixed point is that of the reference's full update, and
//   T = (1 + dtau)*T0.
//   Adjoint FP32 phase: Ta = k0 + kW*TaW + kE*TaE + kS*TaS + kN*TaN with
//   k0 = delta/d, kW = a1p/dx/d, kE = -a2m/dx/d, kS = b1p/dy/d, kN = -b2m/dy/d
//   formed in FP64 and rounded to float; stop when the reference's rule
//   holds or the largest change is at most 2^-21 of the largest |Ta|.
//   The FP32 updates (upwind cell, adjoint) are written with explicit fused
//   multiply-adds, which the oracle restates with fmaf; the rest of the
//   code, the FP64 polish included, is compiled without contraction.
//   Polish: one FP64 iteration of the reference's update.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "satt_internal.h"
#include "satt_strip.h"

namespace satt {
namespace {

// Address-space-qualified pointers for the hot loops: fields in device
// memory (global loads and stores rather than flat ones, which also count
// against the LDS/scalar wait counter) and the teams' LDS channels.
#define GMEM __attribute__((address_space(1)))
#define LMEM __attribute__((address_space(3)))
template <typename T>
__device__ __forceinline__ GMEM T *gm(T *p) { return (GMEM T *)p; }
template <typename T>
__device__ __forceinline__ const GMEM T *gm(const T *p) { return (const GMEM T *)p; }

constexpr float kMixRelStop = 4.76837158203125e-07f;  // 2^-21
// Stop of the mixed forward's FP32 phase, in units of the reference's
// tolerance: measured at 401^2 x 4000 items, 1x costs 4.9 FP32 iterations
// and 4x 4.2, with the FP64 polish still one iteration for all but 0.4% of
// the items; at 8x and 16x the polish's extra iterations cost more time than
// the FP32 iterations saved.
constexpr double kMixFwdTolScale = 4.0;

__device__ __forceinline__ double deg2rad_s(double d) { return d / 180.0 * kPiF; }
__device__ __forceinline__ float weight_r4_s(double v) { return (float)(v < 1.0 ? v : 1.0); }

// Whole-wave lane shifts by one (DPP wave_shr:1 / wave_shl:1 of the GFX9
// family): lane l receives lane l-1 (SHR) or lane l+1 (SHL); the lane that
// has no source receives 0.  One VALU op per 32-bit half instead of an LDS
// permute, which keeps the step-to-step dependency off the LDS latency.
template <int FROM>  // FROM = -1: take lane l-1; +1: take lane l+1
__device__ __forceinline__ int dpp_shift(int v) {
  constexpr int ctrl = FROM < 0 ? 0x138 : 0x130;  // wave_shr:1 : wave_shl:1
  return __builtin_amdgcn_update_dpp(0, v, ctrl, 0xf, 0xf, false);
}
template <int FROM>
__device__ __forceinline__ double lane_shift(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_shift<FROM>((int)(b & 0xffffffffLL));
  const int hi = dpp_shift<FROM>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int FROM>
__device__ __forceinline__ float lane_shift(float v) {
  return __int_as_float(dpp_shift<FROM>(__float_as_int(v)));
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

// The index unchanged, hidden from the optimizer: a load at a selected
// index whose other choice is an index loaded anyway would otherwise become
// that load plus a second one behind a branch.
__device__ __forceinline__ size_t opaque_index(size_t x) {
  asm("" : "+v"(x));
  return x;
}
// A copy of x in a register of its own (early-clobber output): a value kept
// from one step to the next then never shares a register with an operand
// set that the next step refills.
__device__ __forceinline__ double own_reg(double x) {
  double y;
  asm("v_mov_b64 %0, %1" : "=&v"(y) : "v"(x));
  return y;
}

// One cell of the forward sweep body (src/eikon_solver_2d.f90:910-1079) with
// every operand in registers.  Every floating-point expression is the
// reference's, in its order; what the four quadrants have in common is
// evaluated once: px, qx and the a-terms depend only on the x neighbour (W or
// E), py, qy and the b-terms only on the y neighbour (S or N), so e.g.
// ak*(px*px) + bk*(py*py) - 2*ck*px*py is (A1x + B1y) - C1x*py with the same
// operands and roundings.  Quadrants and one-sided candidates at the grid
// boundary, which the reference skips, are evaluated and discarded.
// dis_x = sqrt(fun**2*(b/(a*b-c**2))) and dis_y = sqrt(fun**2*(a/(a*b-c**2)))
// are per-cell constants precomputed by the same expression.
__device__ __forceinline__ double fwd_cell_r(int i, int j, int nx, int ny, double dx, double dy,
                                             double ak, double bk, double ck, double fk,
                                             double disx, double disy, double t0, double t0x,
                                             double t0y, double tc, double tW, double tE,
                                             double tS, double tN, double zW, double zE,
                                             double zS, double zN) {
  const double t0dx = t0 / dx, t0dy = t0 / dy;
  // x neighbour: 0 = W (i-1), 1 = E (i+1); y neighbour: 0 = S (j-1), 1 = N (j+1)
  const double px[2] = {t0x + t0dx, t0x - t0dx};
  const double qx[2] = {-(t0dx * tW), t0dx * tE};
  const double py[2] = {t0y + t0dy, t0y - t0dy};
  const double qy[2] = {-(t0dy * tS), t0dy * tN};
  const bool hx[2] = {i > 0, i < nx - 1};
  const bool hy[2] = {j > 0, j < ny - 1};
  const double ak2 = 2 * ak, bk2 = 2 * bk, ck2 = 2 * ck, ff = fk * fk;
  double A1[2], A2[2], A3[2], C1[2], C2[2], B1[2], B2[2], B3[2];
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    A1[x] = ak * (px[x] * px[x]);
    A2[x] = ak2 * px[x] * qx[x];
    A3[x] = ak * (qx[x] * qx[x]);
    C1[x] = ck2 * px[x];
    C2[x] = ck2 * qx[x];
  }
#pragma unroll
  for (int y = 0; y < 2; ++y) {
    B1[y] = bk * (py[y] * py[y]);
    B2[y] = bk2 * py[y] * qy[y];
    B3[y] = bk * (qy[y] * qy[y]);
  }
  double tnew = tc;
  // quadrants in the reference's order: (W,S), (W,N), (E,S), (E,N); the
  // causality test requires chx >= 0 with W, <= 0 with E, chy >= 0 with S,
  // <= 0 with N
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int x = q >> 1, y = q & 1;
    const double ea = A1[x] + B1[y] - C1[x] * py[y];
    const double eb = A2[x] + B2[y] - ck2 * (px[x] * qy[y] + py[y] * qx[x]);
    const double ec = A3[x] + B3[y] - C2[x] * qy[y] - ff;
    const double delta = eb * eb - 4 * ea * ec;
    const bool use = hx[x] && hy[y] && delta >= 0;
    const double sq = sqrt(delta);
#pragma unroll
    for (int isol = 1; isol <= 2; ++isol) {
      const double tt = isol == 1 ? (-eb + sq) / (2 * ea) : (-eb - sq) / (2 * ea);
      const double Tx = px[x] * tt + qx[x];
      const double Ty = py[y] * tt + qy[y];
      const double chx = ak * Tx - ck * Ty;
      const double chy = bk * Ty - ck * Tx;
      const bool ok = (x == 0 ? chx >= 0 : chx <= 0) && (y == 0 ? chy >= 0 : chy <= 0);
      if (use && ok && tt < tnew) tnew = tt;
    }
  }
  // one-sided candidates (:996-1075): W, E, S, N
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const bool isx = n < 2;
    const int s = n & 1;
    const double p = isx ? px[s] : py[s], q = isx ? qx[s] : qy[s];
    const double dis = isx ? disx : disy;
    const double tn = n == 0 ? tW : n == 1 ? tE : n == 2 ? tS : tN;
    const double t0n = n == 0 ? zW : n == 1 ? zE : n == 2 ? zS : zN;
    const bool use = isx ? hx[s] : hy[s];
    double tt = (dis - q) / p;
    if (use && tt * t0 >= tn * t0n && tt > tn * 0.5 && tt < tnew) tnew = tt;
    tt = (-dis - q) / p;
    if (use && tt * t0 >= tn * t0n && tt > tn * 0.5 && tt < tnew) tnew = tt;
  }
  return tnew;
}

// a*b + c: fused in float (the FP32 phase, where it saves issue slots), two
// roundings in double (the FP64 polish, where the fused chains ran slower).
__device__ __forceinline__ float fmav(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fmav(double a, double b, double c) { return a * b + c; }

// One cell of the mixed mode's upwind update (FP32 phase, V = float; FP64
// polish, V = double) in the sweep direction (XD, YD) (see the precision note
// at the top).  Only the candidates whose neighbours are upwind in this sweep
// are evaluated: the quadrant of the neighbours (i-XD, j) and (i, j-YD), and
// the one-sided candidates from each.  Over the four sweeps of an iteration
// every candidate of the reference's update is evaluated once with its
// neighbours' newest values, and a fixed point is a fixed point of the full
// update.  Operands are offsets dtau = tau - 1; zX, zY are the upwind
// neighbours' T0.
template <typename V, int XD, int YD>
__device__ __forceinline__ V fwd_cell_up_r(int i, int j, int nx, int ny, V rdx, V rdy, V ak, V bk,
                                           V ck, V fk, V disx, V disy, V t0, V t0x, V t0y, V dc,
                                           V dX, V dY, V zX, V zY) {
  const bool hasX = XD > 0 ? i > 0 : i < nx - 1;
  const bool hasY = YD > 0 ? j > 0 : j < ny - 1;
  // px = T0x +- T0/dx, py = T0y +- T0/dy (upwind neighbour side)
  const V px = fmav(t0, XD > 0 ? rdx : -rdx, t0x);
  const V py = fmav(t0, YD > 0 ? rdy : -rdy, t0y);
  V best = dc;
  if (hasX && hasY) {
    // gx = T0x*tX, gy = T0y*tX +- (T0/dy)*(tX - tY) in offsets
    const V t0dy = t0 * rdy;
    const V gx = fmav(t0x, dX, t0x);
    const V gy = fmav(YD > 0 ? t0dy : -t0dy, dX - dY, fmav(t0y, dX, t0y));
    // ea u^2 + eb u + ec = 0: ea = a px^2 + b py^2 - 2c px py,
    // eb = 2(gx (a px - c py) + gy (b py - c px)),
    // ec = gx (a gx - 2c gy) + b gy^2 - f^2
    const V apx = ak * px, bpy = bk * py, cpx = ck * px, cpy = ck * py;
    const V ea = fmav(apx, px, py * fmav(V(-2), cpx, bpy));
    const V eb = V(2) * fmav(gx, apx - cpy, gy * (bpy - cpx));
    const V ec = fmav(gx, fmav(ak, gx, V(-2) * (ck * gy)), fmav(bk * gy, gy, -(fk * fk)));
    const V disc = fmav(eb, eb, V(-4) * ea * ec);
    if (disc >= 0) {
      const V sq = sqrt(disc);
      const V qq = V(-0.5) * (eb >= 0 ? eb + sq : eb - sq);
#pragma unroll
      for (int isol = 1; isol <= 2; ++isol) {
        const V u = isol == 1 ? qq / ea : ec / qq;
        const V dt = dX + u;
        const V Tx = fmav(px, u, gx), Ty = fmav(py, u, gy);
        const V chx = fmav(ak, Tx, -(ck * Ty)), chy = fmav(bk, Ty, -(ck * Tx));
        const bool ok = (XD > 0 ? chx >= 0 : chx <= 0) && (YD > 0 ? chy >= 0 : chy <= 0);
        if (ok && dt < best) best = dt;
      }
    }
  }
  // one-sided candidates (:996-1075): tt*T0 >= tn*T0n and tt > tn*0.5 in
  // offsets; p is the quadrant's px (x neighbour) or py (y neighbour)
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    if (side == 0 ? !hasX : !hasY) continue;
    const V p = side == 0 ? px : py, dis = side == 0 ? disx : disy;
    const V t0d = side == 0 ? t0x : t0y;
    const V dn = side == 0 ? dX : dY, zn = side == 0 ? zX : zY;
    const V rp = V(1) / p, base = fmav(t0d, dn, t0d);
    const V rhs = fmav(dn, zn, zn), half = fmav(V(0.5), dn, V(-0.5));
    V dt = fmav(dis - base, rp, dn);
    if (fmav(dt, t0, t0) >= rhs && dt > half && dt < best) best = dt;
    dt = fmav(-dis - base, rp, dn);
    if (fmav(dt, t0, t0) >= rhs && dt > half && dt < best) best = dt;
  }
  return best;
}

// ---------------------------------------------------------- period skews

// Per-period skewed fields: a, b, c, fun, dis_x, dis_y (kPF fields), both
// skews, in FP64 and (out32 != null) rounded to float.  Thread = (strip,
// row, lane) of one skew of one period.
__global__ void k_period_skew(Grid g, SkewGeom sg, int np, const double *a, const double *b,
                              const double *c, const double *fun, double *out, float *out32) {
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
  if (out32) {  // two sets of three, element by element (CellSet<3>): a, b, c; fun, dis_x, dis_y
    float *b32 = out32 + ((size_t)p * 2 + sk) * kPF * per_skew;
#pragma unroll
    for (int f = 0; f < kPF; ++f) b32[(f / 3) * 3 * per_skew + e * 3 + f % 3] = (float)v[f];
  }
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

// Source slowness fun0 (src/eikon_solver_2d.f90:825-841) and the frozen
// set (:859-872), a rectangle of columns rect[0..1] and rows rect[2..3].
// Returns kOk or the item's error status.
__device__ int fwd_source_rect(const Grid &g, const double *fun_p, double x0, double y0,
                               int lane, double &fun0, int rect[4]) {
  fun0 = 0;
  const int st0 = fwd_fun0_s(g, fun_p, x0, y0, &fun0);
  if (st0 != kOk) return st0;
  int fi0 = INT_MAX, fi1 = -1, fj0 = INT_MAX, fj1 = -1;
  for (int i = lane; i < g.nx; i += 64)
    if (fabs((g.xx[i] - x0) / g.dx) <= 1) { fi0 = min(fi0, i); fi1 = max(fi1, i); }
  for (int j = lane; j < g.ny; j += 64)
    if (fabs((g.yy[j] - y0) / g.dy) <= 1) { fj0 = min(fj0, j); fj1 = max(fj1, j); }
  fi0 = wave_min_i(fi0); fi1 = wave_max_i(fi1); fj0 = wave_min_i(fj0); fj1 = wave_max_i(fj1);
  if (fi0 <= fi1 && fj0 <= fj1 &&
      (fi0 == 0 || fi1 == g.nx - 1 || fj0 == 0 || fj1 == g.ny - 1))
    return kESrcOnBoundary;
  rect[0] = fi0; rect[1] = fi1; rect[2] = fj0; rect[3] = fj1;
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

// The waves of one workgroup working on one item (a team of W waves).  The
// k-th strip of a sweep goes to wave k mod W; the edge column a strip hands
// to the next strip (its final values, and their T0 in the forward) passes
// through LDS, with the count of rows published so far, so that the next
// wave runs its strip as soon as the rows it needs are out (at least L steps
// behind).  Each wave publishes only into its own channel, and a channel row
// is overwritten (by the wave's next strip) only after the consumer has read
// it, because that next strip runs behind the consumer's strip.  W = 1: one
// wave does every strip in order and the channel needs no counts: the
// previous strip has finished, and the strip's own edge lane overwrites a
// row after the strip's first lane has read it.  Either way the edge column
// comes from LDS instead of one scattered global load per step.
// A wait is bounded: past the bound it sets the team's stop flag, every
// later wait of the team returns at once, so the kernel always ends, and the
// kernel gives the item kETeamWait at its end.
struct Team {
  int W, w;           // waves, this wave (wave-uniform)
  int ny, ns;
  int sweep;          // sweeps started (the count set in use is sweep & 1)
  LMEM char *chan;    // [W][ny] edge values of the current strips (V)
  LMEM char *chan0;   // [W][ny] their T0 (forward; V)
  volatile LMEM int *cnt;   // [2][W] rows published by each wave in the sweep
  volatile LMEM int *stop;  // set when a wait ran out
  LMEM double *part;        // [ns][3] per-strip residual terms
  __device__ bool multi() const { return W > 1; }
};

constexpr int kTeamSpinLimit = 1 << 22;  // ~0.5 s of s_sleep per wait

// LDS bytes of a team of W waves on grids up to ny rows and ns strips.
__host__ __device__ inline size_t team_lds_bytes(int W, int ny, int ns, int vbytes) {
  if (W <= 1) return (size_t)2 * ny * vbytes;  // the two channels only
  return (size_t)ns * 3 * 8 + (size_t)2 * W * ny * vbytes + (size_t)(2 * W + 2) * 4;
}

// TEAM false: the one-wave form, with W = 1 known at compile time so that
// no team code is left in the kernel.
template <bool TEAM>
__device__ inline Team team_init(void *lds, int ny, int ns, int vbytes) {
  Team T;
  T.W = TEAM ? (int)(blockDim.x >> 6) : 1;
  T.w = TEAM ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
  T.ny = ny;
  T.ns = ns;
  T.sweep = 0;
  LMEM char *base = (LMEM char *)lds;
  T.part = (LMEM double *)base;
  LMEM char *b = base + (TEAM ? (size_t)ns * 3 * 8 : 0);
  T.chan = b;
  T.chan0 = b + (size_t)T.W * ny * vbytes;
  T.cnt = (volatile LMEM int *)(b + (size_t)2 * T.W * ny * vbytes);
  T.stop = T.cnt + 2 * T.W;
  if (TEAM) {
    if (threadIdx.x < 2 * (unsigned)T.W) T.cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) *T.stop = 0;
    __syncthreads();
  }
  return T;
}

// Start of a sweep: clear this wave's count of the other set (used by the
// previous sweep, which every wave has finished, and by the next one, which
// starts after the barrier that ends this one).
__device__ inline void team_sweep_begin(Team &T) {
  if (T.multi() && (threadIdx.x & 63) == 0) T.cnt[((T.sweep + 1) & 1) * T.W + T.w] = 0;
}
__device__ inline void team_sweep_end(Team &T) {
  ++T.sweep;
  __syncthreads();
}

// Whether a wait of the team ran out (never for W = 1, which has no LDS).
__device__ inline bool team_failed(const Team &T) { return T.multi() && *T.stop; }

// Wait (one lane) until wave pw has published `need` rows in this sweep.
__device__ inline void team_wait(const Team &T, int pw, int need) {
  volatile LMEM int *c = T.cnt + (T.sweep & 1) * T.W + pw;
  int spins = 0;
  while (*c < need && !*T.stop) {
    __builtin_amdgcn_s_sleep(2);
    if (++spins > kTeamSpinLimit) *T.stop = 1;
  }
}

template <typename V>
__device__ inline V team_read(const Team &T, int pw, int u) {
  return ((volatile LMEM V *)T.chan)[pw * T.ny + u];
}
template <typename V>
__device__ inline V team_read0(const Team &T, int pw, int u) {
  return ((volatile LMEM V *)T.chan0)[pw * T.ny + u];
}

// Publish (one lane) the edge value of row u (and, forward, its T0), the
// `count`-th row of this wave's strips in the sweep: the values first, then
// the count (LDS operations of a wave complete in order).
template <typename V>
__device__ inline void team_publish(const Team &T, int u, V v, int count) {
  ((volatile LMEM V *)T.chan)[T.w * T.ny + u] = v;
  if (T.multi()) T.cnt[(T.sweep & 1) * T.W + T.w] = count;
}
template <typename V>
__device__ inline void team_publish2(const Team &T, int u, V v, V z, int count) {
  ((volatile LMEM V *)T.chan)[T.w * T.ny + u] = v;
  ((volatile LMEM V *)T.chan0)[T.w * T.ny + u] = z;
  if (T.multi()) T.cnt[(T.sweep & 1) * T.W + T.w] = count;
}

// Reduction of a sweep's per-strip residual terms (sum, max, max) in the
// sweep's strip order: the same bits for any W.
struct StripAcc {
  double s = 0, m = 0, a = 0;
};
__device__ inline void team_strip_term(const Team &T, int k, double s, double m, double a,
                                       StripAcc &acc) {
  if (T.multi()) {
    if ((threadIdx.x & 63) == 0) {
      T.part[3 * k] = s;
      T.part[3 * k + 1] = m;
      T.part[3 * k + 2] = a;
    }
  } else {
    acc.s = acc.s + s;
    acc.m = m > acc.m ? m : acc.m;
    acc.a = a > acc.a ? a : acc.a;
  }
}
// After the barrier that ends the sweep.
__device__ inline StripAcc team_strip_total(const Team &T, const StripAcc &acc) {
  if (!T.multi()) return acc;
  StripAcc r;
  for (int k = 0; k < T.ns; ++k) {
    r.s = r.s + T.part[3 * k];
    r.m = T.part[3 * k + 1] > r.m ? T.part[3 * k + 1] : r.m;
    r.a = T.part[3 * k + 2] > r.a ? T.part[3 * k + 2] : r.a;
  }
  return r;
}

// A buffer resource over n bytes from p (raw, no stride; the gfx9 format
// word): a buffer load at an offset outside [0, n) returns 0 and a buffer
// store there is dropped, so out-of-range offsets need no branch.
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_over(const GMEM T *p, unsigned n) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)n, 0x00020000);
}
constexpr int kOutOfRange = INT_MIN;  // a buffer offset past any resource
constexpr int kNT = 2;  // a buffer access's cache-policy operand: the non-temporal hint
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// A set of n skewed fields of SE elements each (field k at base + k*SE), as
// the strip passes read and write them at element index e.  FP64: global
// accesses at 64-bit addresses.  FP32: one buffer resource over the whole
// set, the element as a 32-bit byte offset and the field as a scalar offset,
// so that the loads of a step share one offset register and need no address
// arithmetic of their own (plans keep a set within 4 GiB, kMaxSkewElems32).
template <typename V>
struct FieldSet;
template <>
struct FieldSet<double> {
  GMEM double *p;
  size_t SE;
  __device__ FieldSet(const GMEM double *base, size_t se, int) : p((GMEM double *)base), SE(se) {}
  __device__ double ld(int k, size_t e) const { return p[k * SE + e]; }
  __device__ void st(int k, size_t e, double v) const { p[k * SE + e] = v; }
  // the same with the non-temporal hint (streamed data the L2 need not keep)
  __device__ double ldn(int k, size_t e) const { return __builtin_nontemporal_load(&p[k * SE + e]); }
  __device__ void stn(int k, size_t e, double v) const { __builtin_nontemporal_store(v, &p[k * SE + e]); }
};
template <>
struct FieldSet<float> {
  __amdgpu_buffer_rsrc_t r;
  unsigned fb;  // bytes of one field
  __device__ FieldSet(const GMEM float *base, size_t se, int n)
      : r(buffer_over(base, (unsigned)(n * se * sizeof(float)))), fb((unsigned)(se * sizeof(float))) {}
  __device__ float ld(int k, size_t e) const {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)((unsigned)e * 4u), (int)(k * fb), 0));
  }
  __device__ void st(int k, size_t e, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r,
                                          (int)((unsigned)e * 4u), (int)(k * fb), 0);
  }
  // the same with the non-temporal hint (streamed data the L2 need not keep)
  __device__ float ldn(int k, size_t e) const {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)((unsigned)e * 4u), (int)(k * fb), kNT));
  }
  __device__ void stn(int k, size_t e, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r,
                                          (int)((unsigned)e * 4u), (int)(k * fb), kNT);
  }
};

// A set of N float fields of one skew stored element by element: element
// e's N values at e*N .. e*N+N-1, so that one buffer load reads several
// fields (the mixed mode's period, T0 and adjoint coefficient sets).  A
// step's loads of a set share one 32-bit offset (plans keep a set within
// 4 GiB, kMaxSkewElems32); the loads of several consecutive fields are one
// 64-, 96- or 128-bit load.
typedef unsigned u32x3 __attribute__((ext_vector_type(3)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
template <int N>
struct CellSet {
  __amdgpu_buffer_rsrc_t r;
  __device__ CellSet(const GMEM void *base, size_t se)
      : r(buffer_over((const GMEM float *)base, base ? (unsigned)(N * se * sizeof(float)) : 0u)) {}
  __device__ static int off(size_t e, int k) { return (int)(((unsigned)e * N + k) * 4u); }
  __device__ float ld1(size_t e, int k) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off(e, k), 0, 0));
  }
  __device__ u32x2 ld2(size_t e, int k) const {
    return __builtin_amdgcn_raw_buffer_load_b64(r, off(e, k), 0, 0);
  }
  __device__ u32x3 ld3(size_t e, int k) const {
    return __builtin_amdgcn_raw_buffer_load_b96(r, off(e, k), 0, 0);
  }
  __device__ u32x3 ld3n(size_t e, int k) const {
    return __builtin_amdgcn_raw_buffer_load_b96(r, off(e, k), 0, kNT);
  }
  __device__ u32x4 ld4(size_t e, int k) const {
    return __builtin_amdgcn_raw_buffer_load_b128(r, off(e, k), 0, 0);
  }
};
__device__ __forceinline__ float as_f(unsigned u) { return __builtin_bit_cast(float, u); }

// Copies the grid cells of a field from skew FROM into the other skew, one
// whole destination row per step (coalesced stores).  The loads of one row
// gather from 64 source rows; lane l's source rows are its own column shifted
// by 63 - 2l rows, so a source line serves 32 lanes on 32 destination rows
// within 64.  The loads of kU destination rows are issued together, so that
// a line's uses fall within few batches and the batch's latency is paid
// once.  Loads and stores go through buffer resources over the strip: rows
// outside it read 0 and stores to padding cells are dropped, with no branch.
// Padding elements of the destination are left as they are.  The stores
// carry the non-temporal hint (the loads do not: a source line is read again
// by the next batches).
// Rows per batch: 16 in FP32, 8 in FP64 (measured at 401^2 x 4000 items and
// 41^2 x 1600: 32 rows cost the FP32 kernels an occupancy step on the small
// grid and gained nothing on the large one).
#ifndef SATT_COPY_KU_F32
#define SATT_COPY_KU_F32 16
#endif
#ifndef SATT_COPY_KU_F64
#define SATT_COPY_KU_F64 8
#endif
template <typename V, int FROM>
__device__ void skew_copy(const Team &T, const GMEM V *__restrict__ src, GMEM V *__restrict__ dst, int nx,
                          int ny, int R, int ns, int lane) {
  constexpr int TO = 1 - FROM;
  constexpr int kU = sizeof(V) == 4 ? SATT_COPY_KU_F32 : SATT_COPY_KU_F64;
  constexpr int kRow = 64 * (int)sizeof(V);  // bytes of one skew row
  const unsigned strip_bytes = (unsigned)R * kRow;
  for (int c = T.w; c < ns; c += T.W) {
    const bool col = 64 * c + lane < nx;
    // destination row r holds the lane's cell j = r - jof, which sits in
    // source row j + sof
    const int jof = TO == 0 ? lane : 63 - lane;
    const int sof = FROM == 0 ? lane : 63 - lane;
    const __amdgpu_buffer_rsrc_t rs = buffer_over(src + (size_t)c * R * 64, strip_bytes);
    const __amdgpu_buffer_rsrc_t rd = buffer_over(dst + (size_t)c * R * 64, strip_bytes);
    for (int r0 = 0; r0 < R; r0 += kU) {
      // source row of destination row r0 + u: r0 + u - jof + sof
      const int src0 = ((r0 - jof + sof) * 64 + lane) * (int)sizeof(V);
      V v[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if constexpr (sizeof(V) == 4)
          v[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, src0 + u * kRow, 0, 0));
        else
          v[u] = __builtin_bit_cast(double,
                                    __builtin_amdgcn_raw_buffer_load_b64(rs, src0 + u * kRow, 0, 0));
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int j = r0 + u - jof;
        const int off = col && j >= 0 && j < ny ? ((r0 + u) * 64 + lane) * (int)sizeof(V)
                                                : kOutOfRange;
        if constexpr (sizeof(V) == 4)
          __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v[u]), rd, off, 0, kNT);
        else
          __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v[u]), rd, off, 0, kNT);
      }
    }
  }
  __syncthreads();
}

// Skew row of step t (uniform across lanes), see the layout note above.
template <int XD, int YD>
__device__ __forceinline__ int skew_row(int t, int ny, int L) {
  constexpr int SK = XD == YD ? 0 : 1;
  if (SK == 0) return YD > 0 ? t : ny + L - 2 - t;
  return YD > 0 ? t - L + 64 : ny + 62 - t;
}

// Where a lane loads its operands of step t in a strip pass: its own cell
// while that cell is in the grid, else, in the same (clamped) row, the cell
// of the nearest lane whose cell is.  Every load is then valid and none needs
// a branch, and a step touches only the cache lines of its row that hold
// cells of the step.
template <int XD, int YD>
__device__ __forceinline__ size_t step_pos(int c, int t, int lane, int ny, int L, int R) {
  const int r = min(max(skew_row<XD, YD>(t, ny, L), 0), R - 1);
  const int lo = XD > 0 ? max(0, t - ny + 1) : max(0, L - 1 - t);
  const int hi = XD > 0 ? min(L - 1, t) : min(L - 1, L - 2 + ny - t);
  const int l = min(max(min(max(lane, lo), hi), 0), 63);
  return skidx(c, r, l, R);
}

// The operands a lane needs for its cell of step t.  "Far" operands are the
// ones other lanes read one step early (tau, T0) plus the strip-edge and
// convergence values; they are fetched two steps ahead.  "Near" operands are
// the cell's own coefficients, fetched one or two steps ahead.  A lane whose
// cell of that step is outside the grid (pipeline fill and drain, padding
// lanes of a partial strip) loads a padding cell of the strip; the values
// other lanes read from it (lane shifts) are never used for a cell in the
// grid.
template <typename V>
struct FFar {
  V tau, t0;    // old tau and T0 of the lane's cell
  V tdn, zdn;   // downstream neighbour column (edge lane): old tau, T0
  V snap;       // tau at the start of the iteration (last sweep only)
  int j;
  bool act;
};
// The cell's coefficients, the period fields and T0x, T0y, in their stored
// type P: float for the mixed mode's FP64 polish, converted where the update
// reads them (a conversion at the load would wait for the load, steps
// early).
template <typename P>
struct FNear {
  P a, b, c, f, dx_, dy_, t0x, t0y;
};

// Everything one forward solve's strip passes need.
template <typename V>
struct FwdCtx {
  int nx, ny, R, lane;
  size_t SE;
  double dx, dy;           // FP64 update
  V rdx, rdy;              // upwind update: 1/dx, 1/dy in V
  int fi0, fi1, fj0, fj1;  // frozen rectangle
  // period skews (plus, minus), kPF fields each, and item T0 skews (plus,
  // minus), kFF fields each (T0, T0x, T0y), whose T0x and T0y the passes
  // read: in V, except for the mixed mode's FP64 polish (UP in double),
  // which reads the float fields of its FP32 phase: the polish's arithmetic
  // is FP64 on coefficients rounded once to float, half the bytes of FP64
  // ones
  const GMEM void *ps[2];
  const GMEM void *gs[2];
  const GMEM V *is[2];     // item T0 skews (plus, minus) whose T0 the passes read
  GMEM V *snap;            // plus skew
};


// One sweep (XD, YD) over the k-th strip of the sweep, updating tau in place
// in the sweep's own skew: the reference's update (UP false, FP64) or the
// mixed mode's upwind update on dtau (UP true).  MODE 0: plain; 1: record
// the snapshot (first sweep of an iteration); 2: the strip's sum and max of
// |tau-snap|*T0 in double (last sweep) into ss, sm.  PD: how many steps
// ahead the operands are fetched.
template <typename V, int XD, int YD, int MODE, bool UP, int PD>
__device__ void fwd_strip_pass(const FwdCtx<V> &C, const Team &T, int k, GMEM V *tau, double &ss,
                               double &sm) {
  constexpr int SK = XD == YD ? 0 : 1;
  const int nx = C.nx, ny = C.ny, R = C.R, lane = C.lane;
  const int c = XD > 0 ? k : T.ns - 1 - k;
  const size_t SE = C.SE;
  using PS = typename std::conditional<UP && sizeof(V) == 8, float, V>::type;
  // float period and T0 sets are stored element by element (CellSet), FP64
  // ones field by field
  constexpr bool kF = sizeof(PS) == 4;
  const FieldSet<PS> fps((const GMEM PS *)C.ps[SK], SE, kF ? 0 : kPF),
      fgs((const GMEM PS *)C.gs[SK], SE, kF ? 0 : kFF);
  // the float period fields as two sets of three, a, b, c and fun, dis_x,
  // dis_y (the adjoint's coefficient build reads the first only)
  const GMEM float *ps32 = (const GMEM float *)C.ps[SK];
  const CellSet<3> cpa(kF ? ps32 : nullptr, SE), cpf(kF ? ps32 + 3 * SE : nullptr, SE);
  const CellSet<kFF> cgs(kF ? C.gs[SK] : nullptr, SE);
  const FieldSet<V> fis(C.is[SK], SE, kFF), ftau(tau, SE, 1), fsnap(C.snap, SE, 1);
  const int L = min(64, nx - 64 * c);
  const int i = 64 * c + lane;
  const bool lv = lane < L;
  const int lag = XD > 0 ? lane : L - 1 - lane;
  const int nsteps = ny + L - 1;
  // strip edges: the upstream neighbour column is final (its strip ran
  // before this one), the downstream one still holds the previous sweep's
  // values
  const bool up_edge = XD > 0 ? lane == 0 : lane == L - 1;
  const bool dn_edge = XD > 0 ? lane == L - 1 : lane == 0;
  const int iu = i - XD, id = i + XD;
  const bool up_ok = lv && up_edge && iu >= 0 && iu < nx;
  // the upwind update reads upwind neighbours only
  constexpr bool kDown = !UP;
  static_assert(UP || std::is_same<V, double>::value, "the full update is FP64");
  const bool dn_ok = kDown && lv && dn_edge && id >= 0 && id < nx;
  // the upstream strip's edge column comes from the previous wave's channel
  // (rows counted over that wave's strips of the sweep); the lane next to
  // the downstream strip publishes this strip's
  const int pw = (T.w + T.W - 1) % T.W, base_in = ((k - 1) / T.W) * ny;
  const bool pub = k < T.ns - 1 && lag == L - 1;
  const int base_out = (k / T.W) * ny;
  double acc_s = 0, acc_m = 0;

  // Every lane loads at every step (at step_pos), and no global load sits
  // behind a branch: a load that a branch may skip leaves the compiler
  // unsure how many loads are in flight, and it then waits for all of them,
  // the prefetched ones included.
  // The far and near operands of step t (both fetched PD steps ahead).
  auto load_ops = [&](int t, FFar<V> &q, FNear<PS> &n) {
    const int u = t - lag;
    const int j = YD > 0 ? u : ny - 1 - u;
    q.j = j;
    q.act = lv && u >= 0 && u < ny;
    const size_t e = step_pos<XD, YD>(c, t, lane, ny, L, R);
    if constexpr (kF) {
      const u32x3 pa = cpa.ld3(e, 0);
      const u32x3 pf = cpf.ld3(e, 0);
      const u32x3 g3 = cgs.ld3n(e, 0);
      n.a = as_f(pa.x); n.b = as_f(pa.y); n.c = as_f(pa.z);
      n.f = as_f(pf.x); n.dx_ = as_f(pf.y); n.dy_ = as_f(pf.z);
      n.t0x = as_f(g3.y); n.t0y = as_f(g3.z);
      if constexpr (sizeof(V) == 4)
        q.t0 = as_f(g3.x);
      else
        q.t0 = fis.ldn(0, e);  // the polish's FP64 T0
    } else {
      q.t0 = fis.ld(0, e);
      n.a = fps.ld(0, e); n.b = fps.ld(1, e); n.c = fps.ld(2, e); n.f = fps.ld(3, e);
      n.dx_ = fps.ld(4, e); n.dy_ = fps.ld(5, e);
      n.t0x = fgs.ld(1, e); n.t0y = fgs.ld(2, e);
    }
    if constexpr (kF) {
      q.tau = ftau.ldn(0, e);
      if (MODE == 2) q.snap = fsnap.ldn(0, e);
    } else {
      q.tau = ftau.ld(0, e);
      if (MODE == 2) q.snap = fsnap.ld(0, e);
    }
    if constexpr (kDown) {
      const size_t ed = opaque_index(dn_ok ? skpos(SK, id, min(max(j, 0), ny - 1), R) : e);
      q.tdn = ftau.ld(0, ed);
      q.zdn = fis.ld(0, ed);
    }
  };
  // The upstream strip's edge (final tau and its T0) for the cell of step t,
  // from the channel; read at the step itself, ahead of the step's global
  // loads, so that the team's wait loop sits between none of them and their
  // use.
  auto load_up = [&](int t, bool act, V &tu, V &zu) {
    if (up_ok && act) {  // k > 0
      if (T.multi()) team_wait(T, pw, base_in + t - lag + 1);
      tu = team_read<V>(T, pw, t - lag);
      zu = team_read0<V>(T, pw, t - lag);
    }
  };

  // Far and near operands are fetched PD steps ahead, into S = PD+2
  // operand sets that the loop cycles, unrolled by S (step t uses set
  // (t+PD) mod S and fills set (t+2PD) mod S), so that no set is copied: a
  // copy of a set still in flight waits for its loads.  With PD+1 sets the
  // set that a step refills is the previous step's, whose tau and T0 are
  // still live as that step's result, and the compiler rotates registers
  // with copies of fresh loads at the loop's end, which wait for them.
  // The loop starts PD steps early; those steps only fetch (every lane is
  // outside the grid), so the loop has a single entry with nothing in
  // flight: a separate prologue of loads, issued in whatever order the
  // scheduler picks, makes the wait counts derived for the loop assume the
  // worst order, and the loop then waits for all but the newest step's
  // loads.  Up to S-1 steps past the last run with every lane outside the
  // grid too, so they change nothing.
  static_assert(PD >= 2, "the next step's far set is read while this step fills one");
  constexpr int S = PD + (UP ? 2 : 1);
  FFar<V> fs[S] = {};
  FNear<PS> ns[S] = {};
  V tprev = 0, zprev = 0;  // own result and T0 of the previous step
  auto step = [&](int t, const FFar<V> &fc, const FFar<V> &fn, FFar<V> &f2,
                  const FNear<PS> &nc, FNear<PS> &nx2, int tnear) {
    V tU = 0, zU = 0;
    load_up(t, fc.act, tU, zU);
    load_ops(t + PD, f2, nx2);
    (void)tnear;
    // lane shifts read other lanes' registers: they run with every lane
    // enabled, outside the per-cell branch
    V tUx = lane_shift<-XD>(tprev), zUx = lane_shift<-XD>(zprev);
    if (up_ok) { tUx = tU; zUx = zU; }
    V tDx = 0, zDx = 0;
    if constexpr (kDown) {
      tDx = lane_shift<XD>(fn.tau);
      zDx = lane_shift<XD>(fn.t0);
      if (dn_ok) { tDx = fc.tdn; zDx = fc.zdn; }
    }
    const int jc = fc.j;
    V tnew = fc.tau;
    if (fc.act) {
      const bool frozen = i >= C.fi0 && i <= C.fi1 && jc >= C.fj0 && jc <= C.fj1;
      if constexpr (kDown) {
        const V tW = XD > 0 ? tUx : tDx, tE = XD > 0 ? tDx : tUx;
        const V zW = XD > 0 ? zUx : zDx, zE = XD > 0 ? zDx : zUx;
        const V tS = YD > 0 ? tprev : fn.tau, tN = YD > 0 ? fn.tau : tprev;
        const V zS = YD > 0 ? zprev : fn.t0, zN = YD > 0 ? fn.t0 : zprev;
        if (!frozen)
          tnew = fwd_cell_r(i, jc, nx, ny, C.dx, C.dy, nc.a, nc.b, nc.c, nc.f, nc.dx_, nc.dy_,
                            fc.t0, nc.t0x, nc.t0y, fc.tau, tW, tE, tS, tN, zW, zE, zS, zN);
      } else {
        if (!frozen)
          tnew = fwd_cell_up_r<V, XD, YD>(i, jc, nx, ny, C.rdx, C.rdy, (V)nc.a, (V)nc.b, (V)nc.c,
                                          (V)nc.f, (V)nc.dx_, (V)nc.dy_, fc.t0, (V)nc.t0x,
                                          (V)nc.t0y, fc.tau, tUx, tprev, zUx, zprev);
      }
      const size_t e = skidx(c, skew_row<XD, YD>(t, ny, L), lane, R);
      if constexpr (kF) {
        ftau.stn(0, e, tnew);
        if (MODE == 1) fsnap.stn(0, e, fc.tau);
      } else {
        ftau.st(0, e, tnew);
        if (MODE == 1) fsnap.st(0, e, fc.tau);
      }
      if (pub) team_publish2<V>(T, t - lag, tnew, fc.t0, base_out + t - lag + 1);
      if (MODE == 2) {
        const double d = fabs((double)tnew - (double)fc.snap) * (double)fc.t0;
        acc_s += d;
        acc_m = d > acc_m ? d : acc_m;
      }
    }
    if constexpr (kDown) {
      // PD+1 sets: the set refilled next step is this step's
      tprev = own_reg(tnew);
      zprev = own_reg(fc.t0);
    } else {
      tprev = tnew;
      zprev = fc.t0;
    }
  };
  for (int t = -PD; t < nsteps; t += S) {
#pragma unroll
    for (int q = 0; q < S; ++q)
      step(t + q, fs[q], fs[(q + 1) % S], fs[(q + PD) % S], ns[q], ns[(q + PD) % S], t + q + PD);
  }
  if (MODE == 2) {
    ss = wave_sum(acc_s);
    sm = wave_max(acc_m);
  }
  wave_fence();
}

template <typename V, int XD, int YD, int MODE, bool UP, int PD>
__device__ void fwd_sweep(const FwdCtx<V> &C, Team &T, GMEM V *tau, StripAcc &acc) {
  team_sweep_begin(T);
  for (int k = T.w; k < T.ns; k += T.W) {
    double ss = 0, sm = 0;
    fwd_strip_pass<V, XD, YD, MODE, UP, PD>(C, T, k, tau, ss, sm);
    if (MODE == 2) team_strip_term(T, k, ss, sm, 0.0, acc);
  }
  team_sweep_end(T);
}

// One iteration of four sweeps, (x+,y+), (x+,y-), (x-,y+), (x-,y-)
// (src/eikon_solver_2d.f90:884-898), starting and ending in the plus skew:
// the first sweep in the plus skew, the middle two in the minus skew, the
// last in the plus skew, with the field copied between the skews at the
// two family changes.  Returns whether the stop rule (:1091-1115) holds.
template <typename V, bool UP, int PD>
__device__ bool fwd_iteration(const FwdCtx<V> &C, Team &T, GMEM V *tP, GMEM V *tM) {
  StripAcc acc;
  fwd_sweep<V, 1, 1, 1, UP, PD>(C, T, tP, acc);
  skew_copy<V, 0>(T, tP, tM, C.nx, C.ny, C.R, T.ns, C.lane);
  fwd_sweep<V, 1, -1, 0, UP, PD>(C, T, tM, acc);
  fwd_sweep<V, -1, 1, 0, UP, PD>(C, T, tM, acc);
  skew_copy<V, 1>(T, tM, tP, C.nx, C.ny, C.R, T.ns, C.lane);
  fwd_sweep<V, -1, -1, 2, UP, PD>(C, T, tP, acc);
  const StripAcc r = team_strip_total(T, acc);
  const double l1 = r.s / (double)C.nx / (double)C.ny;
  // the FP32 phase stops at kMixFwdTolScale times the reference's tolerance:
  // the FP64 polish then takes the field to the reference's own
  const double tol = std::is_same<V, float>::value ? kMixFwdTolScale * kFwdTol : kFwdTol;
  return fabs(l1) < tol && r.m < tol;
}

// Per-slot FP32 region of the mixed forward, in float skews from double
// skew 9: T0, T0x, T0y plus (0..2) and minus (3..5), each skew's three
// stored element by element (CellSet<3>), dtau plus (6) and minus (7),
// snapshot (8).  Double skews 9..13 are unused by the forward until the
// output T skews (12, 13) are written after the FP32 phase.
constexpr int kSkF32 = 9;

// sin and cos of the grid's latitudes, once per grid.
__global__ void k_grid_trig(Grid g, double *sy, double *cy) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < g.ny) {
    sy[j] = sin(g.yy[j]);
    cy[j] = cos(g.yy[j]);
  }
}

// Per-item forward tables, one thread per column / row: sin(xx - x0),
// cos(xx - x0), the frozen flags abs((xx-x0)/dx) <= 1 and abs((yy-y0)/dy) <= 1
// (src/eikon_solver_2d.f90:859-872), and fun0 (:825-841), sin(y0), cos(y0).
// An item whose source bracket leaves the grid gets its error status.
__global__ __launch_bounds__(64) void k_fwd_prep(const SlotDesc *sd) {
  const SlotDesc D = sd[blockIdx.y];
  const Grid &g = D.g;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  const double x0 = deg2rad_s(D.x0), y0 = deg2rad_s(D.y0);
  double *aux = D.aux;
  if (k < nx) {
    aux[k] = sin(g.xx[k] - x0);
    aux[nx + k] = cos(g.xx[k] - x0);
    aux[2 * nx + k] = fabs((g.xx[k] - x0) / g.dx) <= 1 ? 1.0 : 0.0;
  }
  if (k < ny) aux[3 * nx + k] = fabs((g.yy[k] - y0) / g.dy) <= 1 ? 1.0 : 0.0;
  if (k == 0) {
    double fun0 = 0;
    const int st0 = fwd_fun0_s(g, D.fun, x0, y0, &fun0);
    aux[3 * nx + ny] = fun0;
    aux[3 * nx + ny + 1] = sin(y0);
    aux[3 * nx + ny + 2] = cos(y0);
    if (st0 != kOk) *D.status = st0;
  }
}

// T0, T0x, T0y of one cell (src/eikon_solver_2d.f90:845-857) from the
// row's sin/cos of the latitude and the column's sin/cos of the longitude
// difference: the expressions of t0_cell_s, with the same operand values.
__device__ __forceinline__ void t0_cell_tab(double sy, double cy, double sdx, double cdx,
                                            double sy0, double cy0, double fun0, double &T0,
                                            double &T0x, double &T0y) {
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

// Tiled per-element launches (the forward setup, the adjoint coefficient
// build): a workgroup does a tile of kTileElems consecutive elements of one
// skew of one slot (32 rows of a strip), so that the rows above and below a
// row, which the build's cells read, are mostly the tile's own, and the row
// of each element follows from the tile's first row without a division per
// element.  The workgroups are dealt to the XCDs as the launch group's slots
// are (workgroup w runs on XCD w mod kXcds and slot s belongs to XCD s mod
// kXcds), tile by tile: the items that share an XCD then read each tile of
// their period's fields through its L2 one after another.
constexpr int kTileElems = 2048;

unsigned tiles_per_skew(const GroupDims &gd) {
  return (unsigned)((gd.max_se + kTileElems - 1) / kTileElems);
}
dim3 tile_grid(const GroupDims &gd) {
  return dim3((unsigned)(kXcds * ((gd.count + kXcds - 1) / kXcds) * 2 * tiles_per_skew(gd)));
}

// The slot and tile (skew-major: tile t is tile t mod tps of skew t / tps)
// of this workgroup; false if it has none.
__device__ __forceinline__ bool tile_slot(int count, unsigned tps, int &slot, unsigned &t) {
  const int x = (int)(blockIdx.x % kXcds);
  const unsigned k = blockIdx.x / kXcds;
  const int cnt = (count - x + kXcds - 1) / kXcds;  // slots of XCD x
  if (cnt <= 0) return false;
  t = k / (unsigned)cnt;
  if (t >= 2 * tps) return false;
  slot = x + kXcds * (int)(k % (unsigned)cnt);
  return true;
}

// f(e, cc, r, l) for this thread's elements (strip cc, row r, lane l) of
// tile ts of a skew; a tile's 32 rows cross into the next strip at most once
// (R > 64).
template <typename F>
__device__ __forceinline__ void tile_elems(const SkewGeom &sg, unsigned ts, F &&f) {
  const size_t SE = sg.skew_elems();
  const int R = sg.R;
  const unsigned row0 = ts * (unsigned)(kTileElems / 64);
  const int cc0 = (int)(row0 / (unsigned)R), r0 = (int)(row0 % (unsigned)R);
  const int l = (int)(threadIdx.x & 63);
  for (int q = 0; q < kTileElems / 256; ++q) {
    int r = r0 + 4 * q + (int)(threadIdx.x >> 6), cc = cc0;
    if (r >= R) {
      r -= R;
      ++cc;
    }
    const size_t e = ((size_t)cc * R + r) * 64 + l;
    if (e >= SE) break;
    f(e, cc, r, l);
  }
}

// Forward setup of every item of a launch group, in XCD-dealt tiles:
// T0, T0x, T0y in both skews (and, MIXED, rounded to float), and the tau
// init in the plus skew (:859-872: 1 on the frozen cells, 10 elsewhere;
// MIXED: dtau = tau - 1 in float).  Padding elements get zeros.  A frozen
// cell on the boundary is the reference's 'source on the boundary' error
// (:863-867).
template <bool MIXED>
__global__ __launch_bounds__(256) void k_fwd_setup(const SlotDesc *sd, int count, unsigned tps) {
  int slot;
  unsigned t;
  if (!tile_slot(count, tps, slot, t)) return;
  const SlotDesc D = sd[slot];
  if (*D.status < 0) return;
  const Grid &g = D.g;
  const SkewGeom &sg = D.sg;
  const size_t SE = sg.skew_elems();
  const int sk = (int)(t / tps);
  const int nx = g.nx, ny = g.ny;
  const GMEM double *aux = gm(D.aux);
  const GMEM double *tail = aux + 3 * nx + ny;
  const double fun0 = tail[0], sy0 = tail[1], cy0 = tail[2];
  GMEM double *sk0 = gm(D.skew);
  tile_elems(sg, t % tps, [&](size_t e, int cc, int r, int l) {
    const int i = 64 * cc + l;
    const int j = sk == 0 ? r - l : r + l - 63;
    const bool valid = i < nx && j >= 0 && j < ny;
    double t0 = 0, t0x = 0, t0y = 0;
    bool frz = false;
    if (valid) {
      t0_cell_tab(gm(g.sy)[j], gm(g.cy)[j], aux[i], aux[nx + i], sy0, cy0, fun0, t0, t0x, t0y);
      frz = aux[2 * nx + i] != 0.0 && aux[3 * nx + j] != 0.0;
      if (frz && (i == 0 || i == nx - 1 || j == 0 || j == ny - 1)) *D.status = kESrcOnBoundary;
    }
    GMEM double *b0 = sk0 + (size_t)sk * kFF * SE;
    b0[e] = t0;
    if (!MIXED) {  // the mixed mode's passes read the gradients in float
      b0[SE + e] = t0x;
      b0[2 * SE + e] = t0y;
    }
    if (MIXED) {
      GMEM float *f32 = (GMEM float *)(sk0 + (size_t)kSkF32 * SE);
      GMEM float *b32 = f32 + ((size_t)sk * SE + e) * kFF;  // element by element
      b32[0] = (float)t0; b32[1] = (float)t0x; b32[2] = (float)t0y;
      if (sk == 0) f32[6 * SE + e] = frz ? 0.0f : 9.0f;
    } else if (sk == 0) {
      sk0[kSkTauP * SE + e] = frz ? 1.0 : 10.0;
    }
  });
}

template <typename V>
__device__ void fwd_ctx_init(FwdCtx<V> &C, const Grid &g, const SkewGeom &sg, int lane,
                             const int rect[4]) {
  C.nx = g.nx; C.ny = g.ny; C.R = sg.R; C.lane = lane; C.SE = sg.skew_elems();
  C.dx = g.dx; C.dy = g.dy;
  C.rdx = (V)(1.0 / g.dx); C.rdy = (V)(1.0 / g.dy);
  C.fi0 = rect[0]; C.fi1 = rect[1]; C.fj0 = rect[2]; C.fj1 = rect[3];
}

// FP64 iterations from the field in the plus skew to the stop rule, then
// T = tau*T0 (src/eikon_solver_2d.f90:1124) in both skews and in natural
// layout: the reference's update on tau (UP false), or the mixed mode's
// upwind update on dtau, with T = (1 + dtau)*T0 (UP true).
template <bool UP, int PD>
__device__ void fwd_finish64(const SlotDesc &D, Team &T, int lane, const int rect[4]) {
  const Grid &g = D.g;
  const SkewGeom &sg = D.sg;
  const int nx = g.nx, ny = g.ny, R = sg.R, ns = sg.ns;
  const size_t N = (size_t)nx * ny, SE = sg.skew_elems();
  double *sk0 = D.skew, *Tnat = D.T;
  double *tauP = sk0 + kSkTauP * SE, *tauM = sk0 + kSkTauM * SE;
  double *TP = sk0 + kSkTP * SE, *TM = sk0 + kSkTM * SE;
  FwdCtx<double> C;
  fwd_ctx_init(C, g, sg, lane, rect);
  if (UP) {  // the mixed mode's polish: the float period fields
    C.ps[0] = gm(D.pskew32);
    C.ps[1] = gm(D.pskew32 + (size_t)kPF * SE);
  } else {
    C.ps[0] = gm(D.pskew);
    C.ps[1] = gm(D.pskew + (size_t)kPF * SE);
  }
  C.is[0] = gm(sk0);
  C.is[1] = gm(sk0 + kFF * SE);
  if (UP) {  // and the float T0 gradients
    const GMEM float *f32 = reinterpret_cast<const GMEM float *>(gm(sk0) + (size_t)kSkF32 * SE);
    C.gs[0] = f32;
    C.gs[1] = f32 + (size_t)kFF * SE;
  } else {
    C.gs[0] = C.is[0];
    C.gs[1] = C.is[1];
  }
  C.snap = gm(sk0 + kSkSnapF * SE);
  int iter = 1;
  bool conv = false;
  for (; iter <= kFwdMaxIter; ++iter)
    if (fwd_iteration<double, UP, PD>(C, T, gm(tauP), gm(tauM))) { conv = true; break; }
  if (!conv) iter = kFwdMaxIter;
  auto tau_of = [&](double v) { return UP ? 1.0 + v : v; };
  const GMEM double *gtau = gm(tauP);
  GMEM double *gTP = gm(TP), *gTM = gm(TM), *gT = gm(Tnat);
  for (int cc = T.w; cc < ns; cc += T.W) {
    const int i = 64 * cc + lane;
    for (int r = 0; r < R; ++r) {
      const size_t e = skidx(cc, r, lane, R);
      gTP[e] = tau_of(gtau[e]) * C.is[0][e];
      const int j = r + lane - 63;
      double tm = 0;
      if (i < nx && j >= 0 && j < ny) tm = tau_of(gtau[skpos(0, i, j, R)]) * C.is[1][e];
      gTM[e] = tm;
    }
  }
  __syncthreads();
  for (size_t k = threadIdx.x; k < N; k += blockDim.x) {
    const int i = (int)(k % nx), j = (int)(k / nx);
    gT[k] = gTP[skpos(0, i, j, R)];
  }
  if (threadIdx.x == 0) {
    *D.status = team_failed(T) ? kETeamWait : conv ? kOk : kWMaxIter;
    D.iters[0] = iter;
  }
}

// The frozen rectangle of an item whose setup succeeded.
__device__ bool fwd_item_rect(const SlotDesc &D, int lane, int rect[4]) {
  const double x0 = deg2rad_s(D.x0), y0 = deg2rad_s(D.y0);
  double fun0;
  return fwd_source_rect(D.g, D.fun, x0, y0, lane, fun0, rect) == kOk;
}

// The strip kernels run one item per workgroup of W = blockDim.x/64 waves
// (a team), with team_lds_bytes of dynamic LDS when W > 1, fetching the
// sweeps' operands PD steps ahead.
template <bool TEAM, int PD>
__global__ __launch_bounds__(TEAM ? 64 * kMaxTeam : 64) void k_fwd_strip(const SlotDesc *sd) {
  extern __shared__ double lds[];
  const SlotDesc D = sd[blockIdx.x];
  const int lane = threadIdx.x & 63;
  if (*D.status < 0) return;  // the setup failed
  int rect[4];
  if (!fwd_item_rect(D, lane, rect)) return;
  Team T = team_init<TEAM>(lds, D.g.ny, D.sg.ns, 8);
  fwd_finish64<false, PD>(D, T, lane, rect);
}

// Mixed forward, FP32 phase: the FP32 iterations (count to iters[3]), then
// dtau widened into the FP64 plus skew for the polish.
template <bool TEAM, int PD>
__global__ __launch_bounds__(TEAM ? 64 * kMaxTeam : 64) void k_fwd_mixed32(const SlotDesc *sd) {
  extern __shared__ double lds[];
  const SlotDesc D = sd[blockIdx.x];
  const int lane = threadIdx.x & 63;
  if (*D.status < 0) return;
  const SkewGeom &sg = D.sg;
  const size_t SE = sg.skew_elems();
  int rect[4];
  if (!fwd_item_rect(D, lane, rect)) return;
  Team T = team_init<TEAM>(lds, D.g.ny, sg.ns, 4);
  double *sk0 = D.skew;
  float *f32 = reinterpret_cast<float *>(sk0 + (size_t)kSkF32 * SE);
  FwdCtx<float> C;
  fwd_ctx_init(C, D.g, sg, lane, rect);
  C.ps[0] = gm(D.pskew32);
  C.ps[1] = gm(D.pskew32 + (size_t)kPF * SE);
  C.is[0] = gm(f32);
  C.is[1] = gm(f32 + (size_t)kFF * SE);
  C.gs[0] = C.is[0];
  C.gs[1] = C.is[1];
  C.snap = gm(f32 + 8 * SE);
  float *dP = f32 + 6 * SE, *dM = f32 + 7 * SE;
  int iter = 1;
  for (; iter <= kFwdMaxIter; ++iter)
    if (fwd_iteration<float, true, PD>(C, T, gm(dP), gm(dM))) break;
  if (iter > kFwdMaxIter) iter = kFwdMaxIter;
  double *tauP = sk0 + kSkTauP * SE;
  for (size_t e = threadIdx.x; e < SE; e += blockDim.x) gm(tauP)[e] = (double)gm(dP)[e];
  if (threadIdx.x == 0) {
    D.iters[3] = iter;
    if (team_failed(T)) *D.status = kETeamWait;
  }
}

// Mixed forward, FP64 polish: the upwind update in FP64 from the FP32 result
// to the reference's stop rule, and the outputs.
template <bool TEAM, int PD>
__global__ __launch_bounds__(TEAM ? 64 * kMaxTeam : 64) void k_fwd_polish(const SlotDesc *sd) {
  extern __shared__ double lds[];
  const SlotDesc D = sd[blockIdx.x];
  const int lane = threadIdx.x & 63;
  if (*D.status < 0) return;
  int rect[4];
  if (!fwd_item_rect(D, lane, rect)) return;
  Team T = team_init<TEAM>(lds, D.g.ny, D.sg.ns, 8);
  fwd_finish64<true, PD>(D, T, lane, rect);
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

// The mixed mode's update weights of one interior cell, for the form
// Ta = k0 + kW*TaW + kE*TaE + kS*TaS + kN*TaN: the face coefficients of
// adj_coef_v on the metric as stored in float (widened), with the divisions
// by dx, dy, 4dx and 4dy taken as products with their reciprocals and the
// halvings as products with 0.5; out = {kW, kE, kS, kN} in double (before
// the rounding to float) and 1/d, all 0 where abs(d) < eps.
__device__ __forceinline__ void adj_coef_k(double Tk, double TW, double TE, double TS, double TN,
                                           double TWS, double TWN, double TES, double TEN,
                                           double ak, double aW, double aE, double bk, double bS,
                                           double bN, double ck, double cW, double cE, double cS,
                                           double cN, double rdx, double rdy, double *out) {
  const double rdx4 = 0.25 * rdx, rdy4 = 0.25 * rdy;
  const double a1 = -((Tk - TW) * rdx * (ak + aW) * 0.5) + (TN - TS + TWN - TWS) * rdy4 * (ck + cW) * 0.5;
  const double a2 = -((TE - Tk) * rdx * (aE + ak) * 0.5) + (TEN - TES + TN - TS) * rdy4 * (cE + ck) * 0.5;
  const double b1 = -((Tk - TS) * rdy * (bk + bS) * 0.5) + (TE - TW + TES - TWS) * rdx4 * (ck + cS) * 0.5;
  const double b2 = -((TN - Tk) * rdy * (bN + bk) * 0.5) + (TEN - TWN + TE - TW) * rdx4 * (cN + ck) * 0.5;
  const double a1p = (a1 + fabs(a1)) * 0.5, a1m = (a1 - fabs(a1)) * 0.5;
  const double a2p = (a2 + fabs(a2)) * 0.5, a2m = (a2 - fabs(a2)) * 0.5;
  const double b1p = (b1 + fabs(b1)) * 0.5, b1m = (b1 - fabs(b1)) * 0.5;
  const double b2p = (b2 + fabs(b2)) * 0.5, b2m = (b2 - fabs(b2)) * 0.5;
  const double d = (a2p - a1m) * rdx + (b2p - b1m) * rdy;
  if (fabs(d) < kAdjEps) {
    out[0] = out[1] = out[2] = out[3] = out[4] = 0.0;
    return;
  }
  const double rd = 1.0 / d;
  out[0] = a1p * rdx * rd;
  out[1] = -a2m * rdx * rd;
  out[2] = b1p * rdy * rd;
  out[3] = -b2m * rdy * rd;
  out[4] = rd;
}

// Offsets, in elements of one skew, from the element of an interior cell
// (i, j) at lane l to those of (i - 1, j) and (i + 1, j); the cells
// (i -+ 1, j + dj) are dj rows (64 dj elements) on from those.  A row of a
// strip holds j + l (plus skew) or j - l + 63 (minus skew), and the strips
// are R rows apart, so the neighbour in the strip before (lane 0) or after
// (lane 63) is about a strip away.
__device__ __forceinline__ int nb_west(int sk, int l, int R) {
  const int s = sk == 0 ? 1 : -1;
  return l == 0 ? -64 * R + 63 * 64 * s + 63 : -64 * s - 1;
}
__device__ __forceinline__ int nb_east(int sk, int l, int R) {
  const int s = sk == 0 ? 1 : -1;
  return l == 63 ? 64 * R - 63 * 64 * s - 63 : 64 * s + 1;
}

// The adjoint coefficients of a launch group's items, in XCD-dealt tiles
// (src/eikon_solver_2d.f90:1348-1370), a cell's neighbours from its own skew.
// FP64 mode (k_adj_build): a1, a2, b1, b2, d, delta in skew fields 0..5 (plus
// skew) and 6..11 (minus skew), with delta left 0 for k_adj_source to write
// where the receivers touch it, and Ta = 100 inside, 0 on the boundary
// (:1368-1380), in the plus skew. Mixed mode (k_adj_build32): the update in
// the form Ta = k0 + kW*TaW + kE*TaE + kS*TaS + kN*TaN, its coefficients kW =
// a1p/dx/d, kE = -a2m/dx/d, kS = b1p/dy/d, kN = -b2m/dy/d computed in FP64
// (adj_coef_k, on the metric as stored in float) and stored in float, element
// by element, which both the FP32 phase and the FP64 polish read; k0 left 0
// for k_adj_source; and the float Ta init. Padding elements get zeros. Each
// tile is staged in LDS first: T over the tile and kBuildHalo elements either
// side, a, b, c over the tile and kBuildHaloP either side (both copied with
// 16-byte loads, all of a thread's issued before any is waited for), and the
// columns of the strips either side that lanes 0 and 63 read. The tile's cells
// then read their neighbours from LDS: the loads of a tile are in flight
// together, where a thread loading each cell's own neighbours waits for them
// cell by cell. Elements of the staging windows outside the skew are staged as
// 0 (a load past the buffer's range returns 0): no interior cell reads them.
constexpr int kBuildHalo = 192, kBuildHaloP = 128;
constexpr int kBuildT = kTileElems + 2 * kBuildHalo;   // staged T (doubles)
constexpr int kBuildPE = kTileElems + 2 * kBuildHaloP;  // staged metric elements
constexpr int kTileRows = kTileElems / 64;
// MIXED: the float k set from the float metric (a, b, c element by element);
// else the FP64 mode's coefficients from the FP64 metric (field by field).
template <bool MIXED>
__device__ __forceinline__ void adj_build_staged(const SlotDesc *sd, int count, unsigned tps) {
  using PT = typename std::conditional<MIXED, float, double>::type;
  __shared__ double sT[kBuildT];
  // a, b, c: element by element (MIXED) or field by field
  __shared__ PT sP[3 * kBuildPE];
  // T west of lane 0 and east of lane 63 for the tile's rows -1..kTileRows;
  // a and c there for rows 0..kTileRows-1
  __shared__ double sTw[kTileRows + 2], sTe[kTileRows + 2];
  __shared__ PT sAw[kTileRows], sCw[kTileRows], sAe[kTileRows], sCe[kTileRows];
  int slot;
  unsigned t;
  if (!tile_slot(count, tps, slot, t)) return;  // the same for the whole workgroup
  const SlotDesc D = sd[slot];
  if (*D.status < 0) return;
  const Grid &g = D.g;
  const int nx = g.nx, ny = g.ny, R = D.sg.R;
  const long SE = (long)D.sg.skew_elems();
  const int sk = (int)(t / tps);
  const unsigned ts = t % tps;
  const long e0 = (long)ts * kTileElems;
  GMEM double *sk0 = gm(D.skew);
  const GMEM double *Tf = sk0 + (sk == 0 ? kSkTP : kSkTM) * SE;
  // the metric of the skew: MIXED a, b, c at 3e (floats), else a at e,
  // b at SE + e, c at 2 SE + e (doubles)
  const GMEM PT *Pf;
  if constexpr (MIXED)
    Pf = gm(D.pskew32) + (size_t)sk * kPF * SE;
  else
    Pf = gm(D.pskew) + (size_t)sk * kPF * SE;
  // position of field f of element x in the staged metric and in memory
  auto sp = [](int f, long x) { return MIXED ? 3 * x + f : f * (long)kBuildPE + x; };
  auto gp = [&](int f, long x) { return MIXED ? 3 * x + f : f * SE + x; };
  const int tid = (int)threadIdx.x;
  // SE is a multiple of 64 and the windows start at multiples of 64
  // elements, so no 16-byte chunk straddles the skew's end
  {
    constexpr int kCT = kBuildT / 2;                            // 16-byte chunks of T
    constexpr int kCF = kBuildPE * (int)sizeof(PT) / 16;        // of one metric field
    constexpr int kCP = MIXED ? 3 * kBuildPE / 4 : 3 * kCF;     // of the metric
    constexpr int nT = (kCT + 255) / 256, nP = (kCP + 255) / 256;
    const __amdgpu_buffer_rsrc_t rT = buffer_over(Tf, (unsigned)(SE * 8));
    const __amdgpu_buffer_rsrc_t rP = buffer_over(Pf, (unsigned)(SE * 3 * sizeof(PT)));
    u32x4 vt[nT], vp[nP];
#pragma unroll
    for (int q = 0; q < nT; ++q) {
      const int c = tid + 256 * q;
      const long x = e0 - kBuildHalo + 2 * c;
      const bool in = c < kCT && x >= 0 && x < SE;
      vt[q] = __builtin_amdgcn_raw_buffer_load_b128(rT, in ? (int)(x * 8) : kOutOfRange, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < nP; ++q) {
      const int c = tid + 256 * q;
      bool in;
      long off;  // byte offset
      if constexpr (MIXED) {
        const long x = 3 * (e0 - kBuildHaloP) + 4 * c;
        in = c < kCP && x >= 0 && x < 3 * SE;
        off = x * 4;
      } else {
        const int f = c / kCF;
        const long x = e0 - kBuildHaloP + 2 * (c % kCF);
        in = c < kCP && x >= 0 && x < SE;
        off = (f * SE + x) * 8;
      }
      vp[q] = __builtin_amdgcn_raw_buffer_load_b128(rP, in ? (int)off : kOutOfRange, 0, 0);
    }
    const int ow0 = nb_west(sk, 0, R), oe63 = nb_east(sk, 63, R);
    if (tid < kTileRows + 2) {
      const long x = e0 + (long)(tid - 1) * 64 + ow0;
      sTw[tid] = x >= 0 && x < SE ? Tf[x] : 0.0;
    } else if (tid >= 64 && tid < 64 + kTileRows + 2) {
      const long x = e0 + (long)(tid - 65) * 64 + 63 + oe63;
      sTe[tid - 64] = x >= 0 && x < SE ? Tf[x] : 0.0;
    } else if (tid >= 128 && tid < 128 + kTileRows) {
      const long x = e0 + (long)(tid - 128) * 64 + ow0;
      const bool in = x >= 0 && x < SE;
      sAw[tid - 128] = in ? Pf[gp(0, x)] : PT(0);
      sCw[tid - 128] = in ? Pf[gp(2, x)] : PT(0);
    } else if (tid >= 192 && tid < 192 + kTileRows) {
      const long x = e0 + (long)(tid - 192) * 64 + 63 + oe63;
      const bool in = x >= 0 && x < SE;
      sAe[tid - 192] = in ? Pf[gp(0, x)] : PT(0);
      sCe[tid - 192] = in ? Pf[gp(2, x)] : PT(0);
    }
#pragma unroll
    for (int q = 0; q < nT; ++q) {
      const int c = tid + 256 * q;
      if (c < kCT) *(u32x4 *)&sT[2 * c] = vt[q];
    }
#pragma unroll
    for (int q = 0; q < nP; ++q) {
      const int c = tid + 256 * q;
      if (c < kCP) *(u32x4 *)((char *)sP + 16 * c) = vp[q];
    }
  }
  __syncthreads();
  const int l = tid & 63;
  const int ow = nb_west(sk, l, R), oe = nb_east(sk, l, R);
  const unsigned row0 = ts * (unsigned)kTileRows;
  const int cc0 = (int)(row0 / (unsigned)R), r0 = (int)(row0 % (unsigned)R);
  // one cell at a time: its operands come from LDS, and registers held for
  // several would cost the workgroup a slot on a CU
#pragma unroll 1
  for (int q = 0; q < kTileElems / 256; ++q) {
    const int ro = 4 * q + (tid >> 6);  // row within the tile
    int r = r0 + ro, cc = cc0;
    if (r >= R) {
      r -= R;
      ++cc;
    }
    const long e = e0 + ro * 64 + l;
    if (e >= SE) break;
    const int i = 64 * cc + l;
    const int j = sk == 0 ? r - l : r + l - 63;
    const bool interior = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
    double w[kAF] = {0, 0, 0, 0, 0, 0};
    if (interior) {
      const int xt = kBuildHalo + ro * 64 + l, xp = kBuildHaloP + ro * 64 + l;
      auto tw = [&](int dj) { return l == 0 ? sTw[ro + 1 + dj] : sT[xt + ow + 64 * dj]; };
      auto te = [&](int dj) { return l == 63 ? sTe[ro + 1 + dj] : sT[xt + oe + 64 * dj]; };
      auto P = [&](int f, int x) { return sP[sp(f, x)]; };
      const PT aW = l == 0 ? sAw[ro] : P(0, xp + ow), cW = l == 0 ? sCw[ro] : P(2, xp + ow);
      const PT aE = l == 63 ? sAe[ro] : P(0, xp + oe), cE = l == 63 ? sCe[ro] : P(2, xp + oe);
      if constexpr (MIXED)
        adj_coef_k(sT[xt], tw(0), te(0), sT[xt - 64], sT[xt + 64], tw(-1), tw(1), te(-1), te(1),
                   P(0, xp), aW, aE, P(1, xp), P(1, xp - 64), P(1, xp + 64), P(2, xp), cW, cE,
                   P(2, xp - 64), P(2, xp + 64), 1.0 / g.dx, 1.0 / g.dy, w);
      else
        adj_coef_v(sT[xt], tw(0), te(0), sT[xt - 64], sT[xt + 64], tw(-1), tw(1), te(-1), te(1),
                   P(0, xp), aW, aE, P(1, xp), P(1, xp - 64), P(1, xp + 64), P(2, xp), cW, cE,
                   P(2, xp - 64), P(2, xp + 64), g.dx, g.dy, w);
    }
    if constexpr (MIXED) {
      GMEM float *f32 = (GMEM float *)(sk0 + (size_t)kSkA32 * SE);
      GMEM float *k32 = f32 + ((size_t)sk * SE + e) * 5;  // element by element
      k32[0] = 0.0f;
#pragma unroll
      for (int f = 0; f < 4; ++f) k32[1 + f] = interior ? (float)w[f] : 0.0f;
      if (sk == 0) f32[10 * SE + e] = interior ? 100.0f : 0.0f;
    } else {
      GMEM double *b0 = sk0 + (size_t)sk * kAF * SE;
#pragma unroll
      for (int f = 0; f < kAF; ++f) b0[f * SE + e] = w[f];
      if (sk == 0) sk0[kSkTaP * SE + e] = interior ? 100.0 : 0.0;
    }
  }
}

__global__ __launch_bounds__(256) void k_adj_build32(const SlotDesc *sd, int count,
                                                      unsigned tps) {
  adj_build_staged<true>(sd, count, tps);
}
__global__ __launch_bounds__(256) void k_adj_build(const SlotDesc *sd, int count, unsigned tps) {
  adj_build_staged<false>(sd, count, tps);
}

// The delta source of a launch group's items (src/eikon_solver_2d.f90:
// 1320-1336) into their coefficient fields, one workgroup per item: each
// thread forms the terms of receiver corners (entry 4r + c, in the
// reference's order), then the first entry of every touched cell sums that
// cell's terms in entry order, from 0 as the reference's zeroed field does,
// and writes the cell's source, interior cells only: delta itself in both
// skews (FP64 mode), or k0 = delta/d in the float k set with 1/d as
// k_adj_build32 computes it (mixed).  Only the touched cells are read
// later, so no field is cleared first.  An item with more receivers than
// the workgroup's entries hold is done by one thread in receiver order.
constexpr int kSrcEntries = 2048;
__global__ __launch_bounds__(256) void k_adj_source(const SlotDesc *sd, int which, int mixed) {
  __shared__ int cell[kSrcEntries];
  __shared__ double term[kSrcEntries];
  __shared__ int bad;
  const SlotDesc D = sd[blockIdx.x];
  if (*D.status < 0) return;  // an earlier stage failed
  const Grid &g = D.g;
  const int nx = g.nx, ny = g.ny, R = D.sg.R;
  const size_t SE = D.sg.skew_elems();
  const double *rhs = which ? D.rhs_den : D.rhs_adj;
  const int n = 4 * D.nr;
  GMEM double *sk0 = gm(D.skew);
  // the source of interior cell (i, j) into the item's coefficient fields
  auto put = [&](int i, int j, double dl) {
    if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2) return;
    if (!mixed) {
      for (int sk = 0; sk < 2; ++sk) sk0[(size_t)sk * kAF * SE + 5 * SE + skpos(sk, i, j, R)] = dl;
      return;
    }
    const GMEM double *Ts = sk0 + kSkTP * SE;
    const GMEM float *ps = gm(D.pskew32);  // plus skew: a, b, c element by element
    auto pa = [&](size_t q) { return ps[q * 3]; };
    auto pb = [&](size_t q) { return ps[q * 3 + 1]; };
    auto pc = [&](size_t q) { return ps[q * 3 + 2]; };
    auto at = [&](int di, int dj) { return skpos(0, i + di, j + dj, R); };
    const size_t k0 = at(0, 0), kW = at(-1, 0), kE = at(1, 0), kS = at(0, -1), kN = at(0, 1);
    double wk[5];
    adj_coef_k(Ts[k0], Ts[kW], Ts[kE], Ts[kS], Ts[kN], Ts[at(-1, -1)], Ts[at(-1, 1)],
               Ts[at(1, -1)], Ts[at(1, 1)], pa(k0), pa(kW), pa(kE), pb(k0), pb(kS), pb(kN),
               pc(k0), pc(kW), pc(kE), pc(kS), pc(kN), 1.0 / g.dx, 1.0 / g.dy, wk);
    const float kk = wk[4] != 0.0 ? (float)(dl * wk[4]) : 0.0f;
    GMEM float *k32 = reinterpret_cast<GMEM float *>(sk0 + (size_t)kSkA32 * SE);
    for (int sk = 0; sk < 2; ++sk) k32[((size_t)sk * SE + skpos(sk, i, j, R)) * 5] = kk;
  };
  // receiver q's bracket (i0, j0), its weights and its term's scale
  auto bracket = [&](int q, int &i0, int &j0, float &r1, float &r2, double &sa, double &den) {
    const double xr = deg2rad_s(D.recx[q]), yr = deg2rad_s(D.recy[q]);
    const int idx0 = (int)floor((xr - g.xx[0]) / g.dx + 1);
    const int idy0 = (int)floor((yr - g.yy[0]) / g.dy + 1);
    if (idx0 < 1 || idy0 < 1 || idx0 + 1 > nx || idy0 + 1 > ny) return false;
    i0 = idx0 - 1;
    j0 = idy0 - 1;
    r1 = weight_r4_s((xr - g.xx[i0]) / g.dx);
    r2 = weight_r4_s((yr - g.yy[j0]) / g.dy);
    sa = rhs[q];
    den = g.dx * kREarth * cos(yr) * g.dy * kREarth;
    return true;
  };
  auto corner = [&](int c, double sa, float r1, float r2, double den) {
    const float wa = c < 2 ? 1.0f - r1 : r1, wb = (c & 1) ? r2 : 1.0f - r2;
    return sa * (double)wa * (double)wb / den;
  };
  if (n > kSrcEntries) {  // one thread, receivers in order, on the natural field
    if (threadIdx.x != 0) return;
    double *delta = D.delta;
    for (int pass = 0; pass < 3; ++pass)  // clear, accumulate, scatter
      for (int q = 0; q < D.nr; ++q) {
        int i0, j0;
        float r1, r2;
        double sa, den;
        if (!bracket(q, i0, j0, r1, r2, sa, den)) {
          *D.status = kERecOutOfGrid;
          D.iters[1 + which] = 0;
          return;
        }
        for (int c = 0; c < 4; ++c) {
          const size_t k = i0 + (c >> 1) + (size_t)nx * (j0 + (c & 1));
          if (pass == 0) delta[k] = 0;
          else if (pass == 1) delta[k] = delta[k] + corner(c, sa, r1, r2, den);
          else put(i0 + (c >> 1), j0 + (c & 1), delta[k]);
        }
      }
    return;
  }
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int w = threadIdx.x; w < n; w += blockDim.x) {
    int i0, j0;
    float r1, r2;
    double sa, den;
    const int q = w >> 2, c = w & 3;
    if (!bracket(q, i0, j0, r1, r2, sa, den)) {
      bad = 1;
      cell[w] = -1;
      continue;
    }
    cell[w] = i0 + (c >> 1) + nx * (j0 + (c & 1));
    term[w] = corner(c, sa, r1, r2, den);
  }
  __syncthreads();
  if (bad) {
    if (threadIdx.x == 0) {
      *D.status = kERecOutOfGrid;
      D.iters[1 + which] = 0;
    }
    return;
  }
  for (int w = threadIdx.x; w < n; w += blockDim.x) {
    const int k = cell[w];
    bool first = true;
    for (int v = 0; v < w && first; ++v) first = cell[v] != k;
    if (!first) continue;
    double dl = 0;
    for (int v = w; v < n; ++v)
      if (cell[v] == k) dl = dl + term[v];
    put(k % nx, k / nx, dl);
  }
}

template <typename V>
struct AFar {
  V ta, tdn, snap;
  int j;
  bool act;
};
template <typename V>
struct ANear;
template <>
struct ANear<double> {
  double a1, a2, b1, b2, d, dl;
};
template <>
struct ANear<float> {
  float k0, kW, kE, kS, kN;
};


__device__ __forceinline__ void adj_load_near(const FieldSet<double> &f, size_t e,
                                              ANear<double> &q) {
  q.a1 = f.ld(0, e); q.a2 = f.ld(1, e); q.b1 = f.ld(2, e); q.b2 = f.ld(3, e);
  q.d = f.ld(4, e); q.dl = f.ld(5, e);
}
__device__ __forceinline__ void adj_load_near(const CellSet<5> &f, size_t e, ANear<float> &q) {
  const u32x4 k4 = f.ld4(e, 0);
  q.k0 = as_f(k4.x); q.kW = as_f(k4.y); q.kE = as_f(k4.z); q.kS = as_f(k4.w);
  q.kN = f.ld1(e, 4);
}

// The interior update (src/eikon_solver_2d.f90:1392-1402), and its FP32 form.
__device__ __forceinline__ double adj_update(const ANear<double> &n, double tW, double tE,
                                             double tS, double tN, double dx, double dy) {
  if (n.d == 0.0) return 0;
  const double a1p = (n.a1 + fabs(n.a1)) / 2, a2m = (n.a2 - fabs(n.a2)) / 2;
  const double b1p = (n.b1 + fabs(n.b1)) / 2, b2m = (n.b2 - fabs(n.b2)) / 2;
  const double e = (tW * a1p - tE * a2m) / dx + (tS * b1p - tN * b2m) / dy;
  return (n.dl + e) / n.d;
}
__device__ __forceinline__ float adj_update(const ANear<float> &n, float tW, float tE, float tS,
                                            float tN, double, double) {
  return fmaf(n.kN, tN, fmaf(n.kS, tS, fmaf(n.kE, tE, fmaf(n.kW, tW, n.k0))));
}
// The mixed mode's FP64 polish: the same form in double on the float
// coefficients, accumulated in the same order with a rounding per operation.
__device__ __forceinline__ double adj_update_k64(const ANear<float> &n, double tW, double tE,
                                                 double tS, double tN) {
  double t = (double)n.k0;
  t = (double)n.kW * tW + t;
  t = (double)n.kE * tE + t;
  t = (double)n.kS * tS + t;
  return (double)n.kN * tN + t;
}

template <bool K>
__device__ __forceinline__ typename std::conditional<K, CellSet<5>, FieldSet<double>>::type
make_coef_set(const GMEM void *base, size_t SE) {
  if constexpr (K)
    return CellSet<5>(base, SE);
  else
    return FieldSet<double>((const GMEM double *)base, SE, kAF);
}

template <typename V, typename S = V>
struct AdjCtx {
  int nx, ny, R, lane;
  size_t SE;
  double dx, dy;
  const GMEM S *is[2];  // item coefficient skews, stored in S
  GMEM V *snap;         // plus skew
};

// One adjoint sweep over strip c, in place in the sweep's own skew; MODE as
// for the forward, and MODE 2 also accumulates max |Ta| (acc_a) for the
// FP32 stop rule.
template <typename V, int XD, int YD, int MODE, int PD, typename CS = V>
__device__ void adj_strip_pass(const AdjCtx<V, CS> &C, const Team &T, int k, GMEM V *ta, double &ss,
                               double &sm, double &sa) {
  constexpr int SK = XD == YD ? 0 : 1;
  const int nx = C.nx, ny = C.ny, R = C.R, lane = C.lane;
  const int c = XD > 0 ? k : T.ns - 1 - k;
  const size_t SE = C.SE;
  // the coefficient set: kAF FP64 fields (a1, a2, b1, b2, d, delta), or
  // the mixed mode's 5 float ones (k0, kW, kE, kS, kN)
  constexpr bool kK = sizeof(CS) == 4;
  // the k set element by element (CellSet), the FP64 set field by field
  using CSet = typename std::conditional<kK, CellSet<5>, FieldSet<double>>::type;
  const CSet fis = make_coef_set<kK>(C.is[SK], SE);
  const FieldSet<V> fta(ta, SE, 1), fsnap(C.snap, SE, 1);
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
  // team edges as in the forward pass
  const int pw = (T.w + T.W - 1) % T.W, base_in = ((k - 1) / T.W) * ny;
  const bool pub = k < T.ns - 1 && lag == L - 1;
  const int base_out = (k / T.W) * ny;
  double acc_s = 0, acc_m = 0, acc_a = 0;

  // Every lane loads at every step, at step_pos, with no load behind a
  // branch (see the forward pass).
  auto load_far = [&](int t, AFar<V> &q) {
    const int u = t - lag;
    const int j = YD > 0 ? u : ny - 1 - u;
    q.j = j;
    q.act = lv && u >= 0 && u < ny;
    const int jc = min(max(j, 0), ny - 1);
    const size_t e = step_pos<XD, YD>(c, t, lane, ny, L, R);
    q.ta = fta.ld(0, e);
    q.tdn = fta.ld(0, opaque_index(dn_ok ? skpos(SK, id, jc, R) : e));
    if (MODE == 2) q.snap = fsnap.ld(0, e);
  };
  using NS = ANear<CS>;
  auto load_near = [&](int t, NS &q) {
    adj_load_near(fis, step_pos<XD, YD>(c, t, lane, ny, L, R), q);
  };
  // operand sets as in the forward pass: S = PD+2 sets, far and near
  // operands PD steps ahead, the loop starting PD fetch-only steps early
  static_assert(PD >= 2, "the next step's far set is read while this step fills one");
  constexpr int S = PD + 2;
  AFar<V> fs[S] = {};
  NS ns[S] = {};
  V tprev = 0;
  auto step = [&](int t, const AFar<V> &fc, const AFar<V> &fn, AFar<V> &f2, const NS &nc,
                  NS &nx2, int tnear) {
    // the upstream strip's edge for this step, read ahead of the step's
    // global loads (see the forward pass)
    V tU = 0;
    if (up_ok && fc.act) {  // k > 0
      if (T.multi()) team_wait(T, pw, base_in + t - lag + 1);
      tU = team_read<V>(T, pw, t - lag);
    }
    load_far(t + PD, f2);
    load_near(tnear, nx2);
    V tUx = lane_shift<-XD>(tprev);
    V tDx = lane_shift<XD>(fn.ta);
    if (up_ok) tUx = tU;
    if (dn_ok) tDx = fc.tdn;
    const V tW = XD > 0 ? tUx : tDx, tE = XD > 0 ? tDx : tUx;
    const V tS = YD > 0 ? tprev : fn.ta, tN = YD > 0 ? fn.ta : tprev;
    const int jc = fc.j;
    V tnew = fc.ta;
    if (fc.act) {
      if (interior_col && jc >= 1 && jc <= ny - 2) {
        if constexpr (kK && sizeof(V) == 8)
          tnew = adj_update_k64(nc, tW, tE, tS, tN);
        else
          tnew = adj_update(nc, tW, tE, tS, tN, C.dx, C.dy);
      }
      const size_t e = skidx(c, skew_row<XD, YD>(t, ny, L), lane, R);
      fta.st(0, e, tnew);
      if (pub) team_publish<V>(T, t - lag, tnew, base_out + t - lag + 1);
      if (MODE == 1) fsnap.st(0, e, fc.ta);
      if (MODE == 2) {
        const double dd = fabs((double)tnew - (double)fc.snap);
        acc_s += dd;
        acc_m = dd > acc_m ? dd : acc_m;
        if (std::is_same<V, float>::value) {
          const double av = fabs((double)tnew);
          acc_a = av > acc_a ? av : acc_a;
        }
      }
    }
    tprev = tnew;
  };
  for (int t = -PD; t < nsteps; t += S) {
#pragma unroll
    for (int q = 0; q < S; ++q)
      step(t + q, fs[q], fs[(q + 1) % S], fs[(q + PD) % S], ns[q], ns[(q + PD) % S], t + q + PD);
  }
  if (MODE == 2) {
    ss = wave_sum(acc_s);
    sm = wave_max(acc_m);
    sa = wave_max(acc_a);
  }
  wave_fence();
}

template <typename V, int XD, int YD, int MODE, int PD, typename CS = V>
__device__ void adj_sweep(const AdjCtx<V, CS> &C, Team &T, GMEM V *ta, StripAcc &acc) {
  team_sweep_begin(T);
  for (int k = T.w; k < T.ns; k += T.W) {
    double ss = 0, sm = 0, sa = 0;
    adj_strip_pass<V, XD, YD, MODE, PD, CS>(C, T, k, ta, ss, sm, sa);
    if (MODE == 2) team_strip_term(T, k, ss, sm, sa, acc);
  }
  team_sweep_end(T);
}

// One iteration: (xdirec, ydirec) = (-1,-1), (-1,1), (1,-1), (1,1), x
// ascending first (src/eikon_solver_2d.f90:1383-1390); with SNAP the
// snapshot and the residuals (:1412-1434) are taken.
template <typename V, bool SNAP, int PD, typename CS = V>
__device__ void adj_iteration(const AdjCtx<V, CS> &C, Team &T, GMEM V *tP, GMEM V *tM, double &s,
                              double &m, double &a) {
  StripAcc acc;
  adj_sweep<V, 1, 1, SNAP ? 1 : 0, PD, CS>(C, T, tP, acc);
  skew_copy<V, 0>(T, tP, tM, C.nx, C.ny, C.R, T.ns, C.lane);
  adj_sweep<V, 1, -1, 0, PD, CS>(C, T, tM, acc);
  adj_sweep<V, -1, 1, 0, PD, CS>(C, T, tM, acc);
  skew_copy<V, 1>(T, tM, tP, C.nx, C.ny, C.R, T.ns, C.lane);
  adj_sweep<V, -1, -1, SNAP ? 2 : 0, PD, CS>(C, T, tP, acc);
  const StripAcc r = team_strip_total(T, acc);
  s = r.s;
  m = r.m;
  a = r.a;
}

template <typename V, typename S>
__device__ void adj_ctx_init(AdjCtx<V, S> &C, const Grid &g, const SkewGeom &sg, int lane) {
  C.nx = g.nx; C.ny = g.ny; C.R = sg.R; C.lane = lane; C.SE = sg.skew_elems();
  C.dx = g.dx; C.dy = g.dy;
}

// Natural-layout output from the plus skew, with the mask
// (src/eikon_solver_2d.f90:1449-1467, degrees).
__device__ void adj_output(const Grid &g, const SkewGeom &sg, const double *TaP, double x0,
                           double y0, int apply_mask, double *Tout) {
  const int nx = g.nx;
  const size_t N = (size_t)nx * g.ny;
  const GMEM double *xd = gm(g.xd), *yd = gm(g.yd), *src = gm(TaP);
  GMEM double *out = gm(Tout);
  const double dxd = xd[1] - xd[0], dyd = yd[1] - yd[0];
  for (size_t k = threadIdx.x; k < N; k += blockDim.x) {
    const int i = (int)(k % nx), j = (int)(k / nx);
    double v = src[skpos(0, i, j, sg.R)];
    if (apply_mask && fabs(xd[i] - x0) < 1.5 * dxd && fabs(yd[j] - y0) < 1.5 * dyd) v = 0;
    out[k] = v;
  }
}

template <bool TEAM, int PD>
__global__ __launch_bounds__(TEAM ? 64 * kMaxTeam : 64) void k_adj_strip(const SlotDesc *sd, int which) {
  extern __shared__ double lds[];
  const SlotDesc D = sd[blockIdx.x];
  const int lane = threadIdx.x & 63;
  if (*D.status < 0) return;  // an earlier stage failed
  const Grid &g = D.g;
  const SkewGeom &sg = D.sg;
  const size_t SE = sg.skew_elems();
  Team T = team_init<TEAM>(lds, g.ny, sg.ns, 8);
  double *sk0 = D.skew;
  double *TaP = sk0 + kSkTaP * SE, *TaM = sk0 + kSkTaM * SE;
  AdjCtx<double> C;
  adj_ctx_init(C, g, sg, lane);
  C.is[0] = gm(sk0);
  C.is[1] = gm(sk0 + kAF * SE);
  C.snap = gm(sk0 + kSkSnapA * SE);
  int iter = 1;
  bool conv = false;
  for (; iter <= kAdjMaxIter; ++iter) {
    double s, m, a;
    adj_iteration<double, true, PD>(C, T, gm(TaP), gm(TaM), s, m, a);
    const double l1 = s / (double)(g.nx * g.ny);
    if (fabs(l1) < kAdjTol && m < kAdjTol) { conv = true; break; }
  }
  if (!conv) iter = kAdjMaxIter;
  adj_output(g, sg, TaP, D.x0, D.y0, 1, which ? D.Td : D.Ta);
  if (threadIdx.x == 0) {
    if (team_failed(T)) *D.status = kETeamWait;
    else if (!conv && *D.status == kOk) *D.status = kWMaxIter;
    D.iters[1 + which] = iter;
  }
}

// Mixed adjoint, FP32 phase: iterations on the float coefficients (float
// skews from double skew 16: k0..kN plus 0..4 and minus 5..9, each skew's
// five stored element by element (CellSet<5>), Ta plus 10 and minus 11,
// snapshot 12), then Ta = (double)Ta32 into the FP64 plus skew.
template <bool TEAM, int PD>
__global__ __launch_bounds__(TEAM ? 64 * kMaxTeam : 64) void k_adj_mixed32(const SlotDesc *sd, int which) {
  extern __shared__ double lds[];
  const SlotDesc D = sd[blockIdx.x];
  const int lane = threadIdx.x & 63;
  if (*D.status < 0) return;
  const Grid &g = D.g;
  const SkewGeom &sg = D.sg;
  const size_t SE = sg.skew_elems();
  Team T = team_init<TEAM>(lds, g.ny, sg.ns, 4);
  double *sk0 = D.skew;
  float *f32 = reinterpret_cast<float *>(sk0 + (size_t)kSkA32 * SE);
  float *taP = f32 + 10 * SE, *taM = f32 + 11 * SE;
  AdjCtx<float> C;
  adj_ctx_init(C, g, sg, lane);
  C.is[0] = gm(f32);
  C.is[1] = gm(f32 + 5 * SE);
  C.snap = gm(f32 + 12 * SE);
  int iter = 1;
  bool done = false;
  for (; iter <= kAdjMaxIter; ++iter) {
    double s, m, a;
    adj_iteration<float, true, PD>(C, T, gm(taP), gm(taM), s, m, a);
    const double l1 = s / (double)(g.nx * g.ny);
    if ((fabs(l1) < kAdjTol && m < kAdjTol) || m <= (double)kMixRelStop * a) {
      done = true;
      break;
    }
  }
  if (!done) iter = kAdjMaxIter;
  double *TaP = sk0 + kSkTaP * SE;
  for (size_t e = threadIdx.x; e < SE; e += blockDim.x) gm(TaP)[e] = (double)gm(taP)[e];
  if (threadIdx.x == 0) {
    if (team_failed(T)) *D.status = kETeamWait;
    else if (!done && *D.status == kOk) *D.status = kWMaxIter;
    D.iters[4 + which] = iter;
  }
}

// Mixed adjoint, FP64 polish: one iteration of the update in double on the
// FP32 phase's float coefficients, and the masked output.
template <bool TEAM, int PD>
__global__ __launch_bounds__(TEAM ? 64 * kMaxTeam : 64) void k_adj_polish(const SlotDesc *sd, int which) {
  extern __shared__ double lds[];
  const SlotDesc D = sd[blockIdx.x];
  const int lane = threadIdx.x & 63;
  if (*D.status < 0) return;
  const Grid &g = D.g;
  const SkewGeom &sg = D.sg;
  const size_t SE = sg.skew_elems();
  Team T = team_init<TEAM>(lds, g.ny, sg.ns, 8);
  double *sk0 = D.skew;
  double *TaP = sk0 + kSkTaP * SE, *TaM = sk0 + kSkTaM * SE;
  // the float k set of the FP32 phase (k_adj_build32, k_adj_source)
  const float *f32 = reinterpret_cast<const float *>(sk0 + (size_t)kSkA32 * SE);
  AdjCtx<double, float> C;
  adj_ctx_init(C, g, sg, lane);
  C.is[0] = gm(f32);
  C.is[1] = gm(f32 + 5 * SE);
  C.snap = nullptr;
  double s, m, a;
  adj_iteration<double, false, PD, float>(C, T, gm(TaP), gm(TaM), s, m, a);
  adj_output(g, sg, TaP, D.x0, D.y0, 1, which ? D.Td : D.Ta);
  if (threadIdx.x == 0) {
    if (team_failed(T)) *D.status = kETeamWait;
    D.iters[1 + which] = 1;
  }
}

// The item's own kernel, adj/svel**2 accumulated from zero
// (src/tomo2d.f90:155), one thread per cell.
__global__ __launch_bounds__(256) void k_item_kernel(const SlotDesc *sd) {
  const SlotDesc D = sd[blockIdx.y];
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k >= (size_t)D.g.nx * D.g.ny || !D.kern) return;
  const double v = D.vel[k];
  D.kern[k] = 0.0 + D.Ta[k] / (v * v);
}

}  // namespace

void launch_item_kernel(hipStream_t s, const SlotDesc *sd, const GroupDims &gd) {
  hipLaunchKernelGGL(k_item_kernel, dim3((unsigned)((gd.max_n + 255) / 256), gd.count), dim3(256),
                     0, s, sd);
}

void launch_period_skew(hipStream_t s, const Grid &g, const SkewGeom &sg, int np,
                        const PeriodFields &pf, double *out, float *out32) {
  const size_t SE = sg.skew_elems();
  hipLaunchKernelGGL(k_period_skew, dim3((unsigned)((SE + 255) / 256), 2, np), dim3(256), 0, s,
                     g, sg, np, pf.a, pf.b, pf.c, pf.fun, out, out32);
}

void launch_grid_trig(hipStream_t s, const Grid &g, double *sy, double *cy) {
  hipLaunchKernelGGL(k_grid_trig, dim3((unsigned)((g.ny + 63) / 64)), dim3(64), 0, s, g, sy, cy);
}

namespace {
}  // namespace

void launch_fwd_setup(hipStream_t s, const SlotDesc *sd, const GroupDims &gd, bool mixed) {
  hipLaunchKernelGGL(k_fwd_prep, dim3((unsigned)((gd.max_nxy + 63) / 64), gd.count), dim3(64), 0,
                     s, sd);
  if (mixed)
    hipLaunchKernelGGL(k_fwd_setup<true>, tile_grid(gd), dim3(256), 0, s, sd, gd.count,
                       tiles_per_skew(gd));
  else
    hipLaunchKernelGGL(k_fwd_setup<false>, tile_grid(gd), dim3(256), 0, s, sd, gd.count,
                       tiles_per_skew(gd));
}

namespace {
// Block and LDS of a strip-kernel launch of the group (vbytes: the channel's
// value size).
dim3 team_block(const GroupDims &gd, int st) {
  return dim3(64 * (unsigned)std::max(1, gd.waves[st]));
}
size_t team_lds(const GroupDims &gd, int st, int vbytes) {
  return team_lds_bytes(gd.waves[st], gd.max_ny, gd.max_ns, vbytes);
}
// The kernel form of a stage's launch: team or one wave, and the prefetch
// depth (kDeepPrefetch or the default two steps).
template <typename K>
K pick_form(const GroupDims &gd, int st, K team2, K team_deep, K one2, K one_deep) {
  const bool deep = gd.depth[st] >= kDeepPrefetch;
  return gd.waves[st] > 1 ? (deep ? team_deep : team2) : (deep ? one_deep : one2);
}
dim3 stage_block(const GroupDims &gd, int st) {
  return gd.waves[st] > 1 ? team_block(gd, st) : dim3(64);
}
}  // namespace

int team_waves(int want, int ny, int ns, int stage) {
  const int vbytes = stage == kStageFwd32 || stage == kStageAdj32 ? 4 : 8;
  int w = std::max(1, std::min({want, kMaxTeam, ns}));
  while (w > 1 && team_lds_bytes(w, ny, ns, vbytes) > kTeamLdsMax) --w;
  return w;
}

void launch_fwd_strip(hipStream_t s, const SlotDesc *sd, const GroupDims &gd) {
  const int st = kStageFwd64;
  // the full FP64 cell reads more operands (both x neighbours and their
  // T0): a third operand set would already spill its registers, so it has
  // the two-step form only
  auto k = gd.waves[st] > 1 ? k_fwd_strip<true, 2> : k_fwd_strip<false, 2>;
  hipLaunchKernelGGL(k, dim3(gd.count), stage_block(gd, st), team_lds(gd, st, 8), s, sd);
}

void launch_fwd_mixed32(hipStream_t s, const SlotDesc *sd, const GroupDims &gd) {
  const int st = kStageFwd32;
  auto k = pick_form(gd, st, k_fwd_mixed32<true, 2>, k_fwd_mixed32<true, kDeepPrefetch>,
                     k_fwd_mixed32<false, 2>, k_fwd_mixed32<false, kDeepPrefetch>);
  hipLaunchKernelGGL(k, dim3(gd.count), stage_block(gd, st), team_lds(gd, st, 4), s, sd);
}

void launch_fwd_polish(hipStream_t s, const SlotDesc *sd, const GroupDims &gd) {
  const int st = kStageFwd64;
  auto k = pick_form(gd, st, k_fwd_polish<true, 2>, k_fwd_polish<true, kDeepPrefetch>,
                     k_fwd_polish<false, 2>, k_fwd_polish<false, kDeepPrefetch>);
  hipLaunchKernelGGL(k, dim3(gd.count), stage_block(gd, st), team_lds(gd, st, 8), s, sd);
}

void launch_adj_setup(hipStream_t s, const SlotDesc *sd, const GroupDims &gd, int which,
                      bool mixed) {
  if (mixed)
    hipLaunchKernelGGL(k_adj_build32, tile_grid(gd), dim3(256), 0, s, sd, gd.count,
                       tiles_per_skew(gd));
  else
    hipLaunchKernelGGL(k_adj_build, tile_grid(gd), dim3(256), 0, s, sd, gd.count,
                       tiles_per_skew(gd));
  hipLaunchKernelGGL(k_adj_source, dim3(gd.count), dim3(256), 0, s, sd, which, (int)mixed);
}

void launch_adj_strip(hipStream_t s, const SlotDesc *sd, const GroupDims &gd, int which) {
  const int st = kStageAdj64;
  auto k = pick_form(gd, st, k_adj_strip<true, 2>, k_adj_strip<true, kDeepPrefetch>,
                     k_adj_strip<false, 2>, k_adj_strip<false, kDeepPrefetch>);
  hipLaunchKernelGGL(k, dim3(gd.count), stage_block(gd, st), team_lds(gd, st, 8), s, sd, which);
}

void launch_adj_mixed32(hipStream_t s, const SlotDesc *sd, const GroupDims &gd, int which) {
  const int st = kStageAdj32;
  auto k = pick_form(gd, st, k_adj_mixed32<true, 2>, k_adj_mixed32<true, kDeepPrefetch>,
                     k_adj_mixed32<false, 2>, k_adj_mixed32<false, kDeepPrefetch>);
  hipLaunchKernelGGL(k, dim3(gd.count), stage_block(gd, st), team_lds(gd, st, 4), s, sd, which);
}

void launch_adj_polish(hipStream_t s, const SlotDesc *sd, const GroupDims &gd, int which) {
  const int st = kStageAdj64;
  auto k = pick_form(gd, st, k_adj_polish<true, 2>, k_adj_polish<true, kDeepPrefetch>,
                     k_adj_polish<false, 2>, k_adj_polish<false, kDeepPrefetch>);
  hipLaunchKernelGGL(k, dim3(gd.count), stage_block(gd, st), team_lds(gd, st, 8), s, sd, which);
}

}  // namespace satt
