// The 3-D kernel steps downstream of the stacked 2-D kernels (the att_acqui
// path): the combination of the period kernels with the depth sensitivities,
// and the projection of a model-grid kernel onto the inversion grids and
// back.
//
//   combine_kernels              src/acqui.f90:197-245
//   dalpha_dbeta, drho_dalpha    src/shared/utils.f90:99-145
//   inv_grid_iso, inv2fwd_iso    src/eikon_solver_2d.f90:1538-1691
//                                (locate_within_cube :1470-1535)
//   post_processing_for_kernel   src/acqui.f90:247-304
//   correct_depth (interp1_0)    src/surfker/surfker.f90:260-283
//                                (src/shared/utils.f90:786-805)
//
// Locating a point.  The reference walks from the cube the previous point
// of its (j, i, k) loop ended in, so a point on a face of the inversion grid
// lands in the cube the walk reaches first, and for inv2fwd_iso that choice
// decides the order of the eight weighted terms.  The walk's choice has a
// closed form per axis: with the axis's points visited in ascending order,
// a point reached from below (or the loop's first point, the walk starting
// at cube 1) lands in the lowest cube whose upper node is >= it; a point
// reached from above (the first point of an inner loop that restarts: x at
// the start of each row after the first, z at the start of each column
// after the first) lands in the highest cube whose lower node is <= it.
// Every point is thus located on its own, one thread per point.
//
// inv_grid_iso scatters every interior point into the eight corners of its
// cube, in loop order.  Here a first pass locates every (set, point) in
// parallel and stores its cube and eight terms; then one wave per inversion
// node gathers the terms of the points of the cubes around it, in the same
// (j, i, k) order, so each node's sum takes the reference's terms in the
// reference's order.  Cube indices are packed 10 bits per axis
// (inversion grids of at most 1024 nodes per axis).
#include <hip/hip_runtime.h>

#include "satt_internal.h"

namespace satt {
namespace {

// src/shared/shared_par.f90:16,23
constexpr double kRhoScaling = 0.33;
constexpr double kPrecondThres = 1.0e-4;

// utils.f90:99-145: REAL(4) literals (2*0.4721 is a default-real product),
// widened against the REAL(8) argument; integer powers as gfortran expands
// them (x**3 = x*x*x, x**4 = (x*x)*(x*x)).
__device__ double drho_dalpha(double vp) {
  constexpr double c0 = (double)1.6612f, c1 = (double)(2.0f * 0.4721f);
  constexpr double c2 = (double)(3.0f * 0.0671f), c3 = (double)(4.0f * 0.0043f);
  constexpr double c4 = (double)(5.0f * 0.000106f);
  const double v2 = vp * vp, v3 = v2 * vp, v4 = v2 * v2;
  return c0 - c1 * vp + c2 * v2 - c3 * v3 + c4 * v4;
}

__device__ double dalpha_dbeta(double vs) {
  constexpr double c0 = (double)2.0947f, c1 = (double)(2.0f * 0.8206f);
  constexpr double c2 = (double)(3.0f * 0.2683f), c3 = (double)(4.0f * 0.0251f);
  const double v2 = vs * vs, v3 = v2 * vs;
  return c0 - c1 * vs + c2 * v2 - c3 * v3;
}

// One thread per model cell c = i + nx*(j + ny*k); the periods in order, the
// statements of each period in the reference's order.
__global__ void k_combine(Combine C) {
  const size_t c3 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c3 >= C.n3) return;
  const size_t c2 = c3 % C.n2;
  const int np = C.nperiod;
  const double *a = C.adj_s + (size_t)np * c2, *ad = C.adj_den + (size_t)np * c2;
  const size_t s0 = (size_t)np * c3;
  if (C.abr) {
    double *k = C.ker + 3 * c3, *kd = C.ker_den + 3 * c3;
    double k0 = k[0], k1 = k[1], k2 = k[2], d0 = kd[0], d1 = kd[1], d2 = kd[2];
    for (int p = 0; p < np; ++p) {
      const double svs = C.sen_vs[s0 + p], svp = C.sen_vp[s0 + p];
      k0 = k0 - a[p] * svs;
      d0 = d0 - ad[p] * svs;
      k1 = k1 - a[p] * svp;
      d1 = d1 - ad[p] * svp;
      if (!C.rho_scaling) {
        const double sr = C.sen_rho[s0 + p];
        k2 = k2 - a[p] * sr;
        d2 = d2 - ad[p] * sr;
      }
    }
    if (C.rho_scaling) {
      k2 = kRhoScaling * k0;
      d2 = d0;
    }
    k[0] = k0, k[1] = k1, k[2] = k2;
    kd[0] = d0, kd[1] = d1, kd[2] = d2;
  } else {
    const double dab = dalpha_dbeta(C.vs3d[c3]), dra = drho_dalpha(C.vp3d[c3]);
    double k0 = C.ker[c3], d0 = C.ker_den[c3];
    for (int p = 0; p < np; ++p) {
      const double svs = C.sen_vs[s0 + p], svp = C.sen_vp[s0 + p], sr = C.sen_rho[s0 + p];
      k0 = k0 - a[p] * svs - a[p] * svp * dab - a[p] * sr * dra * dab;
      d0 = d0 - ad[p] * svs - ad[p] * svp * dab - ad[p] * sr * dra * dab;
    }
    C.ker[c3] = k0;
    C.ker_den[c3] = d0;
  }
}

// lowest cube c in [1, n-1] (1-based; nodes v[c-1], v[c]) with x <= v[c]
__device__ int cube_from_below(double x, const double *v, int n) {
  int lo = 1, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (x <= v[mid]) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// highest cube c in [1, n-1] with v[c-1] <= x
__device__ int cube_from_above(double x, const double *v, int n) {
  int lo = 1, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (v[mid - 1] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// locate_within_cube's result for model point (i, j, k) of a loop whose
// first indices are (i0, j0, k0) and counts (ni, nj, nk); the point is
// inside the set's grid (checked by the host).  c: 1-based cube; wt: the
// reference's eight weights in its order.
__device__ void locate(const InvGrid &G, int s, int i, int j, int k, int i0, int j0, int k0,
                       int ni, int nk, int c[3], double wt[8]) {
  const double *X = G.xinv + (size_t)G.nxi * s, *Y = G.yinv + (size_t)G.nyi * s;
  const double *Z = G.zinv + (size_t)G.nzi * s;
  const double x = G.vx[i], y = G.vy[j], z = G.vz[k];
  const bool x_above = i == i0 && j != j0 && ni > 1;
  const bool z_above = k == k0 && (i != i0 || j != j0) && nk > 1;
  c[0] = x_above ? cube_from_above(x, X, G.nxi) : cube_from_below(x, X, G.nxi);
  c[1] = cube_from_below(y, Y, G.nyi);
  c[2] = z_above ? cube_from_above(z, Z, G.nzi) : cube_from_below(z, Z, G.nzi);
  const double wx = (x - X[c[0] - 1]) / (X[c[0]] - X[c[0] - 1]);
  const double wy = (y - Y[c[1] - 1]) / (Y[c[1]] - Y[c[1] - 1]);
  const double wz = (z - Z[c[2] - 1]) / (Z[c[2]] - Z[c[2] - 1]);
  wt[0] = (1.0 - wx) * (1.0 - wy) * (1.0 - wz);
  wt[1] = (1.0 - wx) * wy * (1.0 - wz);
  wt[2] = wx * wy * (1.0 - wz);
  wt[3] = wx * (1.0 - wy) * (1.0 - wz);
  wt[4] = (1.0 - wx) * (1.0 - wy) * wz;
  wt[5] = (1.0 - wx) * wy * wz;
  wt[6] = wx * wy * wz;
  wt[7] = wx * (1.0 - wy) * wz;
}

// corner offsets of weight n (eikon_solver_2d.f90:1585-1602)
__constant__ int kCx[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__constant__ int kCy[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ int kCz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

// inv2fwd_iso: one thread per model point, every point of the model
__global__ void k_inv2fwd(InvGrid G, const double *gk, double *update) {
  const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t n2 = (size_t)G.nx * G.ny;
  if (c >= n2 * G.nz) return;
  const int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / n2);
  const size_t nb = (size_t)G.nxi * G.nyi * G.nzi;
  double u = 0;
  for (int s = 0; s < G.nset; ++s) {
    int cb[3];
    double wt[8];
    locate(G, s, i, j, k, 0, 0, 0, G.nx, G.nz, cb, wt);
    const double *g = gk + nb * s;
    double val = 0.0;
    for (int n = 0; n < 8; ++n) {
      const size_t m = (size_t)G.nxi * G.nyi * (cb[2] - 1 + kCz[n]) +
                       (size_t)G.nxi * (cb[1] - 1 + kCy[n]) + (cb[0] - 1 + kCx[n]);
      val = val + wt[n] * g[m];
    }
    u = u + val;
  }
  update[c] = u;
}

// first index in [lo, hi] with v[idx] >= x (hi + 1 if none)
__device__ int first_at_least(const double *v, int lo, int hi, double x) {
  int a = lo, b = hi + 1;
  while (a < b) {
    const int mid = (a + b) >> 1;
    if (v[mid] >= x) b = mid;
    else a = mid + 1;
  }
  return a;
}

// last index in [lo, hi] with v[idx] <= x (lo - 1 if none)
__device__ int last_at_most(const double *v, int lo, int hi, double x) {
  int a = lo - 1, b = hi;
  while (a < b) {
    const int mid = (a + b + 1) >> 1;
    if (v[mid] <= x) a = mid;
    else b = mid - 1;
  }
  return a;
}

// inv_grid_iso, first pass: one thread per (set, interior model point)
// locates the point and stores its cube (packed) and its eight terms
// adj*dx*dy*dz*wt(n) (eikon_solver_2d.f90:1603).
__global__ void k_inv_terms(InvGrid G, const double *adj, int *cube, double *terms) {
  const int mx = G.nx - 2, my = G.ny - 2, mz = G.nz - 2;
  const size_t ni = (size_t)mx * my * mz;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= ni * G.nset) return;
  const int s = (int)(t / ni);
  const size_t q = t % ni;
  const int i = 1 + (int)(q % mx), j = 1 + (int)((q / mx) % my), k = 1 + (int)(q / ((size_t)mx * my));
  int cb[3];
  double wt[8];
  locate(G, s, i, j, k, 1, 1, 1, mx, mz, cb, wt);
  const double dx = G.vx[1] - G.vx[0], dy = G.vy[1] - G.vy[0], dz = G.vz[1] - G.vz[0];
  const double a = adj[i + (size_t)G.nx * j + (size_t)G.nx * G.ny * k];
  cube[t] = cb[0] | (cb[1] << 10) | (cb[2] << 20);
  double *o = terms + 8 * t;
  for (int n = 0; n < 8; ++n) o[n] = a * dx * dy * dz * wt[n];
}

// Second pass: one wave per (set, inversion node) sums the terms of the
// interior points whose cube has the node as a corner, in the reference's
// (j, i, k) loop order: the lanes fetch 64 consecutive candidates of that
// order at a time, and the chunk is added in lane order.  A candidate whose
// cube misses the node adds +0.0, which leaves the sum as it is (the sum
// starts at +0.0 and can only return to zero as +0.0).
__device__ double lane_value(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

__global__ __launch_bounds__(64) void k_inv_gather(InvGrid G, const int *cube,
                                                   const double *terms, double *gk) {
  const size_t t = blockIdx.x;
  const size_t nb = (size_t)G.nxi * G.nyi * G.nzi;
  const int s = (int)(t / nb);
  const size_t m = t % nb;
  const int a = (int)(m % G.nxi), b = (int)((m / G.nxi) % G.nyi);
  const int e = (int)(m / ((size_t)G.nxi * G.nyi));
  const double *X = G.xinv + (size_t)G.nxi * s, *Y = G.yinv + (size_t)G.nyi * s;
  const double *Z = G.zinv + (size_t)G.nzi * s;
  // model indices with coordinates in [node - 1, node + 1], interior only
  const int ilo = first_at_least(G.vx, 1, G.nx - 2, X[a > 0 ? a - 1 : 0]);
  const int ihi = last_at_most(G.vx, 1, G.nx - 2, X[a < G.nxi - 1 ? a + 1 : a]);
  const int jlo = first_at_least(G.vy, 1, G.ny - 2, Y[b > 0 ? b - 1 : 0]);
  const int jhi = last_at_most(G.vy, 1, G.ny - 2, Y[b < G.nyi - 1 ? b + 1 : b]);
  const int klo = first_at_least(G.vz, 1, G.nz - 2, Z[e > 0 ? e - 1 : 0]);
  const int khi = last_at_most(G.vz, 1, G.nz - 2, Z[e < G.nzi - 1 ? e + 1 : e]);
  const int ni = ihi - ilo + 1, nj = jhi - jlo + 1, nk = khi - klo + 1;
  const int total = ni > 0 && nj > 0 && nk > 0 ? ni * nj * nk : 0;
  const int mx = G.nx - 2, my = G.ny - 2;
  const size_t base = (size_t)s * mx * my * (G.nz - 2);
  const int lane = (int)threadIdx.x;
  double sum = 0;
  for (int p0 = 0; p0 < total; p0 += 64) {
    const int p = p0 + lane;
    double v = 0.0;
    if (p < total) {
      const int k = klo + p % nk, i = ilo + (p / nk) % ni, j = jlo + p / (nk * ni);
      const size_t q = base + (size_t)(i - 1) + (size_t)mx * ((j - 1) + (size_t)my * (k - 1));
      const int c = cube[q];
      const int ox = a - ((c & 1023) - 1), oy = b - (((c >> 10) & 1023) - 1);
      const int oz = e - ((c >> 20) - 1);
      if (ox >= 0 && ox <= 1 && oy >= 0 && oy <= 1 && oz >= 0 && oz <= 1) {
        // weight index of corner (ox, oy, oz): 1..8 in the reference's order
        const int n = oz * 4 + (ox ? (oy ? 2 : 3) : (oy ? 1 : 0));
        v = terms[8 * q + n];
      }
    }
    const int cnt = total - p0 < 64 ? total - p0 : 64;
    for (int l = 0; l < cnt; ++l) sum = sum + lane_value(v, l);
  }
  if (lane == 0) gk[t] = sum;
}

// ker/ker_den component q into contiguous buffers
__global__ void k_component(size_t n, int nker, int q, const double *ker, const double *kd,
                            double *kc, double *kdc) {
  const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c >= n) return;
  kc[c] = ker[q + (size_t)nker * c];
  if (kdc) kdc[c] = kd[q + (size_t)nker * c];
}

// regularize_ker_density (acqui.f90:287-304) applied to component q of ker
// (in place) and to its contiguous copy
__global__ void k_precond(size_t n, int nker, int q, const double *kdc, const double *mx,
                          double coe, double *ker, double *kc) {
  const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c >= n) return;
  double p = fabs(kdc[c]) / *mx;
  p = p < kPrecondThres ? 0.0 : 1 / pow(p, coe);
  const double v = kc[c] * p;
  kc[c] = v;
  ker[q + (size_t)nker * c] = v;
}

// gradient_s(q) += update/nset/chi0*weight (acqui.f90:280-281)
__global__ void k_grad_update(size_t n, int nker, int q, const double *update, int nset,
                              double chi0, double weight, double *grad) {
  const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double u = update[c] / (double)nset / chi0;
  grad[q + (size_t)nker * c] = grad[q + (size_t)nker * c] + u * weight;
}

// correct_depth: one thread per (period, column); each kernel's depth
// profile is resampled at depz/cosd(dip) by interp1_0: the last node <= z
// and the first node > z, past the last node the last value.
__global__ void k_correct_depth(size_t n3, int nz, const double *in, double *out,
                                const double *dip, const double *depz) {
  const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c >= n3) return;
  const int q = blockIdx.y;  // vs, vp, rho
  const double *v = in + (size_t)q * n3 * nz;
  double *o = out + (size_t)q * n3 * nz;
  const double cd = cos(dip[c] * (3.14159265358979323846 / 180.0));
  int x1 = 0;
  for (int k = 0; k < nz; ++k) {
    const double z = depz[k] / cd;
    while (x1 + 1 < nz && depz[x1 + 1] <= z) ++x1;  // newz ascends with k
    double r;
    if (x1 + 1 >= nz || z < depz[0]) {
      r = v[c + n3 * (nz - 1)];
    } else {
      const int x2 = x1 + 1;
      const double vn = fabs(depz[x2] - depz[x1]);
      r = (v[c + n3 * x1] * (depz[x2] - z) + v[c + n3 * x2] * (z - depz[x1])) / vn;
    }
    o[c + n3 * k] = r;
  }
}

__global__ void k_fill(size_t n, double *v) {
  const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c < n) v[c] = 0.0;
}

dim3 blocks(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

void launch_combine(hipStream_t s, const Combine &C) {
  hipLaunchKernelGGL(k_combine, blocks(C.n3), dim3(256), 0, s, C);
}

size_t inv_grid_scratch_bytes(const InvGrid &G) {
  if (G.nx < 3 || G.ny < 3 || G.nz < 3) return 0;
  const size_t n = (size_t)(G.nx - 2) * (G.ny - 2) * (G.nz - 2) * G.nset;
  return n * 8 * sizeof(double) + ((n * sizeof(int) + 255) & ~size_t(255));
}

void launch_inv_grid(hipStream_t s, const InvGrid &G, const double *adj, double *gk,
                     void *scratch) {
  const size_t nb = (size_t)G.nxi * G.nyi * G.nzi * G.nset;
  if (G.nx < 3 || G.ny < 3 || G.nz < 3) {  // no interior point: gk = 0
    hipLaunchKernelGGL(k_fill, blocks(nb), dim3(256), 0, s, nb, gk);
    return;
  }
  const size_t n = (size_t)(G.nx - 2) * (G.ny - 2) * (G.nz - 2) * G.nset;
  int *cube = static_cast<int *>(scratch);
  double *terms = reinterpret_cast<double *>(static_cast<char *>(scratch) +
                                             ((n * sizeof(int) + 255) & ~size_t(255)));
  hipLaunchKernelGGL(k_inv_terms, blocks(n), dim3(256), 0, s, G, adj, cube, terms);
  hipLaunchKernelGGL(k_inv_gather, dim3((unsigned)nb), dim3(64), 0, s, G, cube, terms, gk);
}

void launch_inv2fwd(hipStream_t s, const InvGrid &G, const double *gk, double *update) {
  hipLaunchKernelGGL(k_inv2fwd, blocks((size_t)G.nx * G.ny * G.nz), dim3(256), 0, s, G, gk,
                     update);
}

void launch_component(hipStream_t s, size_t n, int nker, int q, const double *ker,
                      const double *kd, double *kc, double *kdc) {
  hipLaunchKernelGGL(k_component, blocks(n), dim3(256), 0, s, n, nker, q, ker, kd, kc, kdc);
}

void launch_precond(hipStream_t s, size_t n, int nker, int q, const double *kdc,
                    const double *mx, double coe, double *ker, double *kc) {
  hipLaunchKernelGGL(k_precond, blocks(n), dim3(256), 0, s, n, nker, q, kdc, mx, coe, ker, kc);
}

void launch_correct_depth(hipStream_t s, size_t n3, int nz, const double *in, double *out,
                          const double *dip, const double *depz) {
  hipLaunchKernelGGL(k_correct_depth, dim3((unsigned)((n3 + 255) / 256), 3), dim3(256), 0, s, n3,
                     nz, in, out, dip, depz);
}

void launch_grad_update(hipStream_t s, size_t n, int nker, int q, const double *update,
                        int nset, double chi0, double weight, double *grad) {
  hipLaunchKernelGGL(k_grad_update, blocks(n), dim3(256), 0, s, n, nker, q, update, nset, chi0,
                     weight, grad);
}

}  // namespace satt
