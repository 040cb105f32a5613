// Geographic smoothing, gradient, topographic metric and the steepest-
// descent update: the model-side steps on either side of the solver path.
//
//   gaussian_smooth_geo_2  src/shared/utils.f90:1194-1223
//   gradient_2_geo         src/shared/utils.f90:1118-1160
//   get_topo metric        src/grid.f90:114-123,137-141; dip angle src/topo.f90:123-130
//   optimize + steepest_descent  src/tomo2d.f90:203-251
//
// Smoothing.  Output cell (i,j) is sum(w*data)/sum(w) over the window
// |i'-i| <= hx, |j'-j| <= hy(j), with w = exp(-delta^2/(2 sigma^2)) and
// delta the haversine distance in degrees.  On the reference's grids (uniform
// axes from arange) a weight depends on the row pair (j, j') and the column
// offset |i'-i| only, up to the rounding of lon(i')-lon(i): one workgroup
// owns an output row, evaluates the (hx+1) weights of each window row j'
// once into LDS, stages data row j' in LDS, and every thread accumulates its
// cells in the reference's order (j' outer, i' inner, ascending).  The
// transcendental work is thus per (j, j', offset), not per (cell, node); the
// per-node work is one LDS read, one multiply and two adds in FP64.  The
// latitude terms sin((lat-lat')/2)^2, cos(lat), cos(lat') are the
// reference's per pair; the longitude term is sin(dlon_k/2)^2 with dlon_k =
// lon(k)-lon(0).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "satt_internal.h"

namespace satt {
namespace {

// src/shared/shared_par.f90:11-15: pi is a default-real literal
constexpr double kGeoPi = (double)3.14159265359f;
constexpr double kGeoRadius = 6371.0;
constexpr double kGeoDeg2Rad = kGeoPi / 180.0;
constexpr double kGeoKm2Deg = 1.0 / (6371.0 * kGeoPi / 180.0);

// s2[k] = sin(dlon_k * 0.5)**2, dlon_k = (lon(k) - lon(0)) * deg2rad
__global__ void k_geo_s2(const double *lon, int nx, double *s2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nx) return;
  const double s = sin((lon[k] - lon[0]) * kGeoDeg2Rad * 0.5);
  s2[k] = s * s;
}

// One workgroup per (output row j, field f).  LDS: the row pair's weights
// w[0..hxe] and data row j' (nx each).  CPT cells per thread: i = tid + q*bs,
// MAXT threads at most (512 when a thread has several cells, which leaves it
// the registers of its accumulators).
template <int CPT, int MAXT>
__global__ __launch_bounds__(MAXT) void k_smooth_geo2(GeoSmooth P) {
  extern __shared__ double lds[];
  const int j = blockIdx.x, f = blockIdx.y;
  const int nx = P.nx, ny = P.ny, bs = blockDim.x, tid = threadIdx.x;
  const int hx = P.hx[f], hy = P.hy[(size_t)f * ny + j];
  const int hxe = hx < nx - 1 ? hx : nx - 1;
  double *wl = lds, *dl = lds + nx;
  const double *data = P.data + (P.ndata == 1 ? 0 : (size_t)f * nx * ny);
  const double sig = P.sigma[f], two_sig_sq = 2.0 * (sig * sig);
  const int n1 = j - hy > 0 ? j - hy : 0, n2 = j + hy < ny - 1 ? j + hy : ny - 1;
  const double c0 = cos(P.lat[j] * kGeoDeg2Rad);
  double swd[CPT], sw[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) swd[q] = sw[q] = 0.0;
  for (int jj = n1; jj <= n2; ++jj) {
    const double c1 = cos(P.lat[jj] * kGeoDeg2Rad);
    const double sl = sin((P.lat[j] - P.lat[jj]) * kGeoDeg2Rad * 0.5);
    const double sdl = sl * sl;
    for (int k = tid; k <= hxe; k += bs) {
      const double a = sdl + P.s2[k] * c0 * c1;
      const double deg = 2.0 * atan2(sqrt(a), sqrt(1.0 - a));
      const double delta = kGeoRadius * deg * kGeoKm2Deg;
      wl[k] = exp(-(delta * delta) / two_sig_sq);
    }
    const double *row = data + (size_t)jj * nx;
    for (int i = tid; i < nx; i += bs) dl[i] = row[i];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int i = tid + q * bs;
      if (i < nx) {
        const int k0 = -hx > -i ? -hx : -i, k1 = hx < nx - 1 - i ? hx : nx - 1 - i;
        const double *dp = dl + i;
        double a = swd[q], b = sw[q];
        // offsets in ascending order, split at 0 so that both the weight and
        // the data index run linearly; eight nodes' loads are issued before
        // their sums, so one LDS round trip serves eight nodes
        int k = k0;
        for (; k + 8 <= 0; k += 8) {
          double w[8], d[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            w[u] = wl[-(k + u)];
            d[u] = dp[k + u];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            a += w[u] * d[u];
            b += w[u];
          }
        }
        for (; k < 0; ++k) {
          const double w = wl[-k];
          a += w * dp[k];
          b += w;
        }
        for (; k + 8 <= k1 + 1; k += 8) {
          double w[8], d[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            w[u] = wl[k + u];
            d[u] = dp[k + u];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            a += w[u] * d[u];
            b += w[u];
          }
        }
        for (; k <= k1; ++k) {
          const double w = wl[k];
          a += w * dp[k];
          b += w;
        }
        swd[q] = a;
        sw[q] = b;
      }
    }
    __syncthreads();
  }
  double *out = P.out + (size_t)f * nx * ny + (size_t)j * nx;
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int i = tid + q * bs;
    if (i < nx) out[i] = (0.0 + swd[q]) / sw[q];
  }
}

// Two output rows (j, j+1) per workgroup, one cell per thread (nx <= 1024):
// each staged data row serves both rows, so a node costs three LDS reads for
// two cells (the two rows' weights and the shared datum) instead of four, and
// a window row takes one barrier for two output rows.  Each cell's sums keep
// the single-row kernel's order (window rows ascending, offsets ascending).
__global__ __launch_bounds__(1024) void k_smooth_geo2_pair(GeoSmooth P) {
  extern __shared__ double lds[];
  const int j0 = 2 * blockIdx.x, f = blockIdx.y;
  const int nx = P.nx, ny = P.ny, tid = threadIdx.x;
  const bool two = j0 + 1 < ny;
  const int hx = P.hx[f];
  const int hy0 = P.hy[(size_t)f * ny + j0], hy1 = two ? P.hy[(size_t)f * ny + j0 + 1] : 0;
  const int hxe = hx < nx - 1 ? hx : nx - 1;
  double *w0 = lds, *w1 = lds + nx, *dl = lds + 2 * nx;
  const double *data = P.data + (P.ndata == 1 ? 0 : (size_t)f * nx * ny);
  const double sig = P.sigma[f], two_sig_sq = 2.0 * (sig * sig);
  const int a1 = j0 - hy0 > 0 ? j0 - hy0 : 0, a2 = j0 + hy0 < ny - 1 ? j0 + hy0 : ny - 1;
  const int b1 = two ? (j0 + 1 - hy1 > 0 ? j0 + 1 - hy1 : 0) : a1;
  const int b2 = two ? (j0 + 1 + hy1 < ny - 1 ? j0 + 1 + hy1 : ny - 1) : a2;
  const int n1 = a1 < b1 ? a1 : b1, n2 = a2 > b2 ? a2 : b2;
  const double c00 = cos(P.lat[j0] * kGeoDeg2Rad);
  const double c01 = two ? cos(P.lat[j0 + 1] * kGeoDeg2Rad) : c00;
  const int i = tid;
  const int k0 = -hx > -i ? -hx : -i, k1 = hx < nx - 1 - i ? hx : nx - 1 - i;
  double s0 = 0.0, t0 = 0.0, s1 = 0.0, t1 = 0.0;
  for (int jj = n1; jj <= n2; ++jj) {
    const bool in0 = jj >= a1 && jj <= a2, in1 = two && jj >= b1 && jj <= b2;
    const double c1 = cos(P.lat[jj] * kGeoDeg2Rad);
    // the weights of both rows: lanes 0..hxe for row j0, the next hxe+1 for j0+1
    for (int q = tid; q <= 2 * hxe + 1; q += blockDim.x) {
      const int r = q > hxe, k = r ? q - hxe - 1 : q;
      if (r ? !in1 : !in0) continue;
      const double latj = P.lat[j0 + r], cj = r ? c01 : c00;
      const double sl = sin((latj - P.lat[jj]) * kGeoDeg2Rad * 0.5);
      const double a = sl * sl + P.s2[k] * cj * c1;
      const double deg = 2.0 * atan2(sqrt(a), sqrt(1.0 - a));
      const double delta = kGeoRadius * deg * kGeoKm2Deg;
      (r ? w1 : w0)[k] = exp(-(delta * delta) / two_sig_sq);
    }
    const double *row = data + (size_t)jj * nx;
    for (int q = tid; q < nx; q += blockDim.x) dl[q] = row[q];
    __syncthreads();
    if (i < nx) {
      const double *dp = dl + i;
      if (in0 && in1) {
        // four nodes' loads (twelve LDS reads) are issued before their sums
        int k = k0;
        for (; k + 4 <= 0; k += 4) {
          double d[4], x0[4], x1[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            d[u] = dp[k + u];
            x0[u] = w0[-(k + u)];
            x1[u] = w1[-(k + u)];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            s0 += x0[u] * d[u];
            t0 += x0[u];
            s1 += x1[u] * d[u];
            t1 += x1[u];
          }
        }
        for (; k < 0; ++k) {
          const double d = dp[k], x0 = w0[-k], x1 = w1[-k];
          s0 += x0 * d;
          t0 += x0;
          s1 += x1 * d;
          t1 += x1;
        }
        for (; k + 4 <= k1 + 1; k += 4) {
          double d[4], x0[4], x1[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            d[u] = dp[k + u];
            x0[u] = w0[k + u];
            x1[u] = w1[k + u];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            s0 += x0[u] * d[u];
            t0 += x0[u];
            s1 += x1[u] * d[u];
            t1 += x1[u];
          }
        }
        for (; k <= k1; ++k) {
          const double d = dp[k], x0 = w0[k], x1 = w1[k];
          s0 += x0 * d;
          t0 += x0;
          s1 += x1 * d;
          t1 += x1;
        }
      } else {
        const double *w = in0 ? w0 : w1;
        double s = in0 ? s0 : s1, t = in0 ? t0 : t1;
        int k = k0;
        for (; k < 0; ++k) {
          const double x = w[-k];
          s += x * dp[k];
          t += x;
        }
        for (; k <= k1; ++k) {
          const double x = w[k];
          s += x * dp[k];
          t += x;
        }
        if (in0) {
          s0 = s;
          t0 = t;
        } else {
          s1 = s;
          t1 = t;
        }
      }
    }
    __syncthreads();
  }
  if (i < nx) {
    P.out[(size_t)f * nx * ny + (size_t)j0 * nx + i] = (0.0 + s0) / t0;
    if (two) P.out[(size_t)f * nx * ny + (size_t)(j0 + 1) * nx + i] = (0.0 + s1) / t1;
  }
}

// gradient_2_geo of field p (natural layout) at cell k; cosd of each row
// comes from the host.
__device__ __forceinline__ void grad_geo_cell(const GeoGrad &G, const double *f, int i, int j,
                                              double &tx, double &ty) {
  const int nx = G.nx, ny = G.ny;
  const size_t k = (size_t)i + (size_t)nx * j;
  const double cl = G.cosd_lat[j];
  const double gfx = i > 0 ? f[k] - f[k - 1] : 0.0;
  const double gbx = i < nx - 1 ? f[k + 1] - f[k] : 0.0;
  const double gfy = j > 0 ? f[k] - f[k - nx] : 0.0;
  const double gby = j < ny - 1 ? f[k + nx] - f[k] : 0.0;
  if (i == 0) {
    tx = gbx / (kGeoRadius * (G.lon[1] - G.lon[0]) * kGeoDeg2Rad * cl);
  } else {
    const double dxi = kGeoRadius * (G.lon[i] - G.lon[i - 1]) * kGeoDeg2Rad * cl;
    tx = i == nx - 1 ? gfx / dxi : (gfx + gbx) / (2 * dxi);
  }
  if (j == 0) {
    ty = gby / (kGeoRadius * (G.lat[1] - G.lat[0]) * kGeoDeg2Rad);
  } else {
    const double dyj = kGeoRadius * (G.lat[j] - G.lat[j - 1]) * kGeoDeg2Rad;
    ty = j == ny - 1 ? gfy / dyj : (gfy + gby) / (2 * dyj);
  }
}

// One thread per (field p fastest, cell): the period-major outputs are
// written contiguously.
__global__ void k_gradient_geo2(GeoGrad G, int nf, const double *f_nat, double *tx,
                                double *ty) {
  const size_t N = (size_t)G.nx * G.ny;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= N * nf) return;
  const int p = (int)(t % nf);
  const size_t k = t / nf;
  double gx, gy;
  grad_geo_cell(G, f_nat + (size_t)p * N, (int)(k % G.nx), (int)(k / G.nx), gx, gy);
  tx[t] = gx;
  ty[t] = gy;
}

__global__ void k_topo_metric(GeoGrad G, int nf, const double *sm_nat, double *m11, double *m12,
                              double *m22, double *angle) {
  const size_t N = (size_t)G.nx * G.ny;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= N * nf) return;
  const int p = (int)(t % nf);
  const size_t k = t / nf;
  double gx, gy;
  grad_geo_cell(G, sm_nat + (size_t)p * N, (int)(k % G.nx), (int)(k / G.nx), gx, gy);
  const double den = 1 + gx * gx + gy * gy;
  m11[t] = (1 + gy * gy) / den;
  m22[t] = (1 + gx * gx) / den;
  m12[t] = -(gx * gy / den);
  if (angle) angle[t] = atan(sqrt(gx * gx + gy * gy)) * (180.0 / 3.14159265358979323846);
}

__global__ void k_pm_to_nat(size_t N, int nf, const double *pm, double *nat) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (k >= N) return;
  nat[(size_t)p * N + k] = pm[(size_t)p + (size_t)nf * k];
}

// max |v| over n values: per-block maxima, then one block over those
constexpr int kMaxBlocks = 1024;
__global__ __launch_bounds__(256) void k_absmax_part(const double *v, size_t n, double *part) {
  __shared__ double sh[256];
  double m = 0;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const double a = fabs(v[k]);
    m = a > m ? a : m;
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x + s] > sh[threadIdx.x]
                                                    ? sh[threadIdx.x + s] : sh[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// part[nparts] = max over part[0..nparts-1], one block
__global__ __launch_bounds__(256) void k_absmax_final(double *part, int nparts) {
  __shared__ double sh[256];
  double m = 0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) m = part[b] > m ? part[b] : m;
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x + s] > sh[threadIdx.x]
                                                    ? sh[threadIdx.x + s] : sh[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[nparts] = sh[0];
}

// svel = svel*(1 + updatemax*gradient), gradient = ker/maxval(abs(ker))
__global__ void k_sd_update(size_t n, const double *ker, const double *mx, double updatemax,
                            double *svel) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  svel[k] = svel[k] * (1 + updatemax * (ker[k] / *mx));
}

// out = base*(1 + update), update = updatemax*ker/maxval(abs(ker)) with the
// product formed first, as prepare_fwd_linesearch writes it
// (src/acqui_2d.f90:179-180); out may be base
__global__ void k_ls_model(size_t n, const double *base, const double *ker, const double *mx,
                           double updatemax, double *out) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  out[k] = base[k] * (1 + updatemax * ker[k] / *mx);
}

dim3 flat_grid(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

size_t smooth_lds_bytes(int nx) { return 2 * (size_t)nx * sizeof(double); }

void launch_geo_s2(hipStream_t s, const double *lon, int nx, double *s2) {
  hipLaunchKernelGGL(k_geo_s2, flat_grid((size_t)nx), dim3(256), 0, s, lon, nx, s2);
}

void launch_smooth_geo2(hipStream_t s, const GeoSmooth &P, int nfield) {
  const dim3 grid((unsigned)P.ny, (unsigned)nfield);
  const size_t lds = smooth_lds_bytes(P.nx);
  if (P.nx <= 1024) {  // one cell per thread, two output rows per workgroup
    hipLaunchKernelGGL(k_smooth_geo2_pair, dim3((unsigned)((P.ny + 1) / 2), (unsigned)nfield),
                       dim3((P.nx + 63) / 64 * 64), 3 * (size_t)P.nx * sizeof(double), s, P);
  } else if (P.nx <= 2048) {
    hipLaunchKernelGGL((k_smooth_geo2<4, 512>), grid, dim3(512), lds, s, P);
  } else {  // nx <= 4096 (SATT_GEO_MAX_NX)
    hipLaunchKernelGGL((k_smooth_geo2<8, 512>), grid, dim3(512), lds, s, P);
  }
}

void launch_gradient_geo2(hipStream_t s, const GeoGrad &G, int nf, const double *f_nat,
                          double *tx_pm, double *ty_pm) {
  hipLaunchKernelGGL(k_gradient_geo2, flat_grid((size_t)G.nx * G.ny * nf), dim3(256), 0, s, G,
                     nf, f_nat, tx_pm, ty_pm);
}

void launch_topo_metric(hipStream_t s, const GeoGrad &G, int nf, const double *sm_nat,
                        double *m11, double *m12, double *m22, double *angle) {
  hipLaunchKernelGGL(k_topo_metric, flat_grid((size_t)G.nx * G.ny * nf), dim3(256), 0, s, G, nf,
                     sm_nat, m11, m12, m22, angle);
}

void launch_pm_to_nat(hipStream_t s, size_t N, int nf, const double *pm, double *nat) {
  hipLaunchKernelGGL(k_pm_to_nat, dim3((unsigned)((N + 255) / 256), (unsigned)nf), dim3(256), 0,
                     s, N, nf, pm, nat);
}

int absmax_parts(size_t n) { return (int)std::min<size_t>(kMaxBlocks, (n + 255) / 256); }

const double *launch_absmax(hipStream_t s, size_t n, const double *v, double *part) {
  const int nb = absmax_parts(n);
  hipLaunchKernelGGL(k_absmax_part, dim3((unsigned)nb), dim3(256), 0, s, v, n, part);
  hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(256), 0, s, part, nb);
  return part + nb;
}

void launch_sd_update(hipStream_t s, size_t n, const double *ker, double updatemax, double *svel,
                      double *part) {
  const double *mx = launch_absmax(s, n, ker, part);
  hipLaunchKernelGGL(k_sd_update, flat_grid(n), dim3(256), 0, s, n, ker, mx, updatemax, svel);
}

void launch_ls_model(hipStream_t s, size_t n, const double *base, const double *ker,
                     const double *mx, double updatemax, double *out) {
  hipLaunchKernelGGL(k_ls_model, flat_grid(n), dim3(256), 0, s, n, base, ker, mx, updatemax, out);
}

}  // namespace satt
