// The allocation-free per-solve forward solve, satt_fwd_lonlat_noalloc
// (include/satt.h): the kernels of satt_fwd_lonlat on the same operands, with
// every device buffer beyond the caller's arrays carved from one file-scope
// device array instead of hipMalloc.
//
// Each GPU the library runs on holds its own copy of the array for the life
// of the process.  Host code never passes the array's name to a kernel or a
// copy (in host code that name is only a stand-in for the symbol): satt_create
// looks up the array's device address once per device, with that device
// current, and every use goes through that address.  Calls on one GPU share
// the array, so calls that overlap on one GPU race; each call rewrites
// everything it reads, so nothing is left over from the previous call.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/satt.h"
#include "satt_internal.h"

using namespace satt;

namespace {

// Doubles of the scratch: per cell the metric conversion's inputs and
// outputs (spha, sphb, sphc, a, b, c, fun) and the forward's fields (T, tau,
// tau_old, T0, T0x, T0y); per axis node the radian and degree axes and
// sin/cos of the latitudes; one item record.  Every piece starts on a
// 256-byte boundary.
constexpr size_t kAl = 32;  // doubles per 256 bytes
constexpr size_t al(size_t n) { return (n + kAl - 1) / kAl * kAl; }
constexpr size_t kScratchDoubles =
    13 * al((size_t)SATT_NOALLOC_MAX_CELLS) + 6 * al((size_t)SATT_NOALLOC_MAX_AXIS) + kAl;

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices];  // scratch's device address, per device

// The item record of one solve (the layout satt_fwd_lonlat uploads).
struct ItemRec {
  int pidx, rec_off[2], st, it;
  double sx, sy;
};
static_assert(sizeof(ItemRec) <= kAl * sizeof(double), "item record exceeds its piece");

// Radian axes by the reference's xx_deg/180.0*pi (src/eikon_solver_2d.f90:
// 801-804), the expression satt_fwd_lonlat evaluates on the host.
__global__ void k_deg2rad(const double *deg, int n, double *rad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rad[i] = deg[i] / 180.0 * kPiF;
}

}  // namespace

int satt::noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return kEArg;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return kOk;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return kEHip;
  }
  g_scratch[device] = static_cast<double *>(p);
  return kOk;
}

extern "C" int satt_fwd_lonlat_noalloc(satt_ctx *ctx, const double *xx_deg,
                                       const double *yy_deg, int nx, int ny,
                                       const double *spha, const double *sphb,
                                       const double *sphc, const double *fun, double x0_deg,
                                       double y0_deg, double *T, int *iters) {
  if (!ctx) return SATT_E_ARG;
  if (!xx_deg || !yy_deg || nx < 3 || ny < 3 || !spha || !sphb || !sphc || !fun || !T)
    return ctx_fail(ctx, SATT_E_ARG, "satt_fwd_lonlat_noalloc: bad arguments");
  if (nx > SATT_NOALLOC_MAX_AXIS || ny > SATT_NOALLOC_MAX_AXIS ||
      (long long)nx * ny > SATT_NOALLOC_MAX_CELLS)
    return ctx_fail(ctx, SATT_E_ARG,
                    "satt_fwd_lonlat_noalloc: a %d x %d grid exceeds the scratch (at most %d "
                    "cells and %d nodes per axis)",
                    nx, ny, SATT_NOALLOC_MAX_CELLS, SATT_NOALLOC_MAX_AXIS);
  const int dev = ctx_device(ctx);
  double *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (dev >= 0 && dev < kMaxDevices) base = g_scratch[dev];
  }
  if (!base) return ctx_fail(ctx, SATT_E_HIP, "satt_fwd_lonlat_noalloc: no scratch on device %d", dev);
  if (hipSetDevice(dev) != hipSuccess) return ctx_fail(ctx, SATT_E_HIP, "hipSetDevice failed");
  hipStream_t s = ctx_stream(ctx);
  const size_t N = (size_t)nx * ny, NA = al((size_t)SATT_NOALLOC_MAX_CELLS),
               XA = al((size_t)SATT_NOALLOC_MAX_AXIS);
  // the pieces, at fixed offsets for every grid size
  double *xx = base, *yy = base + XA, *xd = base + 2 * XA, *yd = base + 3 * XA,
         *sy = base + 4 * XA, *cy = base + 5 * XA;
  char *item = reinterpret_cast<char *>(base + 6 * XA);
  double *cells = base + 6 * XA + kAl;
  double *sa = cells, *sb = cells + NA, *sc = cells + 2 * NA, *a = cells + 3 * NA,
         *b = cells + 4 * NA, *c = cells + 5 * NA, *f = cells + 6 * NA;
  FwdWork fw{cells + 7 * NA, cells + 8 * NA, cells + 9 * NA, cells + 10 * NA, cells + 11 * NA,
             cells + 12 * NA};

  ItemRec hi{0, {0, 0}, 0, 0, x0_deg, y0_deg};
  hipError_t e = hipSuccess;
  auto cp = [&](void *dst, const void *src, size_t bytes) {
    if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  };
  cp(xd, xx_deg, (size_t)nx * 8);
  cp(yd, yy_deg, (size_t)ny * 8);
  cp(item, &hi, sizeof hi);
  cp(sa, spha, N * 8);
  cp(sb, sphb, N * 8);
  cp(sc, sphc, N * 8);
  cp(f, fun, N * 8);
  if (e != hipSuccess)
    return ctx_fail(ctx, SATT_E_HIP, "satt_fwd_lonlat_noalloc: upload: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(k_deg2rad, dim3((unsigned)((nx + 63) / 64)), dim3(64), 0, s, xd, nx, xx);
  hipLaunchKernelGGL(k_deg2rad, dim3((unsigned)((ny + 63) / 64)), dim3(64), 0, s, yd, ny, yy);
  Grid g{};
  g.nx = nx;
  g.ny = ny;
  // the host's dx, dy from the same expression as the device's axes
  g.dx = xx_deg[1] / 180.0 * kPiF - xx_deg[0] / 180.0 * kPiF;
  g.dy = yy_deg[1] / 180.0 * kPiF - yy_deg[0] / 180.0 * kPiF;
  g.xx = xx;
  g.yy = yy;
  g.xd = xd;
  g.yd = yd;
  g.sy = sy;
  g.cy = cy;
  launch_grid_trig(s, g, sy, cy);
  launch_metric_fields(s, g, sa, sb, sc, a, b, c);
  PeriodFields pf{a, b, c, f, nullptr};
  Items it{};
  it.item0 = 0;
  it.count = 1;
  it.pidx = reinterpret_cast<int *>(item + offsetof(ItemRec, pidx));
  it.rec_off = reinterpret_cast<int *>(item + offsetof(ItemRec, rec_off));
  it.srcx = reinterpret_cast<double *>(item + offsetof(ItemRec, sx));
  it.srcy = reinterpret_cast<double *>(item + offsetof(ItemRec, sy));
  int *d_st = reinterpret_cast<int *>(item + offsetof(ItemRec, st));
  int *d_it = reinterpret_cast<int *>(item + offsetof(ItemRec, it));
  launch_forward(s, g, pf, it, fw, d_it, d_st, 1);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&hi, item, sizeof hi, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(T, fw.T, N * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return ctx_fail(ctx, SATT_E_HIP, "satt_fwd_lonlat_noalloc: %s", hipGetErrorString(e));
  if (iters) *iters = hi.it;
  if (hi.st != SATT_OK) ctx_fail(ctx, hi.st, "satt_fwd_lonlat_noalloc: status %d", hi.st);
  return hi.st;
}
