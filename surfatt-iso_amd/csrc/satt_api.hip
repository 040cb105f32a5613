// C-ABI of libsatt.so (declared in include/satt.h).
//
// The host side keeps the reference's data shapes: fields (nx,ny) column-major,
// model arrays (nperiod,nx,ny) period-major, the isrcs item order and a CSR
// receiver gather.  Device buffers belong to a context (per-solve calls) or to
// a plan (batched calls); nothing the caller passes is retained.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/satt.h"
#include "satt_internal.h"
#include "satt_strip.h"

using namespace satt;

namespace {
// Everything of a satt_batch except the model (svel, m11, m12, m22), with the
// precision: what satt_forward_simulate's resident plan is keyed on.
struct FsKey {
  int dim = 0, nx = 0, ny = 0, nperiod = 0, nitems = 0, npath = 0, precision = -1;
  std::vector<double> axes, items, recs;
  std::vector<int> ints;
  static FsKey of(const satt_batch *B, int precision) {
    FsKey k;
    k.dim = B->dim; k.nx = B->nx; k.ny = B->ny; k.nperiod = B->nperiod;
    k.nitems = B->nitems; k.npath = B->npath; k.precision = precision;
    const int ni = B->nitems > 0 ? B->nitems : 0;
    const int nr = ni > 0 ? B->rec_off[ni] : 0;
    k.axes.assign(B->xgrids, B->xgrids + B->nx);
    k.axes.insert(k.axes.end(), B->ygrids, B->ygrids + B->ny);
    if (ni) {
      k.items.assign(B->item_srcx, B->item_srcx + ni);
      k.items.insert(k.items.end(), B->item_srcy, B->item_srcy + ni);
      k.ints.assign(B->item_pidx, B->item_pidx + ni);
      k.ints.insert(k.ints.end(), B->rec_off, B->rec_off + ni + 1);
    }
    if (nr) {
      for (const double *a : {B->recx, B->recy, B->tt, B->weight}) k.recs.insert(k.recs.end(), a, a + nr);
      k.ints.insert(k.ints.end(), B->idx_path, B->idx_path + nr);
    }
    return k;
  }
  bool operator==(const FsKey &o) const {
    return dim == o.dim && nx == o.nx && ny == o.ny && nperiod == o.nperiod &&
           nitems == o.nitems && npath == o.npath && precision == o.precision &&
           axes == o.axes && items == o.items && recs == o.recs && ints == o.ints;
  }
};
}  // namespace

struct satt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::mutex tune_mu;
  std::map<TuneKey, TuneWaves> tuned;  // team sizes measured by satt_tune_waves
  double last_ms = 0;            // device time of the last model-side call
  // satt_forward_simulate's resident plan and what it was made from: the
  // reference calls forward_simulate once per iteration and per line-search
  // step with the same gathers and a new model (src/tomo2d.f90:104,272), so
  // a call whose batch differs only in the model refreshes it in place
  satt_plan *fs_plan = nullptr;
  FsKey fs_key;
};

bool satt::ctx_tune_lookup(satt_ctx *ctx, const TuneKey &k, TuneWaves &out) {
  std::lock_guard<std::mutex> lk(ctx->tune_mu);
  auto it = ctx->tuned.find(k);
  if (it == ctx->tuned.end()) return false;
  out = it->second;
  return true;
}

void satt::ctx_tune_store(satt_ctx *ctx, const TuneKey &k, const TuneWaves &waves) {
  std::lock_guard<std::mutex> lk(ctx->tune_mu);
  ctx->tuned[k] = waves;
}

namespace {

int fail(satt_ctx *ctx, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                          \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail((ctx), SATT_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

// Device allocation owned by a scope or a plan.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t n) {
    release();
    bytes = n;
    return hipMalloc(&p, n ? n : 1);
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

#define ALLOC(ctx, buf, n)                                                          \
  do {                                                                              \
    if ((buf).alloc(n) != hipSuccess) {                                            \
      (void)hipGetLastError();                                                      \
      return fail((ctx), SATT_E_NOMEM, "device allocation of %zu bytes failed",    \
                  (size_t)(n));                                                     \
    }                                                                               \
  } while (0)

// Grid axes on the device: degrees as given, radians by the reference's
// xx_deg/180.0*pi (src/eikon_solver_2d.f90:801-804).
struct DevGrid {
  DBuf buf;
  Grid g{};
};

int make_grid(satt_ctx *ctx, const double *xd, const double *yd, int nx, int ny,
              DevGrid &dg) {
  if (!xd || !yd || nx < 2 || ny < 2) return fail(ctx, SATT_E_ARG, "bad grid axes");
  std::vector<double> h(2 * (size_t)(nx + ny) + 2 * (size_t)ny);
  double *hx = h.data(), *hy = hx + nx, *hxd = hy + ny, *hyd = hxd + nx;
  for (int i = 0; i < nx; ++i) hx[i] = xd[i] / 180.0 * kPiF;
  for (int j = 0; j < ny; ++j) hy[j] = yd[j] / 180.0 * kPiF;
  std::memcpy(hxd, xd, sizeof(double) * nx);
  std::memcpy(hyd, yd, sizeof(double) * ny);
  ALLOC(ctx, dg.buf, h.size() * sizeof(double));
  HIPCHK(ctx, hipMemcpyAsync(dg.buf.p, h.data(), h.size() * sizeof(double),
                             hipMemcpyHostToDevice, ctx->stream));
  double *d = dg.buf.as<double>();
  dg.g.nx = nx;
  dg.g.ny = ny;
  dg.g.dx = hx[1] - hx[0];
  dg.g.dy = hy[1] - hy[0];
  dg.g.xx = d;
  dg.g.yy = d + nx;
  dg.g.xd = d + nx + ny;
  dg.g.yd = d + 2 * nx + ny;
  double *sy = d + 2 * (nx + ny), *cy = sy + ny;
  dg.g.sy = sy;
  dg.g.cy = cy;
  launch_grid_trig(ctx->stream, dg.g, sy, cy);
  HIPCHK(ctx, hipGetLastError());
  return SATT_OK;
}

int upload(satt_ctx *ctx, DBuf &b, const void *h, size_t bytes) {
  ALLOC(ctx, b, bytes);
  if (bytes) HIPCHK(ctx, hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, ctx->stream));
  return SATT_OK;
}

#define TRY(expr)                 \
  do {                            \
    int st_ = (expr);             \
    if (st_ < 0) return st_;      \
  } while (0)

int sync_status(satt_ctx *ctx) {
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return SATT_OK;
}

const char *status_text(int st) {
  switch (st) {
  case SATT_E_SRC_ON_BOUNDARY: return "source on the boundary, mesh error";
  case SATT_E_REC_OUT_OF_GRID: return "receiver outside the grid";
  case SATT_W_MAXITER: return "maximum iteration count reached";
  default: return "error";
  }
}

}  // namespace

// ------------------------------------------------------------------ context

int satt::ctx_device(const satt_ctx *ctx) { return ctx->device; }
hipStream_t satt::ctx_stream(const satt_ctx *ctx) { return ctx->stream; }
int satt::ctx_fail(satt_ctx *ctx, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

extern "C" const char *satt_version(void) { return "satt 0.1 (gfx950)"; }

extern "C" int satt_create(satt_ctx **out, int device) {
  if (!out) return SATT_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return SATT_E_NODEV;
  }
  if (device < 0 || device >= n) return SATT_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return SATT_E_HIP;
  // the allocation-free entry point's scratch address on this device
  if (noalloc_init(device) != SATT_OK) return SATT_E_HIP;
  satt_ctx *ctx = new satt_ctx();
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return SATT_E_HIP;
  }
  *out = ctx;
  return SATT_OK;
}

extern "C" void satt_destroy(satt_ctx *ctx) {
  if (!ctx) return;
  satt_forward_simulate_release(ctx);
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char *satt_last_error(const satt_ctx *ctx) {
  return ctx ? ctx->err.c_str() : "no context";
}

// ------------------------------------------------------- per-solve (plain)

namespace {

// Single-period fields for the per-solve entry points.
struct OneSolve {
  DevGrid dg;
  DBuf fields;  // a, b, c, fun, spha, sphb, sphc scratch
  PeriodFields pf{};
};

int one_solve_fields(satt_ctx *ctx, OneSolve &o, const double *spha, const double *sphb,
                     const double *sphc, const double *fun) {
  const size_t N = (size_t)o.dg.g.nx * o.dg.g.ny;
  ALLOC(ctx, o.fields, 7 * N * sizeof(double));
  double *d = o.fields.as<double>();
  double *a = d, *b = d + N, *c = d + 2 * N, *f = d + 3 * N, *sa = d + 4 * N,
         *sb = d + 5 * N, *sc = d + 6 * N;
  HIPCHK(ctx, hipMemcpyAsync(sa, spha, N * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(sb, sphb, N * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(sc, sphc, N * 8, hipMemcpyHostToDevice, ctx->stream));
  if (fun) HIPCHK(ctx, hipMemcpyAsync(f, fun, N * 8, hipMemcpyHostToDevice, ctx->stream));
  launch_metric_fields(ctx->stream, o.dg.g, sa, sb, sc, a, b, c);
  o.pf.a = a;
  o.pf.b = b;
  o.pf.c = c;
  o.pf.fun = f;
  o.pf.vel = nullptr;
  return SATT_OK;
}

bool bad_fields(int nx, int ny, const void *a, const void *b, const void *c) {
  return nx < 3 || ny < 3 || !a || !b || !c;
}

}  // namespace

extern "C" int satt_fwd_lonlat(satt_ctx *ctx, const double *xx_deg, const double *yy_deg,
                               int nx, int ny, const double *spha, const double *sphb,
                               const double *sphc, const double *fun, double x0_deg,
                               double y0_deg, double *T, int *iters) {
  if (!ctx) return SATT_E_ARG;
  if (bad_fields(nx, ny, spha, sphb, sphc) || !fun || !T)
    return fail(ctx, SATT_E_ARG, "satt_fwd_lonlat: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)nx * ny;
  OneSolve o;
  TRY(make_grid(ctx, xx_deg, yy_deg, nx, ny, o.dg));
  TRY(one_solve_fields(ctx, o, spha, sphb, sphc, fun));
  struct HostItem { int pidx, rec_off[2], st, it; double sx, sy; } hi{0, {0, 0}, 0, 0, x0_deg, y0_deg};
  DBuf ib, work;
  TRY(upload(ctx, ib, &hi, sizeof hi));
  char *ibase = ib.as<char>();
  Items it{};
  it.item0 = 0;
  it.count = 1;
  it.pidx = reinterpret_cast<int *>(ibase + offsetof(HostItem, pidx));
  it.rec_off = reinterpret_cast<int *>(ibase + offsetof(HostItem, rec_off));
  it.srcx = reinterpret_cast<double *>(ibase + offsetof(HostItem, sx));
  it.srcy = reinterpret_cast<double *>(ibase + offsetof(HostItem, sy));
  int *d_st = reinterpret_cast<int *>(ibase + offsetof(HostItem, st));
  int *d_it = reinterpret_cast<int *>(ibase + offsetof(HostItem, it));
  ALLOC(ctx, work, 6 * N * sizeof(double));
  double *wd = work.as<double>();
  FwdWork fw{wd, wd + N, wd + 2 * N, wd + 3 * N, wd + 4 * N, wd + 5 * N};
  launch_forward(ctx->stream, o.dg.g, o.pf, it, fw, d_it, d_st, 1);
  HIPCHK(ctx, hipMemcpyAsync(&hi, ibase, sizeof hi, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(T, fw.T, N * 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  if (iters) *iters = hi.it;
  if (hi.st != SATT_OK) fail(ctx, hi.st, "satt_fwd_lonlat: %s", status_text(hi.st));
  return hi.st;
}

extern "C" int satt_adj_lonlat(satt_ctx *ctx, const double *xx_deg, const double *yy_deg,
                               int nx, int ny, const double *spha, const double *sphb,
                               const double *sphc, const double *T, const double *xrec_deg,
                               const double *yrec_deg, const double *src_adj, int nr,
                               double *Ta, int *iters) {
  if (!ctx) return SATT_E_ARG;
  if (bad_fields(nx, ny, spha, sphb, sphc) || !T || !Ta || nr < 0 ||
      (nr > 0 && (!xrec_deg || !yrec_deg || !src_adj)))
    return fail(ctx, SATT_E_ARG, "satt_adj_lonlat: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)nx * ny;
  OneSolve o;
  TRY(make_grid(ctx, xx_deg, yy_deg, nx, ny, o.dg));
  TRY(one_solve_fields(ctx, o, spha, sphb, sphc, nullptr));
  struct HostItem { int pidx, rec_off[2], st, it; double sx, sy; } hi{0, {0, nr}, 0, 0, 0.0, 0.0};
  DBuf ib, rec, work;
  TRY(upload(ctx, ib, &hi, sizeof hi));
  std::vector<double> hr(3 * (size_t)nr + 1);
  if (nr) {
    std::memcpy(hr.data(), xrec_deg, nr * 8);
    std::memcpy(hr.data() + nr, yrec_deg, nr * 8);
    std::memcpy(hr.data() + 2 * nr, src_adj, nr * 8);
  }
  TRY(upload(ctx, rec, hr.data(), hr.size() * 8));
  char *ibase = ib.as<char>();
  Items it{};
  it.item0 = 0;
  it.count = 1;
  it.pidx = reinterpret_cast<int *>(ibase + offsetof(HostItem, pidx));
  it.rec_off = reinterpret_cast<int *>(ibase + offsetof(HostItem, rec_off));
  it.srcx = reinterpret_cast<double *>(ibase + offsetof(HostItem, sx));
  it.srcy = reinterpret_cast<double *>(ibase + offsetof(HostItem, sy));
  it.recx = rec.as<double>();
  it.recy = rec.as<double>() + nr;
  int *d_st = reinterpret_cast<int *>(ibase + offsetof(HostItem, st));
  int *d_it = reinterpret_cast<int *>(ibase + offsetof(HostItem, it));
  ALLOC(ctx, work, 10 * N * sizeof(double));
  double *wd = work.as<double>();
  HIPCHK(ctx, hipMemcpyAsync(wd, T, N * 8, hipMemcpyHostToDevice, ctx->stream));
  AdjWork aw{wd, wd + N, wd + 2 * N, wd + 7 * N, wd + 8 * N, rec.as<double>() + 2 * nr};
  launch_adjoint(ctx->stream, o.dg.g, o.pf, it, aw, d_it, d_st, 1, 0);
  HIPCHK(ctx, hipMemcpyAsync(&hi, ibase, sizeof hi, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(Ta, aw.Ta, N * 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  if (iters) *iters = hi.it;
  if (hi.st != SATT_OK) fail(ctx, hi.st, "satt_adj_lonlat: %s", status_text(hi.st));
  return hi.st;
}

extern "C" int satt_mask(satt_ctx *ctx, const double *xx, const double *yy, int nx, int ny,
                         double *table, double x0, double y0) {
  if (!ctx) return SATT_E_ARG;
  if (!table || nx < 2 || ny < 2) return fail(ctx, SATT_E_ARG, "satt_mask: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)nx * ny;
  DevGrid dg;
  TRY(make_grid(ctx, xx, yy, nx, ny, dg));
  DBuf t;
  TRY(upload(ctx, t, table, N * 8));
  launch_mask(ctx->stream, dg.g, t.as<double>(), x0, y0);
  HIPCHK(ctx, hipMemcpyAsync(table, t.p, N * 8, hipMemcpyDeviceToHost, ctx->stream));
  return sync_status(ctx);
}

extern "C" int satt_interp2(satt_ctx *ctx, const double *x, const double *y, int nx, int ny,
                            const double *v, const double *xq, const double *yq, int nq,
                            double *vq) {
  if (!ctx) return SATT_E_ARG;
  if (!v || nq < 0 || (nq > 0 && (!xq || !yq || !vq)) || nx < 2 || ny < 2)
    return fail(ctx, SATT_E_ARG, "satt_interp2: bad arguments");
  if (nq == 0) return SATT_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)nx * ny;
  DevGrid dg;
  TRY(make_grid(ctx, x, y, nx, ny, dg));
  DBuf dv, dq;
  TRY(upload(ctx, dv, v, N * 8));
  std::vector<double> hq(3 * (size_t)nq + 1);
  std::memcpy(hq.data(), xq, nq * 8);
  std::memcpy(hq.data() + nq, yq, nq * 8);
  int st = 0;
  std::memcpy(hq.data() + 3 * (size_t)nq, &st, sizeof st);
  TRY(upload(ctx, dq, hq.data(), hq.size() * 8));
  double *q = dq.as<double>();
  int *d_st = reinterpret_cast<int *>(q + 3 * (size_t)nq);
  launch_interp2(ctx->stream, dg.g, dv.as<double>(), q, q + nq, nq, q + 2 * nq, d_st);
  HIPCHK(ctx, hipMemcpyAsync(vq, q + 2 * nq, nq * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&st, d_st, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  if (st != SATT_OK) return fail(ctx, st, "satt_interp2: %s", status_text(st));
  return SATT_OK;
}

// --------------------------------------------------------------- inspection

// ------------------------------------------------------- model-side steps

namespace {

// cosd of the Fortran runtime: cos of the argument times pi/180 (true pi)
double cosd_host(double x) { return std::cos(x * (3.14159265358979323846 / 180.0)); }

int geo_axes_check(satt_ctx *ctx, const char *who, const double *lons, const double *lats,
                   int nx, int ny) {
  if (!lons || !lats || nx < 2 || ny < 2) return fail(ctx, SATT_E_ARG, "%s: bad grid axes", who);
  if (nx > SATT_GEO_MAX_NX)
    return fail(ctx, SATT_E_ARG, "%s: nx %d above %d", who, nx, SATT_GEO_MAX_NX);
  const double d = lons[1] - lons[0];
  if (!(d > 0) || !(lats[1] - lats[0] > 0))
    return fail(ctx, SATT_E_ARG, "%s: axes must increase", who);
  for (int i = 2; i < nx; ++i)
    if (!(std::fabs((lons[i] - lons[0]) - i * d) <= 1e-6 * d))
      return fail(ctx, SATT_E_ARG, "%s: longitude axis not uniform at %d", who, i);
  for (int j = 1; j < ny; ++j)
    if (!(lats[j] > lats[j - 1])) return fail(ctx, SATT_E_ARG, "%s: axes must increase", who);
  if ((size_t)nx * ny > (size_t)INT_MAX) return fail(ctx, SATT_E_ARG, "%s: grid too large", who);
  return SATT_OK;
}

// fields ride the grid's y dimension in the launches
int geo_fields_check(satt_ctx *ctx, const char *who, int nfield) {
  if (nfield < 1 || nfield > 65535)
    return fail(ctx, SATT_E_ARG, "%s: %d fields (1..65535)", who, nfield);
  return SATT_OK;
}

// Events around the kernels of one call; the elapsed time goes to last_ms.
struct GeoTimer {
  hipEvent_t a = nullptr, b = nullptr;
  ~GeoTimer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

int geo_timer_start(satt_ctx *ctx, GeoTimer &t) {
  HIPCHK(ctx, hipEventCreate(&t.a));
  HIPCHK(ctx, hipEventCreate(&t.b));
  HIPCHK(ctx, hipEventRecord(t.a, ctx->stream));
  return SATT_OK;
}

int geo_timer_stop(satt_ctx *ctx, GeoTimer &t) {
  HIPCHK(ctx, hipEventRecord(t.b, ctx->stream));
  return SATT_OK;
}

int geo_timer_read(satt_ctx *ctx, GeoTimer &t) {
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, t.a, t.b));
  ctx->last_ms = ms;
  return SATT_OK;
}

// Device state of a smoothing: axes, s2, per-field sigma and window
// half-widths (the reference's truncations, src/shared/utils.f90:1204-1212,
// evaluated on the host), the natural-layout output.
struct GeoSmoothBufs {
  DBuf axes, tab, out;
  GeoSmooth P{};
  GeoGrad G{};
};

int geo_smooth_setup(satt_ctx *ctx, const char *who, const double *lons, const double *lats,
                     int nx, int ny, int nfield, const double *sigma, GeoSmoothBufs &B) {
  for (int f = 0; f < nfield; ++f)
    if (!(sigma[f] > 0) || !std::isfinite(sigma[f]))
      return fail(ctx, SATT_E_ARG, "%s: sigma[%d] must be positive", who, f);
  const size_t N = (size_t)nx * ny;
  const double dx = lons[1] - lons[0], dy = lats[1] - lats[0];
  // doubles: lon, lat, cosd(lat), s2, sigma; ints: hx, hy
  std::vector<double> hd(2 * (size_t)nx + 2 * (size_t)ny + nfield);
  std::memcpy(hd.data(), lons, nx * 8);
  std::memcpy(hd.data() + nx, lats, ny * 8);
  for (int j = 0; j < ny; ++j) hd[nx + ny + j] = cosd_host(lats[j]);
  std::memcpy(hd.data() + 2 * (size_t)nx + 2 * ny, sigma, nfield * 8);
  std::vector<int> hi((size_t)nfield * (1 + ny));
  for (int f = 0; f < nfield; ++f) {
    // nx = sigma3/dx, ny = sigma3/cosd(lat)/dy, truncated; windows wider
    // than the grid are the same as the grid
    const double sx = sigma[f] * 3 / dx;
    hi[f] = sx < nx ? (int)sx : nx;
    for (int j = 0; j < ny; ++j) {
      const double sy = sigma[f] * 3 / cosd_host(lats[j]) / dy;
      hi[nfield + (size_t)f * ny + j] = sy >= 0 && sy < ny ? (int)sy : ny;
    }
  }
  TRY(upload(ctx, B.axes, hd.data(), hd.size() * 8));
  ALLOC(ctx, B.tab, (size_t)nx * 8 + hi.size() * sizeof(int));
  double *ax = B.axes.as<double>();
  double *s2 = B.tab.as<double>();
  int *dh = reinterpret_cast<int *>(s2 + nx);
  HIPCHK(ctx, hipMemcpyAsync(dh, hi.data(), hi.size() * sizeof(int), hipMemcpyHostToDevice,
                             ctx->stream));
  ALLOC(ctx, B.out, (size_t)nfield * N * 8);
  launch_geo_s2(ctx->stream, ax, nx, s2);
  B.P.nx = nx;
  B.P.ny = ny;
  B.P.lat = ax + nx;
  B.P.s2 = s2;
  B.P.sigma = ax + 2 * (size_t)nx + 2 * ny;
  B.P.hx = dh;
  B.P.hy = dh + nfield;
  B.P.out = B.out.as<double>();
  B.G.nx = nx;
  B.G.ny = ny;
  B.G.lon = ax;
  B.G.lat = ax + nx;
  B.G.cosd_lat = ax + nx + ny;
  return SATT_OK;
}

}  // namespace

extern "C" double satt_last_device_ms(const satt_ctx *ctx) { return ctx ? ctx->last_ms : 0.0; }

extern "C" int satt_hbm_probe(satt_ctx *ctx, double nbytes, double *read_gbs, double *copy_gbs) {
  if (!ctx || !read_gbs || !copy_gbs || !(nbytes >= (1 << 20)) || nbytes > 64.0 * (1ull << 30))
    return SATT_E_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int cus = 0;
  HIPCHK(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const int st = stream_probe(ctx->stream, cus, (size_t)nbytes, 5, *read_gbs, *copy_gbs);
  if (st < 0) return fail(ctx, st, "satt_hbm_probe: %s", st == SATT_E_NOMEM ? "no memory" : "HIP error");
  return SATT_OK;
}

extern "C" int satt_lds_probe(satt_ctx *ctx, double *read_gbs, double *peak_gbs) {
  if (!ctx || !read_gbs || !peak_gbs) return SATT_E_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int cus = 0, khz = 0;
  HIPCHK(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  HIPCHK(ctx, hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, ctx->device));
  // gfx950's LDS returns up to 256 bytes per clock per CU
  *peak_gbs = 256.0 * cus * (khz * 1e3) / 1e9;
  const int st = lds_probe(ctx->stream, cus, 5, *read_gbs);
  if (st < 0) return fail(ctx, st, "satt_lds_probe: %s", st == SATT_E_NOMEM ? "no memory" : "HIP error");
  return SATT_OK;
}

extern "C" double satt_topo_sigma(double vel, double period, double wavelen_factor) {
  const double pi = (double)3.14159265359f;
  const double km2deg = 1.0 / (6371.0 * pi / 180.0);
  return vel * period * wavelen_factor * km2deg / (2.0 * pi);
}

extern "C" int satt_smooth_geo2(satt_ctx *ctx, const double *lons, const double *lats, int nx,
                                int ny, int nfield, int ndata, const double *data,
                                const double *sigma, double *out) {
  if (!ctx) return SATT_E_ARG;
  TRY(geo_axes_check(ctx, "satt_smooth_geo2", lons, lats, nx, ny));
  TRY(geo_fields_check(ctx, "satt_smooth_geo2", nfield));
  if (nfield < 1 || (ndata != 1 && ndata != nfield) || !data || !sigma || !out)
    return fail(ctx, SATT_E_ARG, "satt_smooth_geo2: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)nx * ny;
  GeoSmoothBufs B;
  TRY(geo_smooth_setup(ctx, "satt_smooth_geo2", lons, lats, nx, ny, nfield, sigma, B));
  DBuf din, dnat;
  TRY(upload(ctx, din, data, (size_t)ndata * N * 8));
  const double *nat = din.as<double>();
  if (ndata > 1) {
    ALLOC(ctx, dnat, (size_t)ndata * N * 8);
    launch_pm_to_nat(ctx->stream, N, ndata, din.as<double>(), dnat.as<double>());
    nat = dnat.as<double>();
  }
  B.P.ndata = ndata;
  B.P.data = nat;
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  launch_smooth_geo2(ctx->stream, B.P, nfield);
  TRY(geo_timer_stop(ctx, t));
  double *res = B.out.as<double>();
  if (nfield > 1) {  // back to period-major in the input buffer's place
    DBuf pm;
    ALLOC(ctx, pm, (size_t)nfield * N * 8);
    launch_to_period_major(ctx->stream, (int)N, nfield, res, pm.as<double>());
    HIPCHK(ctx, hipMemcpyAsync(out, pm.p, (size_t)nfield * N * 8, hipMemcpyDeviceToHost,
                               ctx->stream));
    TRY(sync_status(ctx));
  } else {
    HIPCHK(ctx, hipMemcpyAsync(out, res, N * 8, hipMemcpyDeviceToHost, ctx->stream));
    TRY(sync_status(ctx));
  }
  return geo_timer_read(ctx, t);
}

extern "C" int satt_gradient_geo2(satt_ctx *ctx, const double *lons, const double *lats, int nx,
                                  int ny, int nfield, const double *f, double *tx, double *ty) {
  if (!ctx) return SATT_E_ARG;
  TRY(geo_axes_check(ctx, "satt_gradient_geo2", lons, lats, nx, ny));
  TRY(geo_fields_check(ctx, "satt_gradient_geo2", nfield));
  if (nfield < 1 || !f || !tx || !ty)
    return fail(ctx, SATT_E_ARG, "satt_gradient_geo2: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)nx * ny, M = (size_t)nfield * N;
  std::vector<double> hd(2 * (size_t)nx + 2 * (size_t)ny);
  std::memcpy(hd.data(), lons, nx * 8);
  std::memcpy(hd.data() + nx, lats, ny * 8);
  for (int j = 0; j < ny; ++j) hd[nx + ny + j] = cosd_host(lats[j]);
  DBuf ax, din, work;
  TRY(upload(ctx, ax, hd.data(), hd.size() * 8));
  TRY(upload(ctx, din, f, M * 8));
  ALLOC(ctx, work, 3 * M * 8);
  double *nat = work.as<double>(), *dtx = nat + M, *dty = dtx + M;
  GeoGrad G{nx, ny, ax.as<double>(), ax.as<double>() + nx, ax.as<double>() + nx + ny};
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  launch_pm_to_nat(ctx->stream, N, nfield, din.as<double>(), nat);
  launch_gradient_geo2(ctx->stream, G, nfield, nat, dtx, dty);
  TRY(geo_timer_stop(ctx, t));
  HIPCHK(ctx, hipMemcpyAsync(tx, dtx, M * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ty, dty, M * 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  return geo_timer_read(ctx, t);
}

extern "C" int satt_topo_metric(satt_ctx *ctx, const double *lons, const double *lats, int nx,
                                int ny, const double *z, int nperiod, const double *sigma,
                                double *m11, double *m12, double *m22, double *angle) {
  if (!ctx) return SATT_E_ARG;
  TRY(geo_axes_check(ctx, "satt_topo_metric", lons, lats, nx, ny));
  TRY(geo_fields_check(ctx, "satt_topo_metric", nperiod));
  if (nperiod < 1 || !z || !sigma || !m11 || !m12 || !m22)
    return fail(ctx, SATT_E_ARG, "satt_topo_metric: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)nx * ny, M = (size_t)nperiod * N;
  GeoSmoothBufs B;
  TRY(geo_smooth_setup(ctx, "satt_topo_metric", lons, lats, nx, ny, nperiod, sigma, B));
  DBuf dz, res;
  TRY(upload(ctx, dz, z, N * 8));
  ALLOC(ctx, res, (angle ? 4 : 3) * M * 8);
  double *o11 = res.as<double>(), *o12 = o11 + M, *o22 = o12 + M;
  double *oag = angle ? o22 + M : nullptr;
  B.P.ndata = 1;
  B.P.data = dz.as<double>();
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  launch_smooth_geo2(ctx->stream, B.P, nperiod);
  launch_topo_metric(ctx->stream, B.G, nperiod, B.out.as<double>(), o11, o12, o22, oag);
  TRY(geo_timer_stop(ctx, t));
  HIPCHK(ctx, hipMemcpyAsync(m11, o11, M * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(m12, o12, M * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(m22, o22, M * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (angle)
    HIPCHK(ctx, hipMemcpyAsync(angle, oag, M * 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  return geo_timer_read(ctx, t);
}

extern "C" int satt_steepest_descent(satt_ctx *ctx, const double *lons, const double *lats,
                                     int nx, int ny, int nperiod, const double *adj_s,
                                     double sigma_2d, double updatemax, double *svel,
                                     double *ker_s) {
  if (!ctx) return SATT_E_ARG;
  TRY(geo_axes_check(ctx, "satt_steepest_descent", lons, lats, nx, ny));
  TRY(geo_fields_check(ctx, "satt_steepest_descent", nperiod));
  if (nperiod < 1 || !adj_s || !svel || !ker_s || !std::isfinite(updatemax))
    return fail(ctx, SATT_E_ARG, "satt_steepest_descent: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)nx * ny, M = (size_t)nperiod * N;
  std::vector<double> sig(nperiod, sigma_2d);
  GeoSmoothBufs B;
  TRY(geo_smooth_setup(ctx, "satt_steepest_descent", lons, lats, nx, ny, nperiod, sig.data(), B));
  DBuf din, dsv, work;
  TRY(upload(ctx, din, adj_s, M * 8));
  TRY(upload(ctx, dsv, svel, M * 8));
  ALLOC(ctx, work, (2 * M + absmax_parts(M) + 1) * 8);
  double *nat = work.as<double>(), *kpm = nat + M, *part = kpm + M;
  B.P.ndata = nperiod;
  B.P.data = nat;
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  launch_pm_to_nat(ctx->stream, N, nperiod, din.as<double>(), nat);
  launch_smooth_geo2(ctx->stream, B.P, nperiod);
  launch_to_period_major(ctx->stream, (int)N, nperiod, B.out.as<double>(), kpm);
  launch_sd_update(ctx->stream, M, kpm, updatemax, dsv.as<double>(), part);
  TRY(geo_timer_stop(ctx, t));
  HIPCHK(ctx, hipMemcpyAsync(ker_s, kpm, M * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(svel, dsv.p, M * 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  return geo_timer_read(ctx, t);
}

namespace {

bool increasing(const double *v, int n) {
  for (int k = 0; k < n; ++k)
    if (!std::isfinite(v[k]) || (k > 0 && !(v[k] > v[k - 1]))) return false;
  return true;
}

// Arguments of inv_grid_iso / inv2fwd_iso: increasing axes, and every model
// point the reference's loops visit (the interior ones for inv_grid_iso)
// inside every set's grid, where the reference would stop with "I cannot
// locate the receiver in the domain" (src/eikon_solver_2d.f90:1604-1607).
int inv_check(satt_ctx *ctx, const char *who, const double *xinv, const double *yinv,
              const double *zinv, int nxi, int nyi, int nzi, int nset, int nx, int ny, int nz,
              const double *vx, const double *vy, const double *vz, bool interior) {
  if (!xinv || !yinv || !zinv || !vx || !vy || !vz || nxi < 2 || nyi < 2 || nzi < 2 ||
      nxi > 1024 || nyi > 1024 || nzi > 1024 ||
      nset < 1 || nx < 2 || ny < 2 || nz < 2 || (size_t)nx * ny * nz > (size_t)INT_MAX ||
      (size_t)nxi * nyi * nzi * nset > (size_t)INT_MAX)
    return fail(ctx, SATT_E_ARG, "%s: bad arguments", who);
  if (!increasing(vx, nx) || !increasing(vy, ny) || !increasing(vz, nz))
    return fail(ctx, SATT_E_ARG, "%s: model axes must increase", who);
  const int lo = interior ? 1 : 0;
  const bool visits = !interior || (nx > 2 && ny > 2 && nz > 2);
  const double *v[3] = {vx, vy, vz}, *inv[3] = {xinv, yinv, zinv};
  const int n[3] = {nx, ny, nz}, ni[3] = {nxi, nyi, nzi};
  for (int s = 0; s < nset; ++s)
    for (int d = 0; d < 3; ++d) {
      const double *X = inv[d] + (size_t)ni[d] * s;
      if (!increasing(X, ni[d]))
        return fail(ctx, SATT_E_ARG, "%s: inversion grid set %d: axis %d must increase", who,
                    s + 1, d + 1);
      if (visits && !(v[d][lo] >= X[0] && v[d][n[d] - 1 - lo] <= X[ni[d] - 1]))
        return fail(ctx, SATT_E_ARG,
                    "%s: a model point lies outside inversion grid set %d (axis %d)", who,
                    s + 1, d + 1);
    }
  return SATT_OK;
}

// The model axes and the inversion grids on the device.
int inv_upload(satt_ctx *ctx, DBuf &buf, InvGrid &G, const double *xinv, const double *yinv,
               const double *zinv, int nxi, int nyi, int nzi, int nset, int nx, int ny, int nz,
               const double *vx, const double *vy, const double *vz) {
  const size_t na = (size_t)nx + ny + nz, ni = (size_t)(nxi + nyi + nzi) * nset;
  std::vector<double> h(na + ni);
  double *q = h.data();
  std::memcpy(q, vx, nx * 8);
  std::memcpy(q + nx, vy, ny * 8);
  std::memcpy(q + nx + ny, vz, nz * 8);
  std::memcpy(q + na, xinv, (size_t)nxi * nset * 8);
  std::memcpy(q + na + (size_t)nxi * nset, yinv, (size_t)nyi * nset * 8);
  std::memcpy(q + na + (size_t)(nxi + nyi) * nset, zinv, (size_t)nzi * nset * 8);
  TRY(upload(ctx, buf, h.data(), h.size() * 8));
  const double *d = buf.as<double>();
  G = InvGrid{nx, ny, nz, nxi, nyi, nzi, nset, d, d + nx, d + nx + ny, d + na,
              d + na + (size_t)nxi * nset, d + na + (size_t)(nxi + nyi) * nset};
  return SATT_OK;
}

}  // namespace

extern "C" int satt_combine_kernels(satt_ctx *ctx, int nperiod, int nx, int ny, int nz,
                                    const double *adj_s, const double *adj_density,
                                    const double *sen_vs, const double *sen_vp,
                                    const double *sen_rho, const double *vs3d,
                                    const double *vp3d, int use_alpha_beta_rho, int rho_scaling,
                                    double *ker, double *ker_density) {
  if (!ctx) return SATT_E_ARG;
  const bool abr = use_alpha_beta_rho != 0, rs = rho_scaling != 0;
  if (nperiod < 1 || nx < 1 || ny < 1 || nz < 1 || !adj_s || !adj_density || !sen_vs ||
      !sen_vp || (!sen_rho && !(abr && rs)) || (!abr && (!vs3d || !vp3d)) || !ker ||
      !ker_density || (size_t)nperiod * nx * ny * nz > (size_t)INT_MAX)
    return fail(ctx, SATT_E_ARG, "satt_combine_kernels: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t n2 = (size_t)nx * ny, n3 = n2 * nz, M2 = (size_t)nperiod * n2;
  const size_t M3 = (size_t)nperiod * n3, nk = (abr ? 3 : 1) * n3;
  DBuf b2, bs, bm, bk;
  ALLOC(ctx, b2, 2 * M2 * 8);
  ALLOC(ctx, bs, 3 * M3 * 8);
  ALLOC(ctx, bk, 2 * nk * 8);
  double *a = b2.as<double>(), *sv = bs.as<double>(), *kk = bk.as<double>();
  auto h2d = [&](double *d, const double *h, size_t n) {
    return hipMemcpyAsync(d, h, n * 8, hipMemcpyHostToDevice, ctx->stream);
  };
  HIPCHK(ctx, h2d(a, adj_s, M2));
  HIPCHK(ctx, h2d(a + M2, adj_density, M2));
  HIPCHK(ctx, h2d(sv, sen_vs, M3));
  HIPCHK(ctx, h2d(sv + M3, sen_vp, M3));
  if (sen_rho) HIPCHK(ctx, h2d(sv + 2 * M3, sen_rho, M3));
  HIPCHK(ctx, h2d(kk, ker, nk));
  HIPCHK(ctx, h2d(kk + nk, ker_density, nk));
  Combine C{nperiod, abr, rs, n2, n3, a, a + M2, sv, sv + M3, sen_rho ? sv + 2 * M3 : nullptr,
            nullptr, nullptr, kk, kk + nk};
  if (!abr) {
    ALLOC(ctx, bm, 2 * n3 * 8);
    HIPCHK(ctx, h2d(bm.as<double>(), vs3d, n3));
    HIPCHK(ctx, h2d(bm.as<double>() + n3, vp3d, n3));
    C.vs3d = bm.as<double>();
    C.vp3d = bm.as<double>() + n3;
  }
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  launch_combine(ctx->stream, C);
  HIPCHK(ctx, hipGetLastError());
  TRY(geo_timer_stop(ctx, t));
  HIPCHK(ctx, hipMemcpyAsync(ker, kk, nk * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ker_density, kk + nk, nk * 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  return geo_timer_read(ctx, t);
}

extern "C" int satt_inv_grid_iso(satt_ctx *ctx, const double *xinv, const double *yinv,
                                 const double *zinv, int nxinv, int nyinv, int nzinv, int nset,
                                 const double *adj, int nx, int ny, int nz, double *gk,
                                 const double *vel_x, const double *vel_y,
                                 const double *vel_z) {
  if (!ctx) return SATT_E_ARG;
  TRY(inv_check(ctx, "satt_inv_grid_iso", xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, nx, ny,
                nz, vel_x, vel_y, vel_z, true));
  if (!adj || !gk) return fail(ctx, SATT_E_ARG, "satt_inv_grid_iso: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  DBuf axes, din, dout;
  InvGrid G;
  TRY(inv_upload(ctx, axes, G, xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, nx, ny, nz, vel_x,
                 vel_y, vel_z));
  const size_t N = (size_t)nx * ny * nz, nb = (size_t)nxinv * nyinv * nzinv * nset;
  TRY(upload(ctx, din, adj, N * 8));
  ALLOC(ctx, dout, nb * 8 + inv_grid_scratch_bytes(G));
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  launch_inv_grid(ctx->stream, G, din.as<double>(), dout.as<double>(),
                  dout.as<char>() + nb * 8);
  HIPCHK(ctx, hipGetLastError());
  TRY(geo_timer_stop(ctx, t));
  HIPCHK(ctx, hipMemcpyAsync(gk, dout.p, nb * 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  return geo_timer_read(ctx, t);
}

extern "C" int satt_inv2fwd_iso(satt_ctx *ctx, const double *xinv, const double *yinv,
                                const double *zinv, int nxinv, int nyinv, int nzinv, int nset,
                                int nx, int ny, int nz, const double *gk, const double *vel_x,
                                const double *vel_y, const double *vel_z, double *update) {
  if (!ctx) return SATT_E_ARG;
  TRY(inv_check(ctx, "satt_inv2fwd_iso", xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, nx, ny,
                nz, vel_x, vel_y, vel_z, false));
  if (!gk || !update) return fail(ctx, SATT_E_ARG, "satt_inv2fwd_iso: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  DBuf axes, din, dout;
  InvGrid G;
  TRY(inv_upload(ctx, axes, G, xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, nx, ny, nz, vel_x,
                 vel_y, vel_z));
  const size_t N = (size_t)nx * ny * nz, nb = (size_t)nxinv * nyinv * nzinv * nset;
  TRY(upload(ctx, din, gk, nb * 8));
  ALLOC(ctx, dout, N * 8);
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  launch_inv2fwd(ctx->stream, G, din.as<double>(), dout.as<double>());
  HIPCHK(ctx, hipGetLastError());
  TRY(geo_timer_stop(ctx, t));
  HIPCHK(ctx, hipMemcpyAsync(update, dout.p, N * 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  return geo_timer_read(ctx, t);
}

extern "C" int satt_kernel_postproc(satt_ctx *ctx, int nker, int nx, int ny, int nz,
                                    const double *vel_x, const double *vel_y,
                                    const double *vel_z, int nset, int nxinv, int nyinv,
                                    int nzinv, const double *xinv, const double *yinv,
                                    const double *zinv, double *ker, const double *ker_density,
                                    double kdensity_coe, double chi0, double weight,
                                    double *gradient_s) {
  if (!ctx) return SATT_E_ARG;
  // every model point inside every set (inv2fwd_iso's condition, which
  // covers inv_grid_iso's interior one)
  TRY(inv_check(ctx, "satt_kernel_postproc", xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, nx,
                ny, nz, vel_x, vel_y, vel_z, false));
  const bool pre = kdensity_coe > 0;
  if (nker < 1 || !ker || !gradient_s || (pre && !ker_density) || !std::isfinite(chi0) ||
      !std::isfinite(weight) || !std::isfinite(kdensity_coe) ||
      (size_t)nker * nx * ny * nz > (size_t)INT_MAX)
    return fail(ctx, SATT_E_ARG, "satt_kernel_postproc: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  DBuf axes, dk, dkd, dg, work;
  InvGrid G;
  TRY(inv_upload(ctx, axes, G, xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, nx, ny, nz, vel_x,
                 vel_y, vel_z));
  const size_t N = (size_t)nx * ny * nz, nb = (size_t)nxinv * nyinv * nzinv * nset;
  TRY(upload(ctx, dk, ker, nker * N * 8));
  if (pre) TRY(upload(ctx, dkd, ker_density, nker * N * 8));
  TRY(upload(ctx, dg, gradient_s, nker * N * 8));
  ALLOC(ctx, work, (3 * N + nb + absmax_parts(N) + 1) * 8 + inv_grid_scratch_bytes(G));
  double *kc = work.as<double>(), *kdc = kc + N, *up = kdc + N, *gk = up + N, *part = gk + nb;
  void *scratch = part + absmax_parts(N) + 1;
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  for (int q = 0; q < nker; ++q) {
    launch_component(ctx->stream, N, nker, q, dk.as<double>(), pre ? dkd.as<double>() : nullptr,
                     kc, pre ? kdc : nullptr);
    if (pre) {
      const double *mx = launch_absmax(ctx->stream, N, kdc, part);
      launch_precond(ctx->stream, N, nker, q, kdc, mx, kdensity_coe, dk.as<double>(), kc);
    }
    launch_inv_grid(ctx->stream, G, kc, gk, scratch);
    launch_inv2fwd(ctx->stream, G, gk, up);
    launch_grad_update(ctx->stream, N, nker, q, up, nset, chi0, weight, dg.as<double>());
  }
  HIPCHK(ctx, hipGetLastError());
  TRY(geo_timer_stop(ctx, t));
  if (pre) HIPCHK(ctx, hipMemcpyAsync(ker, dk.p, nker * N * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(gradient_s, dg.p, nker * N * 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  return geo_timer_read(ctx, t);
}

extern "C" int satt_correct_depth(satt_ctx *ctx, int nperiod, int nx, int ny, int nz,
                                  double *ker_vs, double *ker_vp, double *ker_rho,
                                  const double *dip_angle, const double *depz) {
  if (!ctx) return SATT_E_ARG;
  if (nperiod < 1 || nx < 1 || ny < 1 || nz < 2 || !ker_vs || !ker_vp || !ker_rho ||
      !dip_angle || !depz || (size_t)nperiod * nx * ny * nz > (size_t)INT_MAX)
    return fail(ctx, SATT_E_ARG, "satt_correct_depth: bad arguments");
  if (!increasing(depz, nz) || depz[0] < 0)
    return fail(ctx, SATT_E_ARG, "satt_correct_depth: depths must increase from >= 0");
  const size_t n3 = (size_t)nperiod * nx * ny, M = n3 * nz;
  for (size_t c = 0; c < n3; ++c)
    if (!(dip_angle[c] >= 0 && dip_angle[c] < 90))
      return fail(ctx, SATT_E_ARG, "satt_correct_depth: dip angles must lie in [0, 90)");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  DBuf bk, bo, ba;
  ALLOC(ctx, bk, 3 * M * 8);
  ALLOC(ctx, bo, 3 * M * 8);
  double *k = bk.as<double>();
  double *const ks[3] = {ker_vs, ker_vp, ker_rho};
  for (int q = 0; q < 3; ++q)
    HIPCHK(ctx, hipMemcpyAsync(k + q * M, ks[q], M * 8, hipMemcpyHostToDevice, ctx->stream));
  std::vector<double> h(n3 + nz);
  std::memcpy(h.data(), dip_angle, n3 * 8);
  std::memcpy(h.data() + n3, depz, (size_t)nz * 8);
  TRY(upload(ctx, ba, h.data(), h.size() * 8));
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  launch_correct_depth(ctx->stream, n3, nz, k, bo.as<double>(), ba.as<double>(),
                       ba.as<double>() + n3);
  HIPCHK(ctx, hipGetLastError());
  TRY(geo_timer_stop(ctx, t));
  for (int q = 0; q < 3; ++q)
    HIPCHK(ctx, hipMemcpyAsync(ks[q], bo.as<double>() + q * M, M * 8, hipMemcpyDeviceToHost,
                               ctx->stream));
  TRY(sync_status(ctx));
  return geo_timer_read(ctx, t);
}

extern "C" int satt_fwdsurf3d(satt_ctx *ctx, const double *vs3d, int nx, int ny, int nz,
                              const double *depz, int nperiod, const double *periods, int iwave,
                              int igr, double *svel, int *col_status) {
  if (!ctx) return SATT_E_ARG;
  if (!vs3d || !depz || !periods || !svel || nx < 1 || ny < 1 || nz < 2 ||
      nz > SATT_DISP_MAX_NZ || nperiod < 1 || (iwave != 1 && iwave != 2) || igr < 0 ||
      (size_t)nx * ny > (size_t)INT_MAX)
    return fail(ctx, SATT_E_ARG, "satt_fwdsurf3d: bad arguments");
  for (int k = 0; k < nz; ++k) {
    if (!std::isfinite(depz[k]) || (k > 0 && !(depz[k] > depz[k - 1])))
      return fail(ctx, SATT_E_ARG, "satt_fwdsurf3d: depths must increase");
  }
  for (int p = 0; p < nperiod; ++p)
    if (!(periods[p] > 0) || !std::isfinite(periods[p]))
      return fail(ctx, SATT_E_ARG, "satt_fwdsurf3d: periods must be positive");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int ncol = nx * ny;
  const size_t nv = (size_t)ncol * nz, no = (size_t)ncol * nperiod;
  DBuf dvs, dsmall, dout;
  TRY(upload(ctx, dvs, vs3d, nv * 8));
  std::vector<double> hs((size_t)nz + nperiod);
  std::memcpy(hs.data(), depz, nz * 8);
  std::memcpy(hs.data() + nz, periods, nperiod * 8);
  TRY(upload(ctx, dsmall, hs.data(), hs.size() * 8));
  ALLOC(ctx, dout, no * 8 + (size_t)ncol * sizeof(int));
  DispJob J{};
  J.ncol = ncol;
  J.nz = nz;
  J.kmax = nperiod;
  J.iwave = iwave;
  J.igr = igr;
  J.vs = dvs.as<double>();
  J.depz = dsmall.as<double>();
  J.periods = dsmall.as<double>() + nz;
  J.out = dout.as<double>();
  J.status = reinterpret_cast<int *>(dout.as<double>() + no);
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  launch_fwdsurf(ctx->stream, J);
  TRY(geo_timer_stop(ctx, t));
  std::vector<int> st(ncol);
  HIPCHK(ctx, hipMemcpyAsync(svel, J.out, no * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(st.data(), J.status, (size_t)ncol * sizeof(int),
                             hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  TRY(geo_timer_read(ctx, t));
  if (col_status) std::memcpy(col_status, st.data(), (size_t)ncol * sizeof(int));
  for (int c = 0; c < ncol; ++c)
    if (st[c] != SATT_OK)
      return fail(ctx, SATT_E_DISPERSION,
                  "satt_fwdsurf3d: no root of the period equation in column (%d, %d)",
                  c % nx, c / nx);
  return SATT_OK;
}

extern "C" int satt_depthkernel(satt_ctx *ctx, const double *vs3d, int nx, int ny, int nz,
                                const double *depz, int nperiod, const double *periods,
                                int iwave, int igr, double *sen_vs, double *sen_vp,
                                double *sen_rho, double *cphase, int *col_status) {
  if (!ctx) return SATT_E_ARG;
  if (!vs3d || !depz || !periods || !sen_vs || !sen_vp || !sen_rho || nx < 1 || ny < 1 ||
      nz < 2 || nz + 1 > SATT_DISP_MAX_NZ || nperiod < 1 ||
      (size_t)nx * ny > (size_t)INT_MAX / (size_t)nperiod)
    return fail(ctx, SATT_E_ARG, "satt_depthkernel: bad arguments");
  if (iwave != 2 || igr != 0)
    return fail(ctx, SATT_E_ARG,
                "satt_depthkernel: only the Rayleigh phase-velocity kernels are built");
  for (int k = 0; k < nz; ++k)
    if (!std::isfinite(depz[k]) || (k > 0 && !(depz[k] > depz[k - 1])))
      return fail(ctx, SATT_E_ARG, "satt_depthkernel: depths must increase");
  for (int p = 0; p < nperiod; ++p)
    if (!(periods[p] > 0) || !std::isfinite(periods[p]))
      return fail(ctx, SATT_E_ARG, "satt_depthkernel: periods must be positive");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int ncol = nx * ny;
  const size_t nv = (size_t)ncol * nz, no = (size_t)ncol * nperiod, ns = no * nz;
  DBuf dvs, dsmall, dcp, dsen, dwork;
  TRY(upload(ctx, dvs, vs3d, nv * 8));
  std::vector<double> hs((size_t)nz + nperiod);
  std::memcpy(hs.data(), depz, nz * 8);
  std::memcpy(hs.data() + nz, periods, nperiod * 8);
  TRY(upload(ctx, dsmall, hs.data(), hs.size() * 8));
  ALLOC(ctx, dcp, no * 8 + (size_t)ncol * sizeof(int));
  ALLOC(ctx, dsen, 3 * ns * 8);
  ALLOC(ctx, dwork, depth_work_doubles(no, nz + 1) * 8);
  // the phase velocities on the refinegrid model (surfdisp96, fwdsurf's
  // kernel with refinegrid's layers), then sregn96 per (column, period)
  DispJob J{};
  J.ncol = ncol;
  J.nz = nz;
  J.kmax = nperiod;
  J.iwave = 2;
  J.igr = 0;
  J.refine = 1;
  J.vs = dvs.as<double>();
  J.depz = dsmall.as<double>();
  J.periods = dsmall.as<double>() + nz;
  J.out = dcp.as<double>();
  J.status = reinterpret_cast<int *>(dcp.as<double>() + no);
  DepthJob K{};
  K.ncol = ncol;
  K.nz = nz;
  K.kmax = nperiod;
  K.vs = J.vs;
  K.depz = J.depz;
  K.periods = J.periods;
  K.cp = J.out;
  K.col_status = J.status;
  K.sen_vs = dsen.as<double>();
  K.sen_vp = K.sen_vs + ns;
  K.sen_rho = K.sen_vp + ns;
  K.work = dwork.as<double>();
  GeoTimer t;
  TRY(geo_timer_start(ctx, t));
  launch_fwdsurf(ctx->stream, J);
  launch_depthkernel(ctx->stream, K);
  TRY(geo_timer_stop(ctx, t));
  std::vector<int> st(ncol);
  HIPCHK(ctx, hipMemcpyAsync(sen_vs, K.sen_vs, ns * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(sen_vp, K.sen_vp, ns * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(sen_rho, K.sen_rho, ns * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (cphase)
    HIPCHK(ctx, hipMemcpyAsync(cphase, J.out, no * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(st.data(), J.status, (size_t)ncol * sizeof(int),
                             hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  TRY(geo_timer_read(ctx, t));
  if (col_status) std::memcpy(col_status, st.data(), (size_t)ncol * sizeof(int));
  for (int c = 0; c < ncol; ++c)
    if (st[c] != SATT_OK)
      return fail(ctx, SATT_E_DISPERSION,
                  "satt_depthkernel: no root of the period equation in column (%d, %d)",
                  c % nx, c / nx);
  return SATT_OK;
}

extern "C" int satt_fwd_setup_fields(satt_ctx *ctx, const double *xx_deg,
                                     const double *yy_deg, int nx, int ny,
                                     const double *spha, const double *sphb,
                                     const double *sphc, const double *fun, double x0_deg,
                                     double y0_deg, double *a, double *b, double *c,
                                     double *T0, double *T0x, double *T0y) {
  if (!ctx) return SATT_E_ARG;
  if (bad_fields(nx, ny, spha, sphb, sphc) || !fun || !a || !b || !c || !T0 || !T0x || !T0y)
    return fail(ctx, SATT_E_ARG, "satt_fwd_setup_fields: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)nx * ny;
  OneSolve o;
  TRY(make_grid(ctx, xx_deg, yy_deg, nx, ny, o.dg));
  TRY(one_solve_fields(ctx, o, spha, sphb, sphc, fun));
  DBuf work;
  ALLOC(ctx, work, 3 * N * 8 + 8);
  double *w = work.as<double>();
  int *d_st = reinterpret_cast<int *>(w + 3 * N);
  HIPCHK(ctx, hipMemsetAsync(d_st, 0, sizeof(int), ctx->stream));
  launch_fwd_setup_fields(ctx->stream, o.dg.g, o.pf.a, o.pf.fun, x0_deg, y0_deg, w, w + N,
                          w + 2 * N, d_st);
  int st = 0;
  double *outs[6] = {a, b, c, T0, T0x, T0y};
  const double *srcs[6] = {o.pf.a, o.pf.b, o.pf.c, w, w + N, w + 2 * N};
  for (int q = 0; q < 6; ++q)
    HIPCHK(ctx, hipMemcpyAsync(outs[q], srcs[q], N * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&st, d_st, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  if (st != SATT_OK) return fail(ctx, st, "satt_fwd_setup_fields: %s", status_text(st));
  return SATT_OK;
}

extern "C" int satt_adj_setup_fields(satt_ctx *ctx, const double *xx_deg,
                                     const double *yy_deg, int nx, int ny,
                                     const double *spha, const double *sphb,
                                     const double *sphc, const double *xrec_deg,
                                     const double *yrec_deg, const double *src_adj, int nr,
                                     double *a, double *b, double *c, double *delta) {
  if (!ctx) return SATT_E_ARG;
  if (bad_fields(nx, ny, spha, sphb, sphc) || !a || !b || !c || !delta || nr < 0 ||
      (nr > 0 && (!xrec_deg || !yrec_deg || !src_adj)))
    return fail(ctx, SATT_E_ARG, "satt_adj_setup_fields: bad arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)nx * ny;
  OneSolve o;
  TRY(make_grid(ctx, xx_deg, yy_deg, nx, ny, o.dg));
  TRY(one_solve_fields(ctx, o, spha, sphb, sphc, nullptr));
  std::vector<double> hr(3 * (size_t)nr + 1);
  if (nr) {
    std::memcpy(hr.data(), xrec_deg, nr * 8);
    std::memcpy(hr.data() + nr, yrec_deg, nr * 8);
    std::memcpy(hr.data() + 2 * nr, src_adj, nr * 8);
  }
  DBuf rec, work;
  TRY(upload(ctx, rec, hr.data(), hr.size() * 8));
  ALLOC(ctx, work, N * 8 + 8);
  double *dd = work.as<double>();
  int *d_st = reinterpret_cast<int *>(dd + N);
  HIPCHK(ctx, hipMemsetAsync(work.p, 0, N * 8 + 8, ctx->stream));
  const double *r = rec.as<double>();
  launch_adj_delta(ctx->stream, o.dg.g, r, r + nr, r + 2 * nr, nr, dd, d_st);
  int st = 0;
  double *outs[4] = {a, b, c, delta};
  const double *srcs[4] = {o.pf.a, o.pf.b, o.pf.c, dd};
  for (int q = 0; q < 4; ++q)
    HIPCHK(ctx, hipMemcpyAsync(outs[q], srcs[q], N * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&st, d_st, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  if (st != SATT_OK) return fail(ctx, st, "satt_adj_setup_fields: %s", status_text(st));
  return SATT_OK;
}

// ------------------------------------------------------------ batched plan

// Iteration counts per item: forward, adjoint, density adjoint (FP64), then
// the same three for the FP32 phase of the mixed mode.
constexpr int kIterInts = 6;

// Stages timed by the plan's events.
enum : int { kStFwdSetup, kStFwd64, kStFwd32, kStAdjSetup, kStAdj64, kStAdj32, kStOther,
             kStCount };

namespace {

// One grid of a plan (one satt_batch): its model, items, receivers and
// outputs on the device, and the host copies that descriptors and lists are
// built from.
struct Part {
  int dim = 2, nx = 0, ny = 0, np = 0, nitems = 0, npath = 0, nrec = 0;
  size_t N = 0;
  DevGrid dg;
  SkewGeom sg{};
  DBuf model;    // period-major svel, m11, m12, m22
  DBuf pfields;  // per-period a, b, c, fun, vel (natural layout)
  DBuf pskew;    // per-period skewed fields
  DBuf pskew32;  // the same in float (mixed precision)
  DBuf items;    // pidx, rec_off, srcx, srcy
  DBuf recs;     // recx, recy, tt, weight, idx_path
  DBuf rec_out;  // tekio, weight_out, rhs_adj, rhs_den
  DBuf item_out; // item_chi, iters (kIterInts), status
  DBuf tt_fwd;
  DBuf acc;      // stacked kernels: natural, then period-major
  PeriodFields pf{};
  Items it{};
  double *tekio = nullptr, *weight_out = nullptr, *rhs_adj = nullptr, *rhs_den = nullptr;
  double *item_chi = nullptr;
  int *iters = nullptr, *status = nullptr;
  std::vector<int> h_pidx, h_rec_off;
  std::vector<double> h_srcx, h_srcy;
};

// One launch group of the batched form: the descriptors of its slots and,
// per part, the stacking lists (the part's slots of each period, in item
// order).
struct Group {
  GroupDims gd{};
  SlotDesc *desc = nullptr;
  std::vector<std::pair<int, int>> slots;  // (part, item)
  std::vector<const int *> per_off, per_slot;  // per part; null: no slot
};

}  // namespace

struct satt_plan {
  satt_ctx *ctx = nullptr;
  int precision = 0, dim = 2, slot_skews = kSlotSkews;
  std::vector<std::unique_ptr<Part>> parts;
  std::vector<Group> groups;
  DBuf strip;    // slot scratch of the largest launch group
  DBuf desc;     // descriptors of every group, then the per-solve ones
  DBuf lists;    // stacking lists of every group
  DBuf fwdw, adjw;            // one-solve-per-launch scratch (one-grid plans)
  SlotDesc *d_solo = nullptr; // one-solve-per-launch descriptors, one per item
  std::vector<hipEvent_t> ev;
  std::vector<int> marks;  // stage of the interval ev[k]..ev[k+1]
  int last_istotable = 0, last_isadj = 0, ran = 0, last_batched = 0, last_group = -1;
  int graph_on = 0;
  hipGraphExec_t gexec[4] = {};  // replay form, per (istotable, isadj)
  int waves = 1;                 // waves per item asked for (each group caps it)
  int depth[kStages] = {2, 2, 2, 2};  // prefetch depth per sweep stage
  void drop_graphs() {
    for (auto &x : gexec)
      if (x) {
        (void)hipGraphExecDestroy(x);
        x = nullptr;
      }
  }
  satt_plan_stats stats{};
  ~satt_plan() {
    for (auto e : ev) (void)hipEventDestroy(e);
    for (auto x : gexec)
      if (x) (void)hipGraphExecDestroy(x);
  }
};

namespace {

size_t fwd_item_bytes(size_t N) { return 6 * N * sizeof(double); }
size_t adj_item_bytes(size_t N, int dim) { return (dim == 3 ? 10 : 9) * N * sizeof(double); }

int check_batch(satt_ctx *ctx, const satt_batch *B, int k) {
  if (!B || (B->dim != 2 && B->dim != 3) || B->nx < 3 || B->ny < 3 || B->nperiod < 1 ||
      B->nitems < 0 || !B->xgrids || !B->ygrids || !B->svel || !B->m11 || !B->m12 || !B->m22 ||
      (B->nitems > 0 && (!B->item_pidx || !B->item_srcx || !B->item_srcy || !B->rec_off)))
    return fail(ctx, SATT_E_ARG, "batch %d: bad arguments", k);
  if (B->ny > SATT_MAX_NY)
    return fail(ctx, SATT_E_ARG, "batch %d: %d rows, more than SATT_MAX_NY (%d)", k, B->ny,
                SATT_MAX_NY);
  const int nrec = B->nitems > 0 ? B->rec_off[B->nitems] : 0;
  if (nrec < 0 || (nrec > 0 && (!B->recx || !B->recy || !B->tt || !B->weight || !B->idx_path)))
    return fail(ctx, SATT_E_ARG, "batch %d: bad receiver gather", k);
  if (B->nitems > 0 && B->rec_off[0] != 0)
    return fail(ctx, SATT_E_ARG, "batch %d: receiver offsets must start at 0", k);
  for (int i = 0; i < B->nitems; ++i) {
    if (B->item_pidx[i] < 0 || B->item_pidx[i] >= B->nperiod)
      return fail(ctx, SATT_E_ARG, "batch %d item %d: period index %d out of range", k, i,
                  B->item_pidx[i]);
    if (B->rec_off[i + 1] < B->rec_off[i])
      return fail(ctx, SATT_E_ARG, "batch %d item %d: receiver offsets decrease", k, i);
  }
  for (int r = 0; r < nrec; ++r)
    if (B->idx_path[r] < 0 || B->idx_path[r] >= B->npath)
      return fail(ctx, SATT_E_ARG, "batch %d receiver %d: path index out of range", k, r);
  return SATT_OK;
}

int part_period_fields(satt_ctx *ctx, Part &Q) {
  const size_t N = Q.N, M = (size_t)Q.np * N;
  const double *m = Q.model.as<double>();
  double *f = Q.pfields.as<double>();
  launch_period_fields(ctx->stream, Q.dg.g, Q.np, m, m + M, m + 2 * M, m + 3 * M, f, f + M,
                       f + 2 * M, f + 3 * M, f + 4 * M);
  Q.pf.a = f;
  Q.pf.b = f + M;
  Q.pf.c = f + 2 * M;
  Q.pf.fun = f + 3 * M;
  Q.pf.vel = f + 4 * M;
  launch_period_skew(ctx->stream, Q.dg.g, Q.sg, Q.np, Q.pf, Q.pskew.as<double>(),
                     Q.pskew32.p ? Q.pskew32.as<float>() : nullptr);
  HIPCHK(ctx, hipGetLastError());
  return SATT_OK;
}

int part_set_model(satt_ctx *ctx, Part &Q, const double *svel, const double *m11,
                   const double *m12, const double *m22) {
  if (!svel || !m11 || !m12 || !m22) return fail(ctx, SATT_E_ARG, "set_model: null model");
  const size_t M = (size_t)Q.np * Q.N;
  double *m = Q.model.as<double>();
  const double *src[4] = {svel, m11, m12, m22};
  for (int q = 0; q < 4; ++q)
    HIPCHK(ctx, hipMemcpyAsync(m + q * M, src[q], M * 8, hipMemcpyHostToDevice, ctx->stream));
  return part_period_fields(ctx, Q);
}

int part_create(satt_ctx *ctx, Part &Q, const satt_batch *B, int precision) {
  Q.dim = B->dim;
  Q.nx = B->nx;
  Q.ny = B->ny;
  Q.np = B->nperiod;
  Q.nitems = B->nitems;
  Q.npath = B->npath;
  Q.nrec = B->nitems > 0 ? B->rec_off[B->nitems] : 0;
  Q.N = (size_t)B->nx * B->ny;
  const size_t N = Q.N, M = (size_t)Q.np * N;
  const int ni = Q.nitems, nrec = Q.nrec;
  TRY(make_grid(ctx, B->xgrids, B->ygrids, B->nx, B->ny, Q.dg));
  Q.sg = make_skew_geom(B->nx, B->ny);
  const size_t pse = (size_t)Q.np * 2 * kPF * Q.sg.skew_elems();
  if (Q.model.alloc(4 * M * 8) != hipSuccess || Q.pfields.alloc(5 * M * 8) != hipSuccess ||
      Q.pskew.alloc(pse * 8) != hipSuccess ||
      (precision == SATT_PREC_MIXED && Q.pskew32.alloc(pse * 4) != hipSuccess))
    return fail(ctx, SATT_E_NOMEM, "model allocation failed");
  TRY(part_set_model(ctx, Q, B->svel, B->m11, B->m12, B->m22));
  {
    std::vector<char> h((size_t)ni * 4 + (size_t)(ni + 1) * 4 + (size_t)ni * 16 + 16);
    char *q = h.data();
    std::memcpy(q, B->item_pidx, (size_t)ni * 4);
    std::memcpy(q + (size_t)ni * 4, B->rec_off, (size_t)(ni + 1) * 4);
    const size_t off_d = ((size_t)ni * 8 + 4 + 7) & ~size_t(7);
    std::memcpy(q + off_d, B->item_srcx, (size_t)ni * 8);
    std::memcpy(q + off_d + (size_t)ni * 8, B->item_srcy, (size_t)ni * 8);
    TRY(upload(ctx, Q.items, h.data(), h.size()));
    char *d = Q.items.as<char>();
    Q.it.pidx = reinterpret_cast<int *>(d);
    Q.it.rec_off = reinterpret_cast<int *>(d + (size_t)ni * 4);
    Q.it.srcx = reinterpret_cast<double *>(d + off_d);
    Q.it.srcy = reinterpret_cast<double *>(d + off_d + (size_t)ni * 8);
  }
  {
    std::vector<double> h(4 * (size_t)nrec + (nrec + 1) / 2 + 1);
    if (nrec) {
      std::memcpy(h.data(), B->recx, nrec * 8);
      std::memcpy(h.data() + nrec, B->recy, nrec * 8);
      std::memcpy(h.data() + 2 * (size_t)nrec, B->tt, nrec * 8);
      std::memcpy(h.data() + 3 * (size_t)nrec, B->weight, nrec * 8);
      std::memcpy(h.data() + 4 * (size_t)nrec, B->idx_path, nrec * 4);
    }
    TRY(upload(ctx, Q.recs, h.data(), h.size() * 8));
    double *d = Q.recs.as<double>();
    Q.it.recx = d;
    Q.it.recy = d + nrec;
    Q.it.tt = d + 2 * (size_t)nrec;
    Q.it.weight = d + 3 * (size_t)nrec;
    Q.it.idx_path = reinterpret_cast<int *>(d + 4 * (size_t)nrec);
  }
  if (Q.rec_out.alloc(4 * (size_t)(nrec + 1) * 8) != hipSuccess ||
      Q.item_out.alloc((size_t)(ni + 1) * (8 + 4 * kIterInts + 4)) != hipSuccess ||
      Q.tt_fwd.alloc((size_t)(Q.npath + 1) * 8) != hipSuccess ||
      Q.acc.alloc((Q.dim == 3 ? 4 : 2) * M * 8) != hipSuccess)
    return fail(ctx, SATT_E_NOMEM, "output allocation failed");
  double *d = Q.rec_out.as<double>();
  Q.tekio = d;
  Q.weight_out = d + (nrec + 1);
  Q.rhs_adj = d + 2 * (size_t)(nrec + 1);
  Q.rhs_den = d + 3 * (size_t)(nrec + 1);
  Q.item_chi = Q.item_out.as<double>();
  Q.iters = reinterpret_cast<int *>(Q.item_chi + ni + 1);
  Q.status = Q.iters + (size_t)kIterInts * (ni + 1);
  Q.h_pidx.assign(B->item_pidx, B->item_pidx + ni);
  Q.h_rec_off.assign(B->rec_off, B->rec_off + ni + 1);
  Q.h_srcx.assign(B->item_srcx, B->item_srcx + ni);
  Q.h_srcy.assign(B->item_srcy, B->item_srcy + ni);
  return SATT_OK;
}

// The descriptor of item k of part Q, with its scratch at `base` (natural
// fields T, Ta, Td, delta after the skewed fields, then the tables).
SlotDesc make_desc(const Part &Q, int k, double *base, int slot_skews, int mixed) {
  SlotDesc D{};
  const size_t N = Q.N, SE = Q.sg.skew_elems();
  const int p = Q.h_pidx[k], r0 = Q.h_rec_off[k];
  D.g = Q.dg.g;
  D.sg = Q.sg;
  D.pskew = Q.pskew.as<double>() + (size_t)p * 2 * kPF * SE;
  D.pskew32 = mixed ? Q.pskew32.as<float>() + (size_t)p * 2 * kPF * SE : nullptr;
  D.fun = Q.pf.fun + (size_t)p * N;
  D.vel = Q.pf.vel + (size_t)p * N;
  D.kern = nullptr;
  D.skew = base;
  double *nat = base + (size_t)slot_skews * SE;
  D.T = nat + kNatT * N;
  D.Ta = nat + kNatTa * N;
  D.Td = nat + kNatTd * N;
  D.delta = nat + kNatDelta * N;
  D.aux = nat + kSlotNat * N;
  D.slot_skews = slot_skews;
  D.aux_len = aux_len(Q.nx, Q.ny);
  D.x0 = Q.h_srcx[k];
  D.y0 = Q.h_srcy[k];
  D.nr = Q.h_rec_off[k + 1] - r0;
  D.recx = Q.it.recx + r0;
  D.recy = Q.it.recy + r0;
  D.tt = Q.it.tt + r0;
  D.weight = Q.it.weight + r0;
  D.idx_path = Q.it.idx_path + r0;
  D.tekio = Q.tekio + r0;
  D.weight_out = Q.weight_out + r0;
  D.rhs_adj = Q.rhs_adj + r0;
  D.rhs_den = Q.rhs_den + r0;
  D.item_chi = Q.item_chi + k;
  D.tt_fwd = Q.tt_fwd.as<double>();
  D.status = Q.status + k;
  D.iters = Q.iters + (size_t)kIterInts * k;
  return D;
}

// XCD-aware slot order of one launch group.  The device deals workgroups
// to its kXcds chiplets round robin (workgroup w to XCD w mod kXcds), and each
// chiplet has its own L2.  Items of one period read the same per-period
// fields (metric, slowness and the one-sided constants: six of the ten
// fields a forward step loads), so the slots are ordered to give every XCD
// a contiguous run of periods: each grid's items (the grids arrive most
// expensive first) are sorted by period and cut into kXcds balanced runs,
// run x goes to XCD x, and slot w takes the next item of XCD w mod kXcds.
// Items that run together on one XCD then stream the same rows of their
// period's fields through its L2.  Every item still has its own slot and
// scratch, so the order changes no result.
void xcd_order(const satt_plan *P, std::vector<std::pair<int, int>> &slots) {
  const size_t n = slots.size();
  if (n < 2 * (size_t)kXcds) return;
  std::vector<std::deque<std::pair<int, int>>> q(kXcds);
  int rot = 0;  // where the next grid's remainder starts, so that it spreads
  for (size_t a = 0; a < n;) {
    size_t b = a;
    while (b < n && slots[b].first == slots[a].first) ++b;
    std::vector<std::pair<int, int>> run(slots.begin() + a, slots.begin() + b);
    const Part &Q = *P->parts[run[0].first];
    std::stable_sort(run.begin(), run.end(),
                     [&](const std::pair<int, int> &u, const std::pair<int, int> &v) {
                       return Q.h_pidx[u.second] < Q.h_pidx[v.second];
                     });
    const size_t m = run.size(), rem = m % kXcds;
    size_t at = 0;
    for (int k = 0; k < kXcds; ++k) {
      const int x = (rot + k) % kXcds;
      const size_t cnt = m / kXcds + ((size_t)k < rem ? 1 : 0);
      for (size_t c = 0; c < cnt; ++c) q[x].push_back(run[at++]);
    }
    rot = (int)((rot + rem) % kXcds);
    a = b;
  }
  for (size_t w = 0; w < n; ++w) {
    int x = (int)(w % kXcds);
    if (q[x].empty())  // the queues differ by a few items at the end
      for (int y = 0; y < kXcds; ++y)
        if (q[y].size() > q[x].size()) x = y;
    slots[w] = q[x].front();
    q[x].pop_front();
  }
}

// Launch groups: the slots in order (one grid: item order; several grids:
// the most expensive grid first, each grid's items in item order), as many
// per group as the scratch budget holds, then each group's slots in the
// XCD-aware order above.
int plan_groups(satt_plan *P) {
  satt_ctx *ctx = P->ctx;
  const bool mixed = P->precision == SATT_PREC_MIXED;
  std::vector<std::pair<int, int>> order;
  for (int q = 0; q < (int)P->parts.size(); ++q)
    for (int k = 0; k < P->parts[q]->nitems; ++k) order.push_back({q, k});
  auto cost = [&](int q) {
    const Part &Q = *P->parts[q];
    return (double)Q.sg.ns * (Q.ny + 63);
  };
  std::stable_sort(order.begin(), order.end(), [&](const std::pair<int, int> &a,
                                                   const std::pair<int, int> &b) {
    return cost(a.first) > cost(b.first);
  });
  size_t freeb = 0, totalb = 0;
  HIPCHK(ctx, hipMemGetInfo(&freeb, &totalb));
  // 3/4 of the free memory, at most 220 GiB: every item of a 4000-solve
  // 401x401 batch in one group on a 288 GB device, in either precision
  const size_t budget = std::min(freeb / 4 * 3, (size_t)220 << 30) / 8;
  std::vector<size_t> sd(P->parts.size());
  for (size_t q = 0; q < P->parts.size(); ++q)
    sd[q] = slot_doubles(P->parts[q]->nx, P->parts[q]->ny, P->slot_skews);
  size_t biggest = 0;
  for (size_t a = 0; a < order.size();) {
    Group G;
    size_t used = 0;
    while (a < order.size() && G.slots.size() < 8192 &&
           (G.slots.empty() || used + sd[order[a].first] <= budget)) {
      used += sd[order[a].first];
      G.slots.push_back(order[a++]);
    }
    biggest = std::max(biggest, used);
    xcd_order(P, G.slots);
    P->groups.push_back(std::move(G));
  }
  if (P->strip.alloc(std::max<size_t>(biggest, 1) * 8) != hipSuccess)
    return fail(ctx, SATT_E_NOMEM, "scratch allocation of %zu bytes failed", biggest * 8);
  // descriptors and stacking lists
  std::vector<SlotDesc> hd;
  std::vector<int> hl;
  std::vector<size_t> desc_at, list_at;
  for (Group &G : P->groups) {
    desc_at.push_back(hd.size());
    double *base = P->strip.as<double>();
    GroupDims gd{(int)G.slots.size(), 0, 0, 0};
    for (auto &sl : G.slots) {
      const Part &Q = *P->parts[sl.first];
      hd.push_back(make_desc(Q, sl.second, base, P->slot_skews, mixed));
      base += sd[sl.first];
      gd.max_se = std::max(gd.max_se, Q.sg.skew_elems());
      gd.max_n = std::max(gd.max_n, Q.N);
      gd.max_nxy = std::max(gd.max_nxy, std::max(Q.nx, Q.ny));
      gd.max_ny = std::max(gd.max_ny, Q.ny);
      gd.max_ns = std::max(gd.max_ns, Q.sg.ns);
    }
    for (int st = 0; st < kStages; ++st) {
      gd.waves[st] = team_waves(P->waves, gd.max_ny, gd.max_ns, st);
      gd.depth[st] = P->depth[st];
    }
    G.gd = gd;
    for (size_t q = 0; q < P->parts.size(); ++q) {
      const Part &Q = *P->parts[q];
      std::vector<int> off(Q.np + 1, 0), slots;
      for (size_t s = 0; s < G.slots.size(); ++s)
        if (G.slots[s].first == (int)q) off[Q.h_pidx[G.slots[s].second] + 1]++;
      for (int p = 0; p < Q.np; ++p) off[p + 1] += off[p];
      if (off[Q.np] == 0) {
        list_at.push_back(SIZE_MAX);
        continue;
      }
      slots.assign(off[Q.np], 0);
      std::vector<int> fill(off.begin(), off.end() - 1);
      std::vector<std::pair<int, int>> by_item;  // (item, slot): item order within each period
      for (size_t s = 0; s < G.slots.size(); ++s)
        if (G.slots[s].first == (int)q) by_item.push_back({G.slots[s].second, (int)s});
      std::sort(by_item.begin(), by_item.end());
      for (const auto &is : by_item) slots[fill[Q.h_pidx[is.first]]++] = is.second;
      list_at.push_back(hl.size());
      hl.insert(hl.end(), off.begin(), off.end());
      hl.insert(hl.end(), slots.begin(), slots.end());
    }
  }
  // the one-solve-per-launch descriptors (one-grid plans): scratch fwdw/adjw
  const size_t nsolo = P->parts.size() == 1 ? (size_t)P->parts[0]->nitems : 0;
  if (nsolo) {
    const Part &Q = *P->parts[0];
    double *fw = P->fwdw.as<double>(), *aw = P->adjw.as<double>();
    for (size_t k = 0; k < nsolo; ++k) {
      SlotDesc D = make_desc(Q, (int)k, nullptr, 0, 0);
      D.skew = nullptr;
      D.T = fw;
      D.Ta = aw;
      D.Td = Q.dim == 3 ? aw + Q.N : nullptr;
      hd.push_back(D);
    }
  }
  TRY(upload(ctx, P->desc, hd.data(), hd.size() * sizeof(SlotDesc)));
  TRY(upload(ctx, P->lists, hl.data(), std::max<size_t>(hl.size(), 1) * 4));
  size_t li = 0;
  for (size_t g = 0; g < P->groups.size(); ++g) {
    Group &G = P->groups[g];
    G.desc = P->desc.as<SlotDesc>() + desc_at[g];
    for (size_t q = 0; q < P->parts.size(); ++q, ++li) {
      const size_t at = list_at[li];
      const int np = P->parts[q]->np;
      G.per_off.push_back(at == SIZE_MAX ? nullptr : P->lists.as<int>() + at);
      G.per_slot.push_back(at == SIZE_MAX ? nullptr : P->lists.as<int>() + at + np + 1);
    }
  }
  P->d_solo = nsolo ? P->desc.as<SlotDesc>() + (hd.size() - nsolo) : nullptr;
  return sync_status(ctx);
}

}  // namespace

extern "C" int satt_plan_create_multi(satt_ctx *ctx, int nbatch, const satt_batch *batches,
                                      int precision, satt_plan **out) {
  if (!ctx || !out) return SATT_E_ARG;
  *out = nullptr;
  if (nbatch < 1 || !batches) return fail(ctx, SATT_E_ARG, "satt_plan_create: no batch");
  if (precision != SATT_PREC_FP64 && precision != SATT_PREC_MIXED)
    return fail(ctx, SATT_E_ARG, "satt_plan_create: unknown precision %d", precision);
  for (int k = 0; k < nbatch; ++k) {
    TRY(check_batch(ctx, &batches[k], k));
    const size_t se = make_skew_geom(batches[k].nx, batches[k].ny).skew_elems();
    if (precision == SATT_PREC_MIXED && se > kMaxSkewElems32)
      return fail(ctx, SATT_E_ARG,
                  "satt_plan_create: batch %d: %zu skewed elements per field, more than the mixed "
                  "mode's %zu", k, se, (size_t)kMaxSkewElems32);
    if (batches[k].dim != batches[0].dim)
      return fail(ctx, SATT_E_ARG, "satt_plan_create: batches of one plan share dim");
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto *P = new satt_plan();
  auto bail = [&](int st) { delete P; return st; };
  P->ctx = ctx;
  P->precision = precision;
  P->dim = batches[0].dim;
  P->slot_skews = precision == SATT_PREC_MIXED ? kSlotSkewsMixed : kSlotSkews;
  int st;
  for (int k = 0; k < nbatch; ++k) {
    P->parts.emplace_back(new Part());
    if ((st = part_create(ctx, *P->parts.back(), &batches[k], precision)) < 0) return bail(st);
  }
  if (nbatch == 1) {  // one-solve-per-launch scratch: one slot
    const Part &Q = *P->parts[0];
    if (P->fwdw.alloc(fwd_item_bytes(Q.N)) != hipSuccess ||
        P->adjw.alloc(adj_item_bytes(Q.N, Q.dim)) != hipSuccess)
      return bail(fail(ctx, SATT_E_NOMEM, "scratch allocation failed"));
  }
  if ((st = plan_groups(P)) < 0) return bail(st);
  size_t bytes = P->strip.bytes + P->desc.bytes + P->lists.bytes + P->fwdw.bytes + P->adjw.bytes;
  for (auto &Q : P->parts)
    bytes += Q->dg.buf.bytes + Q->model.bytes + Q->pfields.bytes + Q->pskew.bytes +
             Q->pskew32.bytes + Q->items.bytes + Q->recs.bytes + Q->rec_out.bytes +
             Q->item_out.bytes + Q->tt_fwd.bytes + Q->acc.bytes;
  P->stats.device_bytes = (long long)bytes;
  P->stats.chunk = P->groups.empty() ? 0 : (int)P->groups[0].slots.size();
  *out = P;
  return SATT_OK;
}

extern "C" int satt_plan_create(satt_ctx *ctx, const satt_batch *B, int precision,
                                satt_plan **out) {
  return satt_plan_create_multi(ctx, 1, B, precision, out);
}

extern "C" int satt_plan_set_model_part(satt_plan *P, int part, const double *svel,
                                        const double *m11, const double *m12,
                                        const double *m22) {
  if (!P) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  if (part < 0 || part >= (int)P->parts.size())
    return fail(ctx, SATT_E_ARG, "satt_plan_set_model: no part %d", part);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  TRY(part_set_model(ctx, *P->parts[part], svel, m11, m12, m22));
  return sync_status(ctx);
}

extern "C" int satt_plan_set_model(satt_plan *P, const double *svel, const double *m11,
                                   const double *m12, const double *m22) {
  return satt_plan_set_model_part(P, 0, svel, m11, m12, m22);
}

namespace {

// The launch chain of one run on stream s: the per-part output clears, then
// per launch group the forward (setup, sweeps), the receivers (misfit,
// adjoint sources, tt_fwd), the adjoint(s) and the stacking, then the
// period-major kernels.  mark(stage) follows every launch (the stage
// timing; it records nothing while the chain is being captured).
template <typename Mark>
int enqueue_run(satt_plan *P, hipStream_t s, int istotable, int isadj, bool batched, Mark mark) {
  satt_ctx *ctx = P->ctx;
  const bool mixed = P->precision == SATT_PREC_MIXED;
  for (auto &Qp : P->parts) {
    Part &Q = *Qp;
    launch_zero(s, Q.status, (size_t)(Q.nitems + 1) * 4);
    launch_zero(s, Q.iters, (size_t)(Q.nitems + 1) * 4 * kIterInts);
    launch_zero(s, Q.item_chi, (size_t)(Q.nitems + 1) * 8);
    launch_zero(s, Q.tt_fwd.p, Q.tt_fwd.bytes);
    if (isadj) launch_zero(s, Q.acc.p, Q.acc.bytes);
  }
  HIPCHK(ctx, mark(-1));
  const int nadj = P->dim == 3 ? 2 : 1;  // adjoint, then density adjoint
  const int power = P->dim == 3 ? 3 : 2;
  if (batched) {
    for (const Group &G : P->groups) {
      const GroupDims &gd = G.gd;
      launch_fwd_setup(s, G.desc, gd, mixed);
      HIPCHK(ctx, mark(kStFwdSetup));
      if (mixed) {
        launch_fwd_mixed32(s, G.desc, gd);
        HIPCHK(ctx, mark(kStFwd32));
        launch_fwd_polish(s, G.desc, gd);
      } else {
        launch_fwd_strip(s, G.desc, gd);
      }
      HIPCHK(ctx, mark(kStFwd64));
      launch_receivers(s, G.desc, gd.count, istotable);
      HIPCHK(ctx, mark(kStOther));
      if (!isadj) continue;
      for (int q = 0; q < nadj; ++q) {
        launch_adj_setup(s, G.desc, gd, q, mixed);
        HIPCHK(ctx, mark(kStAdjSetup));
        if (mixed) {
          launch_adj_mixed32(s, G.desc, gd, q);
          HIPCHK(ctx, mark(kStAdj32));
          launch_adj_polish(s, G.desc, gd, q);
        } else {
          launch_adj_strip(s, G.desc, gd, q);
        }
        HIPCHK(ctx, mark(kStAdj64));
      }
      for (size_t pq = 0; pq < P->parts.size(); ++pq) {
        if (!G.per_off[pq]) continue;
        Part &Q = *P->parts[pq];
        double *acc = Q.acc.as<double>();
        for (int q = 0; q < nadj; ++q)
          launch_stack(s, Q.dg.g, Q.pf, Q.np, G.per_off[pq], G.per_slot[pq], G.desc, q,
                       acc + (size_t)q * Q.np * Q.N, power);
      }
      HIPCHK(ctx, mark(kStOther));
    }
  } else {
    Part &Q = *P->parts[0];
    const size_t N = Q.N;
    double *fw = P->fwdw.as<double>(), *aw = P->adjw.as<double>();
    FwdWork fwk{fw, fw + N, fw + 2 * N, fw + 3 * N, fw + 4 * N, fw + 5 * N};
    double *Ta = aw, *Td = aw + N, *coef = aw + 2 * N, *delta = aw + 7 * N, *old = aw + 8 * N;
    if (Q.dim == 2) {  // no density field: shift the scratch down by one field
      coef = aw + N;
      delta = aw + 6 * N;
      old = aw + 7 * N;
    }
    for (int k = 0; k < Q.nitems; ++k) {
      Items it = Q.it;
      it.item0 = k;
      it.count = 1;
      launch_forward(s, Q.dg.g, Q.pf, it, fwk, Q.iters + 0, Q.status, kIterInts);
      HIPCHK(ctx, mark(kStFwd64));
      launch_receivers(s, P->d_solo + k, 1, istotable);
      HIPCHK(ctx, mark(kStOther));
      if (!isadj) continue;
      for (int q = 0; q < nadj; ++q) {
        AdjWork awk{fwk.T, q == 0 ? Ta : Td, coef, delta, old, q == 0 ? Q.rhs_adj : Q.rhs_den};
        launch_adjoint(s, Q.dg.g, Q.pf, it, awk, Q.iters + 1 + q, Q.status, kIterInts, 1);
        HIPCHK(ctx, mark(kStAdj64));
      }
      double *acc = Q.acc.as<double>();
      for (int q = 0; q < nadj; ++q)
        launch_stack_item(s, Q.dg.g, Q.pf, Q.it.pidx, k, q == 0 ? Ta : Td,
                          acc + (size_t)q * Q.np * N, power);
      HIPCHK(ctx, mark(kStOther));
    }
  }
  if (isadj) {
    for (auto &Qp : P->parts) {
      Part &Q = *Qp;
      const size_t M = (size_t)Q.np * Q.N;
      double *acc = Q.acc.as<double>(), *pm = acc + (size_t)nadj * M;
      for (int q = 0; q < nadj; ++q)
        launch_to_period_major(s, (int)Q.N, Q.np, acc + (size_t)q * M, pm + (size_t)q * M);
    }
  }
  HIPCHK(ctx, mark(kStOther));
  return SATT_OK;
}

}  // namespace

extern "C" int satt_plan_run(satt_plan *P, int istotable, int isadj, int launch_form) {
  if (!P) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  if (launch_form != SATT_LAUNCH_BATCHED && launch_form != SATT_LAUNCH_PER_SOLVE)
    return fail(ctx, SATT_E_ARG, "satt_plan_run: unknown launch form");
  const bool mixed = P->precision == SATT_PREC_MIXED;
  const bool batched = launch_form == SATT_LAUNCH_BATCHED;
  if (mixed && !batched)
    return fail(ctx, SATT_E_ARG, "satt_plan_run: the one-solve-per-launch form is FP64 only");
  if (!batched && P->parts.size() != 1)
    return fail(ctx, SATT_E_ARG, "satt_plan_run: the one-solve-per-launch form takes one grid");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  // Events on the kernels' stream: one at the start and one after every
  // launch, each interval tagged with the stage it times.
  if (P->ev.empty()) {
    hipEvent_t e;
    HIPCHK(ctx, hipEventCreate(&e));
    P->ev.push_back(e);
  }
  P->marks.clear();
  auto mark = [&](int stage) -> hipError_t {
    if (stage < 0) return hipEventRecord(P->ev[0], s);  // start of the timed chain
    P->marks.push_back(stage);
    while (P->ev.size() <= P->marks.size()) {  // the pool grows as launches are added
      hipEvent_t e;
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      P->ev.push_back(e);
    }
    return hipEventRecord(P->ev[P->marks.size()], s);
  };
  if (batched && P->graph_on) {
    // replay form: the chain of this (istotable, isadj) pair as one graph,
    // captured on its first run; one timed interval around the launch
    hipGraphExec_t &X = P->gexec[(istotable ? 1 : 0) + (isadj ? 2 : 0)];
    if (!X) {
      HIPCHK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int st = enqueue_run(P, s, istotable, isadj, true,
                                 [](int) -> hipError_t { return hipSuccess; });
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &g);
      if (st < 0) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      HIPCHK(ctx, e);
      const hipError_t ei = hipGraphInstantiate(&X, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) {
        X = nullptr;
        HIPCHK(ctx, ei);
      }
    }
    HIPCHK(ctx, mark(-1));
    HIPCHK(ctx, hipGraphLaunch(X, s));
    HIPCHK(ctx, mark(kStOther));
  } else {
    TRY(enqueue_run(P, s, istotable, isadj, batched, mark));
  }
  HIPCHK(ctx, hipGetLastError());
  P->last_istotable = istotable;
  P->last_isadj = isadj;
  P->last_batched = batched;
  P->ran = 1;
  P->last_group = batched ? (int)P->groups.size() - 1 : -1;
  P->stats.launches = batched ? (int)P->groups.size() : P->parts[0]->nitems;
  P->stats.chunk = batched ? (P->groups.empty() ? 0 : (int)P->groups[0].slots.size()) : 1;
  return SATT_OK;
}

extern "C" int satt_plan_set_waves(satt_plan *P, int waves) {
  if (!P || waves < 0) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  P->waves = waves;
  for (Group &G : P->groups) {
    TuneWaves tw{{waves, waves, waves, waves}};
    if (waves == 0) {  // tuned per launch group (its widest grid and its size) and stage
      const int nx = 64 * G.gd.max_ns;
      const TuneKey key = TuneKey::of(nx, G.gd.max_ny, G.gd.count, P->precision);
      if (!ctx_tune_lookup(ctx, key, tw)) {
        satt_tune_result r{};
        const int st = satt_tune_waves(ctx, nx, G.gd.max_ny, G.gd.count, P->precision, &r);
        if (st < 0) return st;
        for (int s = 0; s < kStages; ++s) {
          tw.w[s] = r.stage_waves[s];
          tw.d[s] = r.stage_depth[s];
        }
      }
    } else {
      for (int s = 0; s < kStages; ++s) tw.d[s] = P->depth[s];
    }
    for (int s = 0; s < kStages; ++s) {
      G.gd.waves[s] = team_waves(tw.w[s], G.gd.max_ny, G.gd.max_ns, s);
      G.gd.depth[s] = tw.d[s];
    }
  }
  P->drop_graphs();  // captured with the old launch shapes
  return SATT_OK;
}

extern "C" int satt_plan_set_stage_waves(satt_plan *P, const int *waves) {
  if (!P || !waves) return SATT_E_ARG;
  for (int s = 0; s < kStages; ++s)
    if (waves[s] < 1) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (Group &G : P->groups)
    for (int s = 0; s < kStages; ++s)
      G.gd.waves[s] = team_waves(waves[s], G.gd.max_ny, G.gd.max_ns, s);
  P->drop_graphs();
  return SATT_OK;
}

extern "C" int satt_plan_set_stage_depth(satt_plan *P, const int *depth) {
  if (!P || !depth) return SATT_E_ARG;
  for (int s = 0; s < kStages; ++s)
    if (depth[s] != 2 && depth[s] != kDeepPrefetch) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int s = 0; s < kStages; ++s) P->depth[s] = depth[s];
  for (Group &G : P->groups)
    for (int s = 0; s < kStages; ++s) G.gd.depth[s] = depth[s];
  P->drop_graphs();
  return SATT_OK;
}

// The launch shape of each sweep stage, measured on the plan's own inputs:
// every team size W at the two-step prefetch, then the deep prefetch at each
// stage's best W and at one wave per item; each stage keeps its fastest
// (W, depth), which the plan is left set to.
int satt::tune_plan_shapes(satt_plan *P, satt_tune_result &r) {
  r = satt_tune_result{};
  double best = 0;
  const int graph = P->graph_on;  // stage times are split in the eager form only
  P->graph_on = 0;
  int st = SATT_OK;
  auto run_stats = [&](const int *w, const int *d, satt_plan_stats &s) {
    st = satt_plan_set_stage_waves(P, w);
    if (st >= 0) st = satt_plan_set_stage_depth(P, d);
    for (int rep = 0; rep < 2 && st >= 0; ++rep) {  // the first run warms up
      st = satt_plan_run(P, 1, 1, SATT_LAUNCH_BATCHED);
      if (st >= 0) st = satt_plan_wait(P);
    }
    if (st >= 0) st = satt_plan_get_stats(P, &s);
    return st >= 0;
  };
  auto stage_times = [](const satt_plan_stats &s, double *sm) {
    sm[0] = s.ms_fwd_fp32; sm[1] = s.ms_fwd_fp64; sm[2] = s.ms_adj_fp32; sm[3] = s.ms_adj_fp64;
  };
  const int d2[kStages] = {2, 2, 2, 2};
  double best_stage[kStages] = {0, 0, 0, 0};
  for (int W = 1; W <= kMaxTeam; ++W) {
    bool fresh = W == 1;  // a W that some group does not cap to a smaller one
    for (const Group &G : P->groups)
      fresh = fresh || team_waves(W, G.gd.max_ny, G.gd.max_ns, kStageFwd32) == W;
    if (!fresh) break;
    const int wW[kStages] = {W, W, W, W};
    satt_plan_stats s{};
    if (!run_stats(wW, d2, s)) break;
    r.ms[W - 1] = s.ms_total;
    if (r.waves == 0 || s.ms_total < best) {
      best = s.ms_total;
      r.waves = W;
    }
    double sm[kStages];
    stage_times(s, sm);
    for (int k = 0; k < kStages; ++k) {
      r.stage_ms[W - 1][k] = sm[k];
      if (r.stage_waves[k] == 0 || sm[k] < best_stage[k]) {
        best_stage[k] = sm[k];
        r.stage_waves[k] = W;
      }
    }
  }
  // the deep prefetch (the full FP64 forward has no deep form)
  int dd[kStages];
  for (int k = 0; k < kStages; ++k) {
    r.stage_depth[k] = 2;
    dd[k] = P->precision == SATT_PREC_FP64 && k == kStageFwd64 ? 2 : kDeepPrefetch;
  }
  const int one[kStages] = {1, 1, 1, 1};
  bool all_one = true;
  for (int k = 0; k < kStages; ++k) all_one = all_one && r.stage_waves[k] == 1;
  satt_plan_stats s{};
  if (st >= 0 && run_stats(r.stage_waves, dd, s)) stage_times(s, r.stage_ms_deep[0]);
  if (st >= 0 && !all_one && run_stats(one, dd, s)) stage_times(s, r.stage_ms_deep[1]);
  if (all_one)
    for (int k = 0; k < kStages; ++k) r.stage_ms_deep[1][k] = r.stage_ms_deep[0][k];
  P->graph_on = graph;
  if (st < 0) return st;
  for (int k = 0; k < kStages; ++k) {
    if (dd[k] == 2) continue;
    if (r.stage_ms_deep[0][k] > 0 && r.stage_ms_deep[0][k] < best_stage[k]) {
      best_stage[k] = r.stage_ms_deep[0][k];
      r.stage_depth[k] = kDeepPrefetch;
    }
    if (r.stage_ms_deep[1][k] > 0 && r.stage_ms_deep[1][k] < best_stage[k]) {
      best_stage[k] = r.stage_ms_deep[1][k];
      r.stage_waves[k] = 1;
      r.stage_depth[k] = kDeepPrefetch;
    }
  }
  TRY(satt_plan_set_stage_waves(P, r.stage_waves));
  return satt_plan_set_stage_depth(P, r.stage_depth);
}

extern "C" int satt_plan_tune_waves(satt_plan *P, satt_tune_result *out) {
  if (!P) return SATT_E_ARG;
  satt_tune_result r{};
  TRY(tune_plan_shapes(P, r));
  if (out) *out = r;
  return SATT_OK;
}

extern "C" int satt_plan_set_graph(satt_plan *P, int on) {
  if (!P) return SATT_E_ARG;
  P->graph_on = on ? 1 : 0;
  return SATT_OK;
}

extern "C" int satt_plan_wait(satt_plan *P) {
  if (!P) return SATT_E_ARG;
  HIPCHK(P->ctx, hipSetDevice(P->ctx->device));
  return sync_status(P->ctx);
}

namespace {

// Timing and iteration statistics of the last run, over every part.
int plan_stats(satt_plan *P) {
  satt_ctx *ctx = P->ctx;
  float ms = 0;
  double st_ms[kStCount] = {};
  for (size_t k = 0; k < P->marks.size(); ++k) {
    HIPCHK(ctx, hipEventElapsedTime(&ms, P->ev[k], P->ev[k + 1]));
    st_ms[P->marks[k]] += ms;
  }
  ms = 0;
  if (!P->marks.empty())
    HIPCHK(ctx, hipEventElapsedTime(&ms, P->ev[0], P->ev[P->marks.size()]));
  P->stats.ms_fwd_fp64 = st_ms[kStFwd64];
  P->stats.ms_fwd_fp32 = st_ms[kStFwd32];
  P->stats.ms_adj_fp64 = st_ms[kStAdj64];
  P->stats.ms_adj_fp32 = st_ms[kStAdj32];
  P->stats.ms_fwd_setup = st_ms[kStFwdSetup];
  P->stats.ms_adj_setup = st_ms[kStAdjSetup];
  P->stats.ms_forward = st_ms[kStFwdSetup] + st_ms[kStFwd64] + st_ms[kStFwd32];
  P->stats.ms_adjoint = st_ms[kStAdjSetup] + st_ms[kStAdj64] + st_ms[kStAdj32];
  P->stats.ms_total = ms;
  long long fs = 0, as = 0, fs32 = 0, as32 = 0;
  for (auto &Qp : P->parts) {
    const Part &Q = *Qp;
    std::vector<int> its((size_t)kIterInts * (Q.nitems + 1));
    HIPCHK(ctx, hipMemcpy(its.data(), Q.iters, its.size() * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < Q.nitems; ++i) {
      const int *q = &its[(size_t)kIterInts * i];
      fs += 4LL * q[0];
      as += 4LL * (q[1] + q[2]);
      fs32 += 4LL * q[3];
      as32 += 4LL * (q[4] + q[5]);
    }
  }
  P->stats.fwd_sweeps = fs;
  P->stats.adj_sweeps = as;
  P->stats.fwd_sweeps32 = fs32;
  P->stats.adj_sweeps32 = as32;
  return SATT_OK;
}

}  // namespace

extern "C" int satt_plan_results_part(satt_plan *P, int part, satt_result *R) {
  if (!P) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  if (part < 0 || part >= (int)P->parts.size())
    return fail(ctx, SATT_E_ARG, "satt_plan_results: no part %d", part);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  TRY(sync_status(ctx));
  if (P->ran) TRY(plan_stats(P));
  const Part &Q = *P->parts[part];
  const int ni = Q.nitems;
  const size_t M = (size_t)Q.np * Q.N;
  std::vector<int> st((size_t)ni + 1), its((size_t)kIterInts * (ni + 1));
  std::vector<double> chi((size_t)ni + 1);
  HIPCHK(ctx, hipMemcpy(st.data(), Q.status, (size_t)(ni + 1) * 4, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(its.data(), Q.iters, its.size() * 4, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(chi.data(), Q.item_chi, (size_t)(ni + 1) * 8, hipMemcpyDeviceToHost));
  int worst = SATT_OK, worst_item = -1;
  for (int i = 0; i < ni; ++i) {
    if (st[i] < 0 && worst >= 0) { worst = st[i]; worst_item = i; }
    else if (st[i] > 0 && worst == SATT_OK) { worst = st[i]; worst_item = i; }
  }
  if (R) {
    if (R->item_status) std::memcpy(R->item_status, st.data(), (size_t)ni * 4);
    for (int i = 0; i < ni; ++i)
      for (int q = 0; q < 3; ++q) {
        if (R->iters) R->iters[3 * i + q] = its[(size_t)kIterInts * i + q];
        if (R->iters_fp32) R->iters_fp32[3 * i + q] = its[(size_t)kIterInts * i + 3 + q];
      }
    if (R->item_chi) std::memcpy(R->item_chi, chi.data(), (size_t)ni * 8);
    if (R->chi) {
      // chi_local = chi_local + ma%chi over items in order (src/tomo2d.f90:193)
      double c = 0;
      for (int i = 0; i < ni; ++i) c = c + chi[i];
      *R->chi = c;
    }
    if (R->tt_fwd && Q.npath > 0) {
      if (P->last_istotable)
        HIPCHK(ctx, hipMemcpy(R->tt_fwd, Q.tt_fwd.p, (size_t)Q.npath * 8, hipMemcpyDeviceToHost));
      else
        std::memset(R->tt_fwd, 0, (size_t)Q.npath * 8);
    }
    if (R->weight_out && Q.nrec > 0) {
      const double *src = P->last_isadj ? Q.weight_out : Q.it.weight;
      HIPCHK(ctx, hipMemcpy(R->weight_out, src, (size_t)Q.nrec * 8, hipMemcpyDeviceToHost));
    }
    if (P->last_isadj) {
      const int nadj = Q.dim == 3 ? 2 : 1;
      const double *pm = Q.acc.as<double>() + (size_t)nadj * M;
      if (R->adj_s) HIPCHK(ctx, hipMemcpy(R->adj_s, pm, M * 8, hipMemcpyDeviceToHost));
      if (R->adj_density && Q.dim == 3)
        HIPCHK(ctx, hipMemcpy(R->adj_density, pm + M, M * 8, hipMemcpyDeviceToHost));
    }
  }
  if (worst < 0)
    return fail(ctx, worst, "item %d: %s", worst_item, status_text(worst));
  if (worst > 0) fail(ctx, worst, "item %d: %s", worst_item, status_text(worst));
  return worst;
}

extern "C" int satt_plan_results(satt_plan *P, satt_result *R) {
  return satt_plan_results_part(P, 0, R);
}

// att_tomo_2d%linesearch (src/tomo2d.f90:253-284): each sub-iteration forms
// the trial model of prepare_fwd_linesearch (src/acqui_2d.f90:173-183) on
// the device, in the plan's own model buffer, and runs the plan forward only
// (forward_simulate(chi, .false., .false.), :271); the base model and the
// kernel stay resident over the sub-iterations.
extern "C" int satt_plan_linesearch(satt_plan *P, const double *svel, const double *ker_s,
                                    double chi0, double maxshrink, int max_sub_niter,
                                    double *updatemax, double *svel_out, double *chi_out,
                                    int *nsub) {
  if (!P) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  if (P->parts.size() != 1)
    return fail(ctx, SATT_E_ARG, "satt_plan_linesearch: the plan must hold one grid");
  if (!svel || !ker_s || !updatemax || !svel_out || !chi_out || !nsub ||
      !std::isfinite(*updatemax) || !std::isfinite(maxshrink) || std::isnan(chi0))
    return fail(ctx, SATT_E_ARG, "satt_plan_linesearch: bad arguments");
  Part &Q = *P->parts[0];
  const size_t M = (size_t)Q.np * Q.N;
  *nsub = 0;
  *chi_out = chi0;
  if (max_sub_niter < 1) {  // the loop does not run: the model stays svel
    std::memcpy(svel_out, svel, M * 8);
    return SATT_OK;
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  DBuf work;
  ALLOC(ctx, work, (2 * M + absmax_parts(M) + 1) * 8);
  double *base = work.as<double>(), *ker = base + M, *part = ker + M;
  HIPCHK(ctx, hipMemcpyAsync(base, svel, M * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ker, ker_s, M * 8, hipMemcpyHostToDevice, ctx->stream));
  const double *mx = launch_absmax(ctx->stream, M, ker, part);
  double mxh = 0;
  HIPCHK(ctx, hipMemcpyAsync(&mxh, mx, 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  if (!(mxh > 0) || !std::isfinite(mxh))
    return fail(ctx, SATT_E_ARG, "satt_plan_linesearch: ker_s is zero or not finite");
  int warn = SATT_OK;
  for (int sit = 1; sit <= max_sub_niter; ++sit) {
    launch_ls_model(ctx->stream, M, base, ker, mx, *updatemax, Q.model.as<double>());
    HIPCHK(ctx, hipGetLastError());
    TRY(part_period_fields(ctx, Q));
    TRY(satt_plan_run(P, 0, 0, SATT_LAUNCH_BATCHED));
    double chi = 0;
    satt_result R{};
    R.chi = &chi;
    const int st = satt_plan_results_part(P, 0, &R);
    if (st < 0) return st;
    if (st > 0) warn = st;
    *nsub = sit;
    *chi_out = chi;
    if (chi > chi0)
      *updatemax = *updatemax * maxshrink;  // :275-278
    else
      break;  // :279-282
  }
  // acqui%svel = acqui%ag%svel (:285-287): the last model tried
  HIPCHK(ctx, hipMemcpyAsync(svel_out, Q.model.p, M * 8, hipMemcpyDeviceToHost, ctx->stream));
  TRY(sync_status(ctx));
  return warn;
}

extern "C" int satt_plan_item_field(satt_plan *P, int item, int which, double *out) {
  if (!P) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  const Part &Q = *P->parts[0];
  if (!out || which < SATT_FIELD_T || which > SATT_FIELD_TD ||
      (which == SATT_FIELD_TD && Q.dim != 3))
    return fail(ctx, SATT_E_ARG, "satt_plan_item_field: bad field");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  TRY(sync_status(ctx));
  const double *src = nullptr;
  if (P->ran && P->last_batched && P->last_group >= 0) {
    const Group &G = P->groups[P->last_group];
    for (size_t s = 0; s < G.slots.size(); ++s)
      if (G.slots[s].first == 0 && G.slots[s].second == item) {
        SlotDesc D;
        HIPCHK(ctx, hipMemcpy(&D, G.desc + s, sizeof D, hipMemcpyDeviceToHost));
        src = which == SATT_FIELD_T ? D.T : which == SATT_FIELD_TA ? D.Ta : D.Td;
      }
  } else if (P->ran && !P->last_batched && item == Q.nitems - 1) {
    const double *fw = P->fwdw.as<double>(), *aw = P->adjw.as<double>();
    src = which == SATT_FIELD_T ? fw : which == SATT_FIELD_TA ? aw : aw + Q.N;
  }
  if (!src) return fail(ctx, SATT_E_ARG, "satt_plan_item_field: item %d is not resident", item);
  HIPCHK(ctx, hipMemcpy(out, src, Q.N * 8, hipMemcpyDeviceToHost));
  return SATT_OK;
}

extern "C" int satt_plan_get_stats(satt_plan *P, satt_plan_stats *S) {
  if (!P || !S) return SATT_E_ARG;
  if (P->ran) {
    HIPCHK(P->ctx, hipSetDevice(P->ctx->device));
    TRY(sync_status(P->ctx));
    TRY(plan_stats(P));
  }
  *S = P->stats;
  return SATT_OK;
}

extern "C" void satt_plan_destroy(satt_plan *P) {
  if (!P) return;
  (void)hipSetDevice(P->ctx->device);
  (void)hipStreamSynchronize(P->ctx->stream);
  delete P;
}

extern "C" void satt_forward_simulate_release(satt_ctx *ctx) {
  if (!ctx || !ctx->fs_plan) return;
  satt_plan_destroy(ctx->fs_plan);
  ctx->fs_plan = nullptr;
  ctx->fs_key = FsKey{};
}

extern "C" int satt_forward_simulate(satt_ctx *ctx, const satt_batch *B, int istotable,
                                     int isadj, int precision, satt_result *R) {
  if (!ctx) return SATT_E_ARG;
  TRY(check_batch(ctx, B, 0));
  if (precision != SATT_PREC_FP64 && precision != SATT_PREC_MIXED)
    return fail(ctx, SATT_E_ARG, "satt_forward_simulate: unknown precision %d", precision);
  FsKey key = FsKey::of(B, precision);
  int st;
  if (ctx->fs_plan && ctx->fs_key == key) {
    // the same gathers: only the model is new
    TRY(satt_plan_set_model(ctx->fs_plan, B->svel, B->m11, B->m12, B->m22));
  } else {
    satt_forward_simulate_release(ctx);  // the old plan's memory goes first
    satt_plan *P = nullptr;
    TRY(satt_plan_create(ctx, B, precision, &P));
    // the team size of each sweep stage, tuned once on the plan's own inputs
    satt_tune_result tr{};
    if ((st = satt_plan_tune_waves(P, &tr)) < 0) {
      satt_plan_destroy(P);
      return st;
    }
    ctx->fs_plan = P;
    ctx->fs_key = std::move(key);
  }
  st = satt_plan_run(ctx->fs_plan, istotable, isadj, SATT_LAUNCH_BATCHED);
  if (st >= 0) st = satt_plan_results(ctx->fs_plan, R);
  return st;
}
