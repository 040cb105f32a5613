// C-ABI of libsatt.so (declared in include/satt.h).
//
// The host side keeps the reference's data shapes: fields (nx,ny) column-major,
// model arrays (nperiod,nx,ny) period-major, the isrcs item order and a CSR
// receiver gather.  Device buffers belong to a context (per-solve calls) or to
// a plan (batched calls); nothing the caller passes is retained.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/satt.h"
#include "satt_internal.h"
#include "satt_strip.h"

using namespace satt;

struct satt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
};

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

struct satt_plan {
  satt_ctx *ctx = nullptr;
  int dim = 2, nx = 0, ny = 0, np = 0, nitems = 0, npath = 0, nrec = 0, precision = 0;
  size_t N = 0;
  int chunk = 1, nchunks = 0;
  DevGrid dg;
  DBuf model;   // period-major svel, m11, m12, m22
  DBuf pfields; // per-period a, b, c, fun, vel (natural layout)
  DBuf items;   // pidx, rec_off, srcx, srcy
  DBuf recs;    // recx, recy, tt, weight, idx_path
  DBuf rec_out; // tekio, weight_out, rhs_adj, rhs_den
  DBuf item_out;// item_chi, iters(3), status
  DBuf tt_fwd;
  DBuf acc;     // stacked kernels, natural then period-major staging
  DBuf lists;   // per chunk: per_off (np+1), per_slot (chunk)
  DBuf fwdw, adjw;  // one-solve-per-launch scratch (one slot)
  DBuf pskew;   // per-period skewed fields (strip pipeline)
  DBuf pskew32; // the same rounded to float (mixed precision)
  DBuf strip;   // per-slot strip scratch: skew fields, natural fields, tables
  int slot_skews = kSlotSkews, aux_len = 0;
  SkewGeom sg{};
  PeriodFields pf{};
  Items it{};
  double *tekio = nullptr, *weight_out = nullptr, *rhs_adj = nullptr, *rhs_den = nullptr;
  double *item_chi = nullptr;
  int *iters = nullptr, *status = nullptr;
  std::vector<int> h_per_off, h_per_slot;  // host copies of the chunk lists
  int *d_per_off = nullptr, *d_per_slot = nullptr;
  std::vector<hipEvent_t> ev;
  std::vector<int> marks;  // stage of the interval ev[k]..ev[k+1]
  int last_istotable = 0, last_isadj = 0, ran = 0;
  int last_item0 = 0, last_count = 0, last_batched = 0;  // resident slots
  satt_plan_stats stats{};
  ~satt_plan() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

namespace {

size_t fwd_item_bytes(size_t N) { return 6 * N * sizeof(double); }
size_t adj_item_bytes(size_t N, int dim) { return (dim == 3 ? 10 : 9) * N * sizeof(double); }
// strip scratch of one slot: its skewed fields, kSlotNat natural fields and
// the forward tables
// fields (see satt_strip.h)
size_t strip_slot_doubles(size_t N, const SkewGeom &sg, int slot_skews, int aux) {
  return (size_t)slot_skews * sg.skew_elems() + (size_t)kSlotNat * N + (size_t)aux;
}

int plan_period_fields(satt_plan *P) {
  satt_ctx *ctx = P->ctx;
  const size_t N = P->N;
  const double *m = P->model.as<double>();
  const size_t M = (size_t)P->np * N;
  double *f = P->pfields.as<double>();
  launch_period_fields(ctx->stream, P->dg.g, P->np, m, m + M, m + 2 * M, m + 3 * M, f, f + M,
                       f + 2 * M, f + 3 * M, f + 4 * M);
  P->pf.a = f;
  P->pf.b = f + M;
  P->pf.c = f + 2 * M;
  P->pf.fun = f + 3 * M;
  P->pf.vel = f + 4 * M;
  if (P->pskew.p)
    launch_period_skew(ctx->stream, P->dg.g, P->sg, P->np, P->pf, P->pskew.as<double>(),
                       P->pskew32.p ? P->pskew32.as<float>() : nullptr);
  HIPCHK(ctx, hipGetLastError());
  return SATT_OK;
}

}  // namespace

extern "C" int satt_plan_create(satt_ctx *ctx, const satt_batch *B, int precision,
                                satt_plan **out) {
  if (!ctx || !out) return SATT_E_ARG;
  *out = nullptr;
  if (!B || (B->dim != 2 && B->dim != 3) || B->nx < 3 || B->ny < 3 || B->nperiod < 1 ||
      B->nitems < 0 || !B->svel || !B->m11 || !B->m12 || !B->m22 ||
      (B->nitems > 0 && (!B->item_pidx || !B->item_srcx || !B->item_srcy || !B->rec_off)))
    return fail(ctx, SATT_E_ARG, "satt_plan_create: bad batch");
  if (precision != SATT_PREC_FP64 && precision != SATT_PREC_MIXED)
    return fail(ctx, SATT_E_ARG, "satt_plan_create: unknown precision %d", precision);
  const int nrec = B->nitems > 0 ? B->rec_off[B->nitems] : 0;
  if (nrec < 0 || (nrec > 0 && (!B->recx || !B->recy || !B->tt || !B->weight || !B->idx_path)))
    return fail(ctx, SATT_E_ARG, "satt_plan_create: bad receiver gather");
  for (int i = 0; i < B->nitems; ++i) {
    if (B->item_pidx[i] < 0 || B->item_pidx[i] >= B->nperiod)
      return fail(ctx, SATT_E_ARG, "item %d: period index %d out of range", i, B->item_pidx[i]);
    if (B->rec_off[i + 1] < B->rec_off[i])
      return fail(ctx, SATT_E_ARG, "item %d: receiver offsets decrease", i);
  }
  for (int r = 0; r < nrec; ++r)
    if (B->idx_path[r] < 0 || B->idx_path[r] >= B->npath)
      return fail(ctx, SATT_E_ARG, "receiver %d: path index out of range", r);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto *P = new satt_plan();
  auto bail = [&](int st) { delete P; return st; };
  P->ctx = ctx;
  P->dim = B->dim;
  P->nx = B->nx;
  P->ny = B->ny;
  P->np = B->nperiod;
  P->nitems = B->nitems;
  P->npath = B->npath;
  P->nrec = nrec;
  P->precision = precision;
  P->N = (size_t)B->nx * B->ny;
  const size_t N = P->N, M = (size_t)P->np * N;
  int st;
  if ((st = make_grid(ctx, B->xgrids, B->ygrids, B->nx, B->ny, P->dg)) < 0) return bail(st);
  P->sg = make_skew_geom(B->nx, B->ny);
  const size_t pskew_elems = (size_t)P->np * 2 * kPF * P->sg.skew_elems();
  if (P->model.alloc(4 * M * 8) != hipSuccess || P->pfields.alloc(5 * M * 8) != hipSuccess ||
      P->pskew.alloc(pskew_elems * 8) != hipSuccess ||
      (precision == SATT_PREC_MIXED && P->pskew32.alloc(pskew_elems * 4) != hipSuccess))
    return bail(fail(ctx, SATT_E_NOMEM, "model allocation failed"));
  if ((st = satt_plan_set_model(P, B->svel, B->m11, B->m12, B->m22)) < 0) return bail(st);

  // items and receivers
  const int ni = P->nitems;
  {
    std::vector<char> h((size_t)ni * 4 + (size_t)(ni + 1) * 4 + (size_t)ni * 16 + 16);
    char *q = h.data();
    std::memcpy(q, B->item_pidx, (size_t)ni * 4);
    std::memcpy(q + (size_t)ni * 4, B->rec_off, (size_t)(ni + 1) * 4);
    size_t off_d = ((size_t)ni * 8 + 4 + 7) & ~size_t(7);
    std::memcpy(q + off_d, B->item_srcx, (size_t)ni * 8);
    std::memcpy(q + off_d + (size_t)ni * 8, B->item_srcy, (size_t)ni * 8);
    if ((st = upload(ctx, P->items, h.data(), h.size())) < 0) return bail(st);
    char *d = P->items.as<char>();
    P->it.pidx = reinterpret_cast<int *>(d);
    P->it.rec_off = reinterpret_cast<int *>(d + (size_t)ni * 4);
    P->it.srcx = reinterpret_cast<double *>(d + off_d);
    P->it.srcy = reinterpret_cast<double *>(d + off_d + (size_t)ni * 8);
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
    if ((st = upload(ctx, P->recs, h.data(), h.size() * 8)) < 0) return bail(st);
    double *d = P->recs.as<double>();
    P->it.recx = d;
    P->it.recy = d + nrec;
    P->it.tt = d + 2 * (size_t)nrec;
    P->it.weight = d + 3 * (size_t)nrec;
    P->it.idx_path = reinterpret_cast<int *>(d + 4 * (size_t)nrec);
  }
  if (P->rec_out.alloc(4 * (size_t)(nrec + 1) * 8) != hipSuccess ||
      P->item_out.alloc((size_t)(ni + 1) * (8 + 4 * kIterInts + 4)) != hipSuccess ||
      P->tt_fwd.alloc((size_t)(P->npath + 1) * 8) != hipSuccess ||
      P->acc.alloc((P->dim == 3 ? 4 : 2) * M * 8) != hipSuccess)
    return bail(fail(ctx, SATT_E_NOMEM, "output allocation failed"));
  {
    double *d = P->rec_out.as<double>();
    P->tekio = d;
    P->weight_out = d + (nrec + 1);
    P->rhs_adj = d + 2 * (size_t)(nrec + 1);
    P->rhs_den = d + 3 * (size_t)(nrec + 1);
    P->item_chi = P->item_out.as<double>();
    P->iters = reinterpret_cast<int *>(P->item_chi + ni + 1);
    P->status = P->iters + (size_t)kIterInts * (ni + 1);
  }

  // chunk size from free device memory
  size_t freeb = 0, totalb = 0;
  HIPCHK(ctx, hipMemGetInfo(&freeb, &totalb));
  // one-solve-per-launch scratch: one slot
  if (P->fwdw.alloc(fwd_item_bytes(N)) != hipSuccess ||
      P->adjw.alloc(adj_item_bytes(N, P->dim)) != hipSuccess)
    return bail(fail(ctx, SATT_E_NOMEM, "scratch allocation failed"));
  // batched: as many slots as 3/4 of the free memory holds, at most 220 GiB
  // (every item of a 4000-solve 401x401 batch in one launch on a 288 GB
  // device, in either precision)
  HIPCHK(ctx, hipMemGetInfo(&freeb, &totalb));
  P->slot_skews = precision == SATT_PREC_MIXED ? kSlotSkewsMixed : kSlotSkews;
  P->aux_len = aux_len(B->nx, B->ny);
  const size_t per_item = strip_slot_doubles(N, P->sg, P->slot_skews, P->aux_len) * 8;
  size_t budget = std::min(freeb / 4 * 3, (size_t)220 << 30);
  size_t ch = std::max<size_t>(1, budget / per_item);
  ch = std::min<size_t>(ch, 8192);
  ch = std::min<size_t>(ch, (size_t)std::max(ni, 1));
  P->chunk = (int)ch;
  P->nchunks = ni > 0 ? (ni + P->chunk - 1) / P->chunk : 0;
  if (P->strip.alloc(per_item * ch) != hipSuccess)
    return bail(fail(ctx, SATT_E_NOMEM, "scratch allocation failed (chunk %zu)", ch));

  // per chunk: the chunk's slots of each period, in item order
  P->h_per_off.assign((size_t)P->nchunks * (P->np + 1), 0);
  P->h_per_slot.assign((size_t)std::max(ni, 1), 0);
  for (int cix = 0; cix < P->nchunks; ++cix) {
    const int i0 = cix * P->chunk, i1 = std::min(ni, i0 + P->chunk);
    int *off = &P->h_per_off[(size_t)cix * (P->np + 1)];
    for (int i = i0; i < i1; ++i) off[B->item_pidx[i] + 1]++;
    for (int p = 0; p < P->np; ++p) off[p + 1] += off[p];
    std::vector<int> fill(off, off + P->np);
    for (int i = i0; i < i1; ++i) P->h_per_slot[i0 + fill[B->item_pidx[i]]++] = i - i0;
  }
  {
    const size_t a = P->h_per_off.size() * 4, b = P->h_per_slot.size() * 4;
    if (P->lists.alloc(a + b + 8) != hipSuccess)
      return bail(fail(ctx, SATT_E_NOMEM, "list allocation failed"));
    P->d_per_off = P->lists.as<int>();
    P->d_per_slot = P->d_per_off + P->h_per_off.size();
    if (a) HIPCHK(ctx, hipMemcpyAsync(P->d_per_off, P->h_per_off.data(), a, hipMemcpyHostToDevice, ctx->stream));
    if (b) HIPCHK(ctx, hipMemcpyAsync(P->d_per_slot, P->h_per_slot.data(), b, hipMemcpyHostToDevice, ctx->stream));
  }
  if ((st = sync_status(ctx)) < 0) return bail(st);
  P->stats.chunk = P->chunk;
  P->stats.device_bytes =
      (long long)(P->dg.buf.bytes + P->model.bytes + P->pfields.bytes + P->items.bytes +
                  P->recs.bytes + P->rec_out.bytes + P->item_out.bytes + P->tt_fwd.bytes +
                  P->acc.bytes + P->lists.bytes + P->fwdw.bytes + P->adjw.bytes +
                  P->pskew.bytes + P->pskew32.bytes + P->strip.bytes);
  *out = P;
  return SATT_OK;
}

extern "C" int satt_plan_set_model(satt_plan *P, const double *svel, const double *m11,
                                   const double *m12, const double *m22) {
  if (!P) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  if (!svel || !m11 || !m12 || !m22) return fail(ctx, SATT_E_ARG, "satt_plan_set_model: null");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t M = (size_t)P->np * P->N;
  double *m = P->model.as<double>();
  const double *src[4] = {svel, m11, m12, m22};
  for (int q = 0; q < 4; ++q)
    HIPCHK(ctx, hipMemcpyAsync(m + q * M, src[q], M * 8, hipMemcpyHostToDevice, ctx->stream));
  TRY(plan_period_fields(P));
  return sync_status(ctx);
}

extern "C" int satt_plan_run(satt_plan *P, int istotable, int isadj, int launch_form) {
  if (!P) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  if (launch_form != SATT_LAUNCH_BATCHED && launch_form != SATT_LAUNCH_PER_SOLVE)
    return fail(ctx, SATT_E_ARG, "satt_plan_run: unknown launch form");
  const bool mixed = P->precision == SATT_PREC_MIXED;
  if (mixed && launch_form != SATT_LAUNCH_BATCHED)
    return fail(ctx, SATT_E_ARG, "satt_plan_run: the one-solve-per-launch form is FP64 only");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t N = P->N, M = (size_t)P->np * N;
  const int ni = P->nitems;
  // Events on the kernels' stream: one at the start and one after every
  // launch, each interval tagged with the stage it times.
  const int groups = launch_form == SATT_LAUNCH_BATCHED ? P->nchunks : ni;
  if (P->ev.empty()) {
    hipEvent_t e;
    HIPCHK(ctx, hipEventCreate(&e));
    P->ev.push_back(e);
  }
  P->marks.clear();
  auto mark = [&](int stage) -> hipError_t {
    P->marks.push_back(stage);
    while (P->ev.size() <= P->marks.size()) {  // the pool grows as launches are added
      hipEvent_t e;
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      P->ev.push_back(e);
    }
    return hipEventRecord(P->ev[P->marks.size()], s);
  };
  HIPCHK(ctx, hipMemsetAsync(P->status, 0, (size_t)(ni + 1) * 4, s));
  HIPCHK(ctx, hipMemsetAsync(P->iters, 0, (size_t)(ni + 1) * 4 * kIterInts, s));
  HIPCHK(ctx, hipMemsetAsync(P->item_chi, 0, (size_t)(ni + 1) * 8, s));
  HIPCHK(ctx, hipMemsetAsync(P->tt_fwd.p, 0, P->tt_fwd.bytes, s));
  if (isadj) HIPCHK(ctx, hipMemsetAsync(P->acc.p, 0, P->acc.bytes, s));
  HIPCHK(ctx, hipEventRecord(P->ev[0], s));

  const bool batched = launch_form == SATT_LAUNCH_BATCHED;
  const int cap = batched ? P->chunk : 1;
  // one-solve-per-launch buffers (workgroup wavefront kernels)
  double *fw = P->fwdw.as<double>(), *aw = P->adjw.as<double>();
  FwdWork fwk{fw, fw + N, fw + 2 * N, fw + 3 * N, fw + 4 * N, fw + 5 * N};
  double *Ta = aw, *Td = aw + N, *coef = aw + 2 * N, *delta = aw + 7 * N, *old = aw + 8 * N;
  if (P->dim == 2) {  // no density field: shift scratch down by one slot
    coef = aw + N;
    delta = aw + 6 * N;
    old = aw + 7 * N;
  }
  // batched buffers (strip-pipeline kernels)
  StripWork sw{};
  if (batched) {
    const size_t SE = P->sg.skew_elems();
    double *b0 = P->strip.as<double>();
    double *nat = b0 + (size_t)P->slot_skews * SE * cap;
    sw.skew = b0;
    sw.T = nat + (size_t)kNatT * N * cap;
    sw.Ta = nat + (size_t)kNatTa * N * cap;
    sw.Td = nat + (size_t)kNatTd * N * cap;
    sw.delta = nat + (size_t)kNatDelta * N * cap;
    sw.aux = nat + (size_t)kSlotNat * N * cap;
    sw.aux_len = P->aux_len;
    sw.slot_skews = P->slot_skews;
    Ta = sw.Ta;
    Td = sw.Td;
    fwk.T = sw.T;
  }
  double *acc_s = P->acc.as<double>(), *acc_d = acc_s + M;
  for (int g0 = 0; g0 < groups; ++g0) {
    Items it = P->it;
    const int *per_off, *per_slot;
    if (launch_form == SATT_LAUNCH_BATCHED) {
      it.item0 = g0 * P->chunk;
      it.count = std::min(P->chunk, ni - it.item0);
      per_off = P->d_per_off + (size_t)g0 * (P->np + 1);
      per_slot = P->d_per_slot + it.item0;
    } else {
      it.item0 = g0;
      it.count = 1;
      per_off = nullptr;
      per_slot = nullptr;
    }
    const double *pk = P->pskew.as<double>();
    if (batched) {
      launch_fwd_setup(s, P->dg.g, P->sg, P->pf, it, sw, P->status, mixed);
      HIPCHK(ctx, mark(kStFwdSetup));
    }
    if (batched && mixed) {
      launch_fwd_mixed32(s, P->dg.g, P->sg, P->pskew32.as<float>(), P->pf, it, sw, P->iters + 3,
                         P->status, kIterInts);
      HIPCHK(ctx, mark(kStFwd32));
      launch_fwd_polish(s, P->dg.g, P->sg, pk, P->pf, it, sw, P->iters + 0, P->status, kIterInts);
    } else if (batched) {
      launch_fwd_strip(s, P->dg.g, P->sg, pk, P->pf, it, sw, P->iters + 0, P->status, kIterInts);
    } else {
      launch_forward(s, P->dg.g, P->pf, it, fwk, P->iters + 0, P->status, kIterInts);
    }
    HIPCHK(ctx, mark(kStFwd64));
    RecWork rw{fwk.T, P->tekio, P->weight_out, P->rhs_adj, P->rhs_den, P->item_chi,
               istotable ? P->tt_fwd.as<double>() : nullptr};
    launch_receivers(s, P->dg.g, it, rw, P->status);
    HIPCHK(ctx, mark(kStOther));
    if (isadj) {
      for (int q = 0; q < (P->dim == 3 ? 2 : 1); ++q) {  // adjoint, then density adjoint
        const double *rhs = q == 0 ? P->rhs_adj : P->rhs_den;
        const int nat = q == 0 ? kNatTa : kNatTd;
        if (batched) {
          launch_adj_setup(s, P->dg.g, P->sg, pk, it, sw, rhs, P->iters + 1 + q, P->status,
                           kIterInts, mixed);
          HIPCHK(ctx, mark(kStAdjSetup));
        }
        if (batched && mixed) {
          launch_adj_mixed32(s, P->dg.g, P->sg, it, sw, P->iters + 4 + q, P->status, kIterInts);
          HIPCHK(ctx, mark(kStAdj32));
          launch_adj_polish(s, P->dg.g, P->sg, it, sw, nat, P->iters + 1 + q, P->status,
                            kIterInts, 1);
        } else if (batched) {
          launch_adj_strip(s, P->dg.g, P->sg, it, sw, nat, P->iters + 1 + q, P->status, kIterInts,
                           1);
        } else {
          AdjWork awk{fwk.T, q == 0 ? Ta : Td, coef, delta, old, rhs};
          launch_adjoint(s, P->dg.g, P->pf, it, awk, P->iters + 1 + q, P->status, kIterInts, 1);
        }
        HIPCHK(ctx, mark(kStAdj64));
      }
    }
    if (isadj) {
      const int power = P->dim == 3 ? 3 : 2;
      if (per_off) {
        launch_stack(s, P->dg.g, P->pf, P->np, per_off, per_slot, Ta, acc_s, power);
        if (P->dim == 3) launch_stack(s, P->dg.g, P->pf, P->np, per_off, per_slot, Td, acc_d, power);
      } else {
        launch_stack_item(s, P->dg.g, P->pf, P->it.pidx, g0, Ta, acc_s, power);
        if (P->dim == 3) launch_stack_item(s, P->dg.g, P->pf, P->it.pidx, g0, Td, acc_d, power);
      }
    }
    HIPCHK(ctx, mark(kStOther));
  }
  if (isadj) {
    double *pm = P->acc.as<double>() + (P->dim == 3 ? 2 : 1) * M;
    launch_to_period_major(s, (int)N, P->np, acc_s, pm);
    if (P->dim == 3) launch_to_period_major(s, (int)N, P->np, acc_d, pm + M);
  }
  HIPCHK(ctx, mark(kStOther));
  HIPCHK(ctx, hipGetLastError());
  P->last_istotable = istotable;
  P->last_isadj = isadj;
  P->ran = groups;
  P->last_batched = batched;
  P->last_item0 = groups > 0 ? (batched ? (groups - 1) * P->chunk : ni - 1) : 0;
  P->last_count = groups > 0 ? (batched ? ni - P->last_item0 : 1) : 0;
  P->stats.chunk = cap;
  P->stats.launches = groups;
  return SATT_OK;
}

extern "C" int satt_plan_wait(satt_plan *P) {
  if (!P) return SATT_E_ARG;
  HIPCHK(P->ctx, hipSetDevice(P->ctx->device));
  return sync_status(P->ctx);
}

extern "C" int satt_plan_results(satt_plan *P, satt_result *R) {
  if (!P) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  TRY(sync_status(ctx));
  const int ni = P->nitems;
  const size_t M = (size_t)P->np * P->N;
  // timing: the intervals between consecutive events, by stage
  {
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
  }
  std::vector<int> st((size_t)ni + 1), its((size_t)kIterInts * (ni + 1));
  std::vector<double> chi((size_t)ni + 1);
  HIPCHK(ctx, hipMemcpy(st.data(), P->status, (size_t)(ni + 1) * 4, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(its.data(), P->iters, its.size() * 4, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(chi.data(), P->item_chi, (size_t)(ni + 1) * 8, hipMemcpyDeviceToHost));
  long long fs = 0, as = 0, fs32 = 0, as32 = 0;
  int worst = SATT_OK, worst_item = -1;
  for (int i = 0; i < ni; ++i) {
    const int *q = &its[(size_t)kIterInts * i];
    fs += 4LL * q[0];
    as += 4LL * (q[1] + q[2]);
    fs32 += 4LL * q[3];
    as32 += 4LL * (q[4] + q[5]);
    if (st[i] < 0 && worst >= 0) { worst = st[i]; worst_item = i; }
    else if (st[i] > 0 && worst == SATT_OK) { worst = st[i]; worst_item = i; }
  }
  P->stats.fwd_sweeps = fs;
  P->stats.adj_sweeps = as;
  P->stats.fwd_sweeps32 = fs32;
  P->stats.adj_sweeps32 = as32;
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
    if (R->tt_fwd && P->npath > 0) {
      if (P->last_istotable)
        HIPCHK(ctx, hipMemcpy(R->tt_fwd, P->tt_fwd.p, (size_t)P->npath * 8, hipMemcpyDeviceToHost));
      else
        std::memset(R->tt_fwd, 0, (size_t)P->npath * 8);
    }
    if (R->weight_out && P->nrec > 0) {
      const double *src = P->last_isadj ? P->weight_out : P->it.weight;
      HIPCHK(ctx, hipMemcpy(R->weight_out, src, (size_t)P->nrec * 8, hipMemcpyDeviceToHost));
    }
    if (P->last_isadj) {
      double *pm = P->acc.as<double>() + (P->dim == 3 ? 2 : 1) * M;
      if (R->adj_s) HIPCHK(ctx, hipMemcpy(R->adj_s, pm, M * 8, hipMemcpyDeviceToHost));
      if (R->adj_density && P->dim == 3)
        HIPCHK(ctx, hipMemcpy(R->adj_density, pm + M, M * 8, hipMemcpyDeviceToHost));
    }
  }
  if (worst < 0)
    return fail(ctx, worst, "item %d: %s", worst_item, status_text(worst));
  if (worst > 0) fail(ctx, worst, "item %d: %s", worst_item, status_text(worst));
  return worst;
}

extern "C" int satt_plan_item_field(satt_plan *P, int item, int which, double *out) {
  if (!P) return SATT_E_ARG;
  satt_ctx *ctx = P->ctx;
  if (!out || which < SATT_FIELD_T || which > SATT_FIELD_TD ||
      (which == SATT_FIELD_TD && P->dim != 3))
    return fail(ctx, SATT_E_ARG, "satt_plan_item_field: bad field");
  if (!P->ran || item < P->last_item0 || item >= P->last_item0 + P->last_count)
    return fail(ctx, SATT_E_ARG, "satt_plan_item_field: item %d is not resident", item);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  TRY(sync_status(ctx));
  const size_t N = P->N;
  const double *src;
  if (P->last_batched) {
    const int cap = P->chunk;
    const size_t SE = P->sg.skew_elems();
    const double *nat = P->strip.as<double>() + (size_t)P->slot_skews * SE * cap;
    const int nat_field = which == SATT_FIELD_T ? kNatT : which == SATT_FIELD_TA ? kNatTa : kNatTd;
    src = nat + ((size_t)nat_field * cap + (item - P->last_item0)) * N;
  } else {
    const double *fw = P->fwdw.as<double>(), *aw = P->adjw.as<double>();
    src = which == SATT_FIELD_T ? fw : which == SATT_FIELD_TA ? aw : aw + N;
  }
  HIPCHK(ctx, hipMemcpy(out, src, N * 8, hipMemcpyDeviceToHost));
  return SATT_OK;
}

extern "C" int satt_plan_get_stats(satt_plan *P, satt_plan_stats *S) {
  if (!P || !S) return SATT_E_ARG;
  *S = P->stats;
  return SATT_OK;
}

extern "C" void satt_plan_destroy(satt_plan *P) {
  if (!P) return;
  (void)hipSetDevice(P->ctx->device);
  (void)hipStreamSynchronize(P->ctx->stream);
  delete P;
}

extern "C" int satt_forward_simulate(satt_ctx *ctx, const satt_batch *B, int istotable,
                                     int isadj, int precision, satt_result *R) {
  satt_plan *P = nullptr;
  int st = satt_plan_create(ctx, B, precision, &P);
  if (st < 0) return st;
  st = satt_plan_run(P, istotable, isadj, SATT_LAUNCH_BATCHED);
  if (st >= 0) st = satt_plan_results(P, R);
  satt_plan_destroy(P);
  return st;
}
