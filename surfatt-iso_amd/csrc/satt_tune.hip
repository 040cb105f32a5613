// Autotuner of the sweep kernels' launch shape: the team size W (waves per
// item) and the prefetch depth, for a grid shape and a launch-group size.
//
// One wave per item is the throughput form when the items fill the device;
// a team of W waves cuts an item's latency on grids of several strips, at
// the price of W wave slots and some waiting in the pipeline's fill.  Where
// the balance lies depends on the grid (strips, rows), the number of items
// launched together and the precision, so it is measured: a synthetic batch
// of that shape runs once per candidate W, then with the deep prefetch (which
// pays where few waves share a SIMD), and each stage's fastest shape is kept
// in the context's table, which plans (satt_plan_set_waves(plan, 0)) and
// servers (waves 0) consult.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/satt.h"
#include "satt_internal.h"
#include "satt_strip.h"

using namespace satt;

namespace {

struct Lcg {
  uint64_t s;
  double next() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) / 9007199254740992.0;
  }
};

}  // namespace

extern "C" int satt_tune_waves(satt_ctx *ctx, int nx, int ny, int nitems, int precision,
                               satt_tune_result *out) {
  if (!ctx || nx < 8 || ny < 8 || nitems < 1 || nitems > 65536 ||
      (precision != SATT_PREC_FP64 && precision != SATT_PREC_MIXED))
    return SATT_E_ARG;
  // synthetic batch: one period, a smooth +-5% velocity, flat metric, 8
  // receivers per item, sources and receivers inside the middle 80%
  const double d = 0.01;
  const size_t N = (size_t)nx * ny;
  std::vector<double> xg(nx), yg(ny), svel(N), m11(N, 1.0), m12(N, 0.0), m22(N, 1.0);
  for (int i = 0; i < nx; ++i) xg[i] = 100.0 + d * i;
  for (int j = 0; j < ny; ++j) yg[j] = 20.0 + d * j;
  const double pi = 3.14159265358979323846;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      svel[i + (size_t)nx * j] = 3.0 * (1 + 0.05 * std::sin(2 * pi * i / nx) *
                                                std::sin(2 * pi * j / ny));
  const int nr = 8;
  std::vector<int> pidx(nitems, 0), off(nitems + 1), path((size_t)nitems * nr);
  std::vector<double> sx(nitems), sy(nitems), rx((size_t)nitems * nr), ry((size_t)nitems * nr),
      tt((size_t)nitems * nr, 0.0), w((size_t)nitems * nr, 1.0);
  Lcg g{0x5a77u + (uint64_t)nx * 131 + (uint64_t)ny};
  auto px = [&]() { return xg[0] + d * (nx - 1) * (0.1 + 0.8 * g.next()); };
  auto py = [&]() { return yg[0] + d * (ny - 1) * (0.1 + 0.8 * g.next()); };
  for (int k = 0; k < nitems; ++k) {
    sx[k] = px();
    sy[k] = py();
    off[k] = k * nr;
    for (int q = 0; q < nr; ++q) {
      rx[(size_t)k * nr + q] = px();
      ry[(size_t)k * nr + q] = py();
      path[(size_t)k * nr + q] = k * nr + q;
    }
  }
  off[nitems] = nitems * nr;
  satt_batch B{};
  B.dim = 2;
  B.nx = nx;
  B.ny = ny;
  B.nperiod = 1;
  B.xgrids = xg.data();
  B.ygrids = yg.data();
  B.svel = svel.data();
  B.m11 = m11.data();
  B.m12 = m12.data();
  B.m22 = m22.data();
  B.nitems = nitems;
  B.item_pidx = pidx.data();
  B.item_srcx = sx.data();
  B.item_srcy = sy.data();
  B.rec_off = off.data();
  B.recx = rx.data();
  B.recy = ry.data();
  B.tt = tt.data();
  B.weight = w.data();
  B.idx_path = path.data();
  B.npath = nitems * nr;
  satt_plan *P = nullptr;
  int st = satt_plan_create(ctx, &B, precision, &P);
  if (st < 0) return st;
  satt_tune_result r{};
  st = tune_plan_shapes(P, r);
  satt_plan_destroy(P);
  if (st < 0) return st;
  TuneWaves tw;
  for (int k = 0; k < kStages; ++k) {
    tw.w[k] = r.stage_waves[k];
    tw.d[k] = r.stage_depth[k];
  }
  ctx_tune_store(ctx, TuneKey::of(nx, ny, nitems, precision), tw);
  if (out) *out = r;
  return SATT_OK;
}
