// Serving: a request loop over independent fwd+adj solves (include/satt.h,
// satt_server_*).
//
// Requests carry their own grid and model.  A scheduler thread owns every
// HIP call: it admits waiting requests in arrival order while the device pool
// and the pinned staging pool have room and a launch stream is free, runs
// each admitted group as one batched launch sequence (the descriptor-driven
// kernels of the batched plan) on its own stream, and completes requests as
// their group's end event fires.  Groups on different streams overlap on the
// device, so requests that arrive while others run start without waiting for
// them.  Memory comes from pools reserved at creation: no allocation or free
// (each of which can synchronize the device) happens while groups run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/satt.h"
#include "satt_internal.h"
#include "satt_strip.h"

using namespace satt;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, capacity) with coalescing; 256-byte granules.
class Pool {
 public:
  explicit Pool(size_t cap = 0) { reset(cap); }
  void reset(size_t cap) {
    free_.clear();
    if (cap) free_[0] = cap;
  }
  bool alloc(size_t bytes, size_t *off) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<size_t, size_t> free_;  // offset -> length
};

size_t al32(size_t n) { return (n + 31) & ~(size_t)31; }  // doubles, 256-byte steps

// Where each piece of one request lives inside its device block (offsets in
// doubles).  The inputs come first and are uploaded in one copy.
struct Layout {
  size_t in_axes, in_model, in_recs, in_item, in_end;  // input: axes, model, receivers, item record
  size_t sc, pf, psk, psk32, rec_out, kern, slot, total;
};

Layout layout(int nx, int ny, int nr, int slot_skews) {
  Layout L{};
  const size_t N = (size_t)nx * ny, SE = make_skew_geom(nx, ny).skew_elems();
  size_t at = 0;
  L.in_axes = at;  at += al32(2 * (size_t)(nx + ny));    // xx, yy (rad), xd, yd (deg)
  L.in_model = at; at += al32(4 * N);                    // svel, m11, m12, m22
  L.in_recs = at;  at += al32(4 * (size_t)nr + (nr + 1) / 2);  // recx, recy, tt, weight, idx
  L.in_item = at;  at += 32;                             // chi, iters (6 int), status
  L.in_end = at;
  L.sc = at;       at += al32(2 * (size_t)ny);           // sin, cos of the latitudes
  L.pf = at;       at += al32(5 * N);                    // a, b, c, fun, vel
  L.psk = at;      at += al32(2 * (size_t)kPF * SE);
  L.psk32 = at;    at += al32((size_t)kPF * SE);         // 2*kPF*SE floats
  L.rec_out = at;  at += al32(4 * (size_t)nr);           // tekio, weight_out, rhs_adj, rhs_den
  L.kern = at;     at += al32(N);
  L.slot = at;     at += slot_doubles(nx, ny, slot_skews);
  L.total = at;
  return L;
}

struct Req {
  long long ticket = 0;
  satt_request q{};
  std::vector<double> axes, model, recs;  // host copies of the inputs
  satt_response *resp = nullptr;
  Clock::time_point t_submit, t_launch;
  Layout L{};
  size_t dev_off = 0, pin_in = 0, pin_out = 0, out_bytes = 0;
  int cls = 0;  // 1: a large grid (see kLargeSteps)
  bool done = false;
};

// Requests whose sweep takes at least this many wave steps (strips x rows:
// 401x401 takes 3248) run in launch groups of their own, on their own
// streams: a group ends with its slowest item, and the device has few
// hardware queues, so small requests must not wait behind large ones.
constexpr long kLargeSteps = 2000;

struct Batch {
  int stream = 0;
  std::vector<Req *> reqs;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  size_t desc_dev = 0, desc_pin = 0;  // the stream's descriptor area
  bool failed = false;
};

}  // namespace

struct satt_server {
  satt_ctx *ctx = nullptr;
  int device = 0, precision = 0, max_batch = 256, nstreams = 4, slot_skews = kSlotSkews;
  int waves = 0;           // waves per request in the sweep kernels (0: tuned, else 4)
  char *dev = nullptr;     // device pool
  char *pin = nullptr;     // pinned staging pool
  size_t dev_cap = 0, pin_cap = 0;
  // Each stream owns a descriptor area of max_batch descriptors, carved from
  // both pools at creation, so that admitting a group needs only the
  // requests' own blocks; dev_avail/pin_avail are what requests can use.
  size_t desc_stride = 0, desc_dev0 = 0, desc_pin0 = 0, dev_avail = 0, pin_avail = 0;
  Pool dev_pool, pin_pool;
  std::vector<hipStream_t> streams;
  std::vector<bool> stream_busy;
  std::vector<hipEvent_t> spare_events;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> incoming;          // submitted, not yet seen by the scheduler
  std::map<long long, std::unique_ptr<Req>> reqs;
  long long next_ticket = 1, in_device = 0;
  bool stop = false;
  int hip_error = 0;
  satt_server_stats st{};
  std::thread thr;

  void run();
  bool launch(Batch &b);
  void finish(Batch &b);
};

namespace {

bool hip_ok(hipError_t e) { return e == hipSuccess; }

// Grid of a request inside its block.
Grid request_grid(const Req &r, double *blk) {
  Grid g{};
  g.nx = r.q.nx;
  g.ny = r.q.ny;
  const double *ax = r.axes.data();
  g.dx = ax[1] - ax[0];
  g.dy = ax[r.q.nx + 1] - ax[r.q.nx];
  double *d = blk + r.L.in_axes;
  g.xx = d;
  g.yy = d + g.nx;
  g.xd = d + g.nx + g.ny;
  g.yd = d + 2 * g.nx + g.ny;
  g.sy = blk + r.L.sc;
  g.cy = blk + r.L.sc + g.ny;
  return g;
}

}  // namespace

// One launch group: inputs up, per-request fields, the batched kernels,
// outputs down, end event.  Returns false on a HIP error.
bool satt_server::launch(Batch &b) {
  hipStream_t s = streams[b.stream];
  const bool mixed = precision == SATT_PREC_MIXED;
  std::vector<SlotDesc> desc;
  GroupDims gd{(int)b.reqs.size(), 0, 0, 0};
  if (!hip_ok(hipEventRecord(b.ev_start, s))) return false;
  for (Req *r : b.reqs) {
    double *blk = reinterpret_cast<double *>(dev + r->dev_off);
    double *pin_in = reinterpret_cast<double *>(pin + r->pin_in);
    const int nx = r->q.nx, ny = r->q.ny, nr = r->q.nrec;
    const size_t N = (size_t)nx * ny;
    const Layout &L = r->L;
    // stage the inputs: axes, model, receivers, zeroed item record
    std::memcpy(pin_in + L.in_axes, r->axes.data(), r->axes.size() * 8);
    std::memcpy(pin_in + L.in_model, r->model.data(), r->model.size() * 8);
    std::memcpy(pin_in + L.in_recs, r->recs.data(), r->recs.size() * 8);
    std::memset(pin_in + L.in_item, 0, 32 * 8);
    if (!hip_ok(hipMemcpyAsync(blk, pin_in, L.in_end * 8, hipMemcpyHostToDevice, s))) return false;
    const Grid g = request_grid(*r, blk);
    const SkewGeom sg = make_skew_geom(nx, ny);
    launch_grid_trig(s, g, blk + L.sc, blk + L.sc + ny);
    double *mo = blk + L.in_model, *pf = blk + L.pf;
    launch_period_fields(s, g, 1, mo, mo + N, mo + 2 * N, mo + 3 * N, pf, pf + N, pf + 2 * N,
                         pf + 3 * N, pf + 4 * N);
    PeriodFields P{pf, pf + N, pf + 2 * N, pf + 3 * N, pf + 4 * N};
    float *psk32 = mixed ? reinterpret_cast<float *>(blk + L.psk32) : nullptr;
    launch_period_skew(s, g, sg, 1, P, blk + L.psk, psk32);
    // the request's descriptor
    SlotDesc D{};
    D.g = g;
    D.sg = sg;
    D.pskew = blk + L.psk;
    D.pskew32 = psk32;
    D.fun = P.fun;
    D.vel = P.vel;
    D.kern = blk + L.kern;
    D.skew = blk + L.slot;
    double *nat = D.skew + (size_t)slot_skews * sg.skew_elems();
    D.T = nat + kNatT * N;
    D.Ta = nat + kNatTa * N;
    D.Td = nat + kNatTd * N;
    D.delta = nat + kNatDelta * N;
    D.aux = nat + kSlotNat * N;
    D.slot_skews = slot_skews;
    D.aux_len = aux_len(nx, ny);
    D.x0 = r->q.srcx;
    D.y0 = r->q.srcy;
    D.nr = nr;
    double *rc = blk + L.in_recs;
    D.recx = rc;
    D.recy = rc + nr;
    D.tt = rc + 2 * nr;
    D.weight = rc + 3 * nr;
    D.idx_path = reinterpret_cast<const int *>(rc + 4 * nr);
    double *ro = blk + L.rec_out;
    D.tekio = ro;
    D.weight_out = ro + nr;
    D.rhs_adj = ro + 2 * nr;
    D.rhs_den = ro + 3 * nr;
    double *item = blk + L.in_item;
    D.item_chi = item;
    D.iters = reinterpret_cast<int *>(item + 1);
    D.status = reinterpret_cast<int *>(item + 4);
    D.tt_fwd = ro + 2 * nr;  // per-path scatter target, unused here
    desc.push_back(D);
    gd.max_se = std::max(gd.max_se, sg.skew_elems());
    gd.max_n = std::max(gd.max_n, N);
    gd.max_nxy = std::max(gd.max_nxy, std::max(nx, ny));
    gd.max_ny = std::max(gd.max_ny, ny);
    gd.max_ns = std::max(gd.max_ns, sg.ns);
  }
  TuneWaves tw{{waves, waves, waves, waves}};
  if (waves == 0) {  // the team sizes tuned for one request of this shape, if any
    const TuneKey key = TuneKey::of(64 * gd.max_ns, gd.max_ny, 1, precision);
    if (!ctx_tune_lookup(ctx, key, tw)) tw = TuneWaves{{4, 4, 4, 4}};
  }
  for (int st = 0; st < kStages; ++st) {
    gd.waves[st] = team_waves(tw.w[st], gd.max_ny, gd.max_ns, st);
    gd.depth[st] = tw.d[st];
  }
  std::memcpy(pin + b.desc_pin, desc.data(), desc.size() * sizeof(SlotDesc));
  SlotDesc *sd = reinterpret_cast<SlotDesc *>(dev + b.desc_dev);
  if (!hip_ok(hipMemcpyAsync(sd, pin + b.desc_pin, desc.size() * sizeof(SlotDesc),
                             hipMemcpyHostToDevice, s)))
    return false;
  launch_fwd_setup(s, sd, gd, mixed);
  if (mixed) {
    launch_fwd_mixed32(s, sd, gd);
    launch_fwd_polish(s, sd, gd);
  } else {
    launch_fwd_strip(s, sd, gd);
  }
  launch_receivers(s, sd, gd.count, 0);
  launch_adj_setup(s, sd, gd, 0, mixed);
  if (mixed) {
    launch_adj_mixed32(s, sd, gd, 0);
    launch_adj_polish(s, sd, gd, 0);
  } else {
    launch_adj_strip(s, sd, gd, 0);
  }
  launch_item_kernel(s, sd, gd);
  if (!hip_ok(hipGetLastError())) return false;
  for (size_t k = 0; k < b.reqs.size(); ++k) {
    Req *r = b.reqs[k];
    double *blk = reinterpret_cast<double *>(dev + r->dev_off);
    char *po = pin + r->pin_out;
    const size_t N = (size_t)r->q.nx * r->q.ny;
    // tekio (nr) | item record (32) | kernel (N) | T (N, when asked for)
    const Layout &L = r->L;
    // until the copy below lands, the request's record reads as a HIP error
    double *po_item = reinterpret_cast<double *>(po + (size_t)r->q.nrec * 8);
    std::memset(po_item, 0, 32 * 8);
    *reinterpret_cast<int *>(po_item + 4) = SATT_E_HIP;
    if (!hip_ok(hipMemcpyAsync(po, blk + L.rec_out, (size_t)r->q.nrec * 8,
                               hipMemcpyDeviceToHost, s)) ||
        !hip_ok(hipMemcpyAsync(po + (size_t)r->q.nrec * 8, blk + L.in_item, 32 * 8,
                               hipMemcpyDeviceToHost, s)) ||
        !hip_ok(hipMemcpyAsync(po + ((size_t)r->q.nrec + 32) * 8, blk + L.kern, N * 8,
                               hipMemcpyDeviceToHost, s)))
      return false;
    if (r->resp->T &&
        !hip_ok(hipMemcpyAsync(po + ((size_t)r->q.nrec + 32 + N) * 8, desc[k].T, N * 8,
                               hipMemcpyDeviceToHost, s)))
      return false;
  }
  return hip_ok(hipEventRecord(b.ev_end, s));
}

// Results of a finished group into the callers' responses.
void satt_server::finish(Batch &b) {
  float ms = 0;
  const bool timed = !b.failed && hipEventElapsedTime(&ms, b.ev_start, b.ev_end) == hipSuccess;
  const Clock::time_point now = Clock::now();
  std::lock_guard<std::mutex> lk(mu);
  if (timed) st.device_busy_ms += ms;
  for (Req *r : b.reqs) {
    const char *po = pin + r->pin_out;
    const int nr = r->q.nrec;
    const size_t N = (size_t)r->q.nx * r->q.ny;
    const double *item = reinterpret_cast<const double *>(po + (size_t)nr * 8);
    const int *its = reinterpret_cast<const int *>(item + 1);
    const int status = *reinterpret_cast<const int *>(item + 4);
    satt_response *R = r->resp;
    if (R->tt_rec) std::memcpy(R->tt_rec, po, (size_t)nr * 8);
    if (R->kernel) std::memcpy(R->kernel, po + ((size_t)nr + 32) * 8, N * 8);
    if (R->T) std::memcpy(R->T, po + ((size_t)nr + 32 + N) * 8, N * 8);
    R->chi = item[0];
    R->iters[0] = its[0];
    R->iters[1] = its[1];
    R->iters_fp32[0] = its[3];
    R->iters_fp32[1] = its[4];
    R->status = b.failed ? SATT_E_HIP : status;
    R->queue_ms = std::chrono::duration<double, std::milli>(r->t_launch - r->t_submit).count();
    R->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
    dev_pool.release(r->dev_off, r->L.total * 8);
    pin_pool.release(r->pin_in, r->L.in_end * 8);
    pin_pool.release(r->pin_out, r->out_bytes);
    r->done = true;
    r->axes.clear();
    r->model.clear();
    r->recs.clear();
    in_device--;
    st.completed++;
  }
  stream_busy[b.stream] = false;
  spare_events.push_back(b.ev_start);
  spare_events.push_back(b.ev_end);
  cv_done.notify_all();
}

void satt_server::run() {
  (void)hipSetDevice(device);
  std::deque<Req *> waiting[2];  // small, large
  std::vector<Batch> flying;
  // streams [0, nlarge) take large requests, the rest small ones
  const int nlarge = nstreams >= 2 ? nstreams / 2 : 0;
  for (;;) {
    bool progressed = false;
    {
      std::unique_lock<std::mutex> lk(mu);
      while (!incoming.empty()) {
        Req *r = incoming.front();
        waiting[nlarge > 0 ? r->cls : 0].push_back(r);
        incoming.pop_front();
      }
      if (stop && waiting[0].empty() && waiting[1].empty() && flying.empty()) break;
    }
    // completions
    for (size_t k = 0; k < flying.size();) {
      const hipError_t q = hipEventQuery(flying[k].ev_end);
      if (q == hipErrorNotReady) { ++k; continue; }
      if (q != hipSuccess) {
        // the group's outputs were never copied back: every request in it
        // reports the error instead of the staging pool's stale contents
        flying[k].failed = true;
        std::lock_guard<std::mutex> lk(mu);
        hip_error = SATT_E_HIP;
      }
      finish(flying[k]);
      flying.erase(flying.begin() + k);
      progressed = true;
    }
    // admissions: arrival order within a class, while the pools and a stream
    // of the class have room
    for (int c = 0; c < 2; ++c)
    while (!waiting[c].empty()) {
      std::deque<Req *> &wq = waiting[c];
      const int s0 = nlarge == 0 ? 0 : (c == 1 ? 0 : nlarge);
      const int s1 = nlarge == 0 ? nstreams : (c == 1 ? nlarge : nstreams);
      int si = -1;
      for (int k = s0; k < s1; ++k)
        if (!stream_busy[k]) { si = k; break; }
      if (si < 0) break;
      Batch b;
      b.stream = si;
      b.desc_dev = desc_dev0 + (size_t)si * desc_stride;
      b.desc_pin = desc_pin0 + (size_t)si * desc_stride;
      {
        std::lock_guard<std::mutex> lk(mu);
        const size_t most = std::min<size_t>(wq.size(), (size_t)max_batch);
        while (!wq.empty() && b.reqs.size() < most) {
          Req *r = wq.front();
          size_t d, pi, po;
          if (!dev_pool.alloc(r->L.total * 8, &d)) break;
          if (!pin_pool.alloc(r->L.in_end * 8, &pi)) {
            dev_pool.release(d, r->L.total * 8);
            break;
          }
          if (!pin_pool.alloc(r->out_bytes, &po)) {
            dev_pool.release(d, r->L.total * 8);
            pin_pool.release(pi, r->L.in_end * 8);
            break;
          }
          r->dev_off = d;
          r->pin_in = pi;
          r->pin_out = po;
          r->t_launch = Clock::now();
          b.reqs.push_back(r);
          wq.pop_front();
        }
        if (b.reqs.empty()) break;
        in_device += (long long)b.reqs.size();
        st.max_in_flight = std::max(st.max_in_flight, in_device);
        st.launches++;
        stream_busy[si] = true;
        if (spare_events.size() >= 2) {
          b.ev_start = spare_events.back(); spare_events.pop_back();
          b.ev_end = spare_events.back(); spare_events.pop_back();
        }
      }
      if (!b.ev_start) (void)hipEventCreate(&b.ev_start);
      if (!b.ev_end) (void)hipEventCreate(&b.ev_end);
      if (!launch(b)) {
        (void)hipStreamSynchronize(streams[si]);
        {
          std::lock_guard<std::mutex> lk(mu);
          hip_error = SATT_E_HIP;
        }
        b.failed = true;
      }
      flying.push_back(b);
      progressed = true;
    }
    if (!progressed) {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait_for(lk, std::chrono::microseconds(flying.empty() ? 2000 : 100));
    }
  }
}

extern "C" int satt_server_create(satt_ctx *ctx, const satt_server_config *cfg,
                                  satt_server **out) {
  if (!ctx || !out) return SATT_E_ARG;
  *out = nullptr;
  satt_server_config c{};
  if (cfg) c = *cfg;
  if (c.precision != SATT_PREC_FP64 && c.precision != SATT_PREC_MIXED) return SATT_E_ARG;
  auto S = std::make_unique<satt_server>();
  S->ctx = ctx;
  S->device = ctx_device(ctx);
  S->precision = c.precision;
  S->slot_skews = c.precision == SATT_PREC_MIXED ? kSlotSkewsMixed : kSlotSkews;
  S->max_batch = c.max_batch > 0 ? c.max_batch : 256;
  // 6 streams (3 for large grids, 3 for small ones) measured best for the
  // mixed-size request stream with teams: p99 325 ms against 498-572 ms with
  // 4 or 8 streams at 40 requests/s
  S->nstreams = c.streams > 0 ? std::min(c.streams, 16) : 6;
  S->waves = c.waves > 0 ? c.waves : 0;
  if (hipSetDevice(S->device) != hipSuccess) return SATT_E_HIP;
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return SATT_E_HIP;
  S->dev_cap = c.device_bytes > 0 ? (size_t)c.device_bytes : freeb / 2;
  S->dev_cap &= ~(size_t)255;
  // inputs and outputs of requests in flight: 1/8 of the pool, 64 MiB to 4 GiB
  S->pin_cap = std::min(std::max<size_t>(S->dev_cap / 8, (size_t)64 << 20), (size_t)4 << 30) &
               ~(size_t)255;
  if (hipMalloc(&S->dev, S->dev_cap) != hipSuccess) return SATT_E_NOMEM;
  if (hipHostMalloc(&S->pin, S->pin_cap, hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(S->dev);
    return SATT_E_NOMEM;
  }
  S->dev_pool.reset(S->dev_cap);
  S->pin_pool.reset(S->pin_cap);
  S->desc_stride = ((size_t)S->max_batch * sizeof(SlotDesc) + 255) & ~(size_t)255;
  const size_t desc_all = (size_t)S->nstreams * S->desc_stride;
  if (desc_all * 4 > S->dev_cap || desc_all * 4 > S->pin_cap ||
      !S->dev_pool.alloc(desc_all, &S->desc_dev0) || !S->pin_pool.alloc(desc_all, &S->desc_pin0)) {
    (void)hipFree(S->dev);
    (void)hipHostFree(S->pin);
    return SATT_E_ARG;  // a pool too small for its own descriptors
  }
  S->dev_avail = S->dev_cap - desc_all;
  S->pin_avail = S->pin_cap - desc_all;
  S->streams.resize(S->nstreams);
  S->stream_busy.assign(S->nstreams, false);
  for (auto &st : S->streams)
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return SATT_E_HIP;
  satt_server *raw = S.release();
  raw->thr = std::thread([raw] { raw->run(); });
  *out = raw;
  return SATT_OK;
}

extern "C" int satt_server_submit(satt_server *S, const satt_request *q, satt_response *resp,
                                  long long *ticket) {
  if (!S || !q || !resp) return SATT_E_ARG;
  const bool no_metric = !q->m11 && !q->m12 && !q->m22;
  if (S->precision == SATT_PREC_MIXED && q->nx >= 3 && q->ny >= 3 &&
      make_skew_geom(q->nx, q->ny).skew_elems() > kMaxSkewElems32)
    return SATT_E_ARG;
  if (q->nx < 3 || q->ny < 3 || q->ny > SATT_MAX_NY || !q->xgrids || !q->ygrids || !q->svel || q->nrec < 0 ||
      (!no_metric && (!q->m11 || !q->m12 || !q->m22)) ||
      (q->nrec > 0 && (!q->recx || !q->recy || !q->tt || !q->weight)))
    return SATT_E_ARG;
  auto r = std::make_unique<Req>();
  r->q = *q;
  r->resp = resp;
  r->t_submit = Clock::now();
  const int nx = q->nx, ny = q->ny, nr = q->nrec;
  const size_t N = (size_t)nx * ny;
  // axes: radians by the reference's xx_deg/180.0*pi, then the degrees
  r->axes.resize(2 * (size_t)(nx + ny));
  for (int i = 0; i < nx; ++i) r->axes[i] = q->xgrids[i] / 180.0 * kPiF;
  for (int j = 0; j < ny; ++j) r->axes[nx + j] = q->ygrids[j] / 180.0 * kPiF;
  std::memcpy(&r->axes[nx + ny], q->xgrids, nx * 8);
  std::memcpy(&r->axes[2 * nx + ny], q->ygrids, ny * 8);
  r->model.resize(4 * N);
  std::memcpy(r->model.data(), q->svel, N * 8);
  for (size_t k = 0; k < N; ++k) {
    r->model[N + k] = no_metric ? 1.0 : q->m11[k];
    r->model[2 * N + k] = no_metric ? 0.0 : q->m12[k];
    r->model[3 * N + k] = no_metric ? 1.0 : q->m22[k];
  }
  r->recs.assign(4 * (size_t)nr + (nr + 1) / 2, 0.0);
  if (nr) {
    std::memcpy(&r->recs[0], q->recx, nr * 8);
    std::memcpy(&r->recs[nr], q->recy, nr * 8);
    std::memcpy(&r->recs[2 * (size_t)nr], q->tt, nr * 8);
    std::memcpy(&r->recs[3 * (size_t)nr], q->weight, nr * 8);
    int *idx = reinterpret_cast<int *>(&r->recs[4 * (size_t)nr]);
    for (int k = 0; k < nr; ++k) idx[k] = k;
  }
  r->L = layout(nx, ny, nr, S->slot_skews);
  r->cls = (long)make_skew_geom(nx, ny).ns * (ny + 63) >= kLargeSteps ? 1 : 0;
  r->out_bytes = ((size_t)nr + 32 + 2 * N) * 8;
  resp->status = 0;
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->hip_error) return S->hip_error;
  if (r->L.total * 8 > S->dev_avail || (r->L.in_end * 8 + r->out_bytes) > S->pin_avail / 2)
    return SATT_E_ARG;  // larger than the pool
  r->ticket = S->next_ticket++;
  if (ticket) *ticket = r->ticket;
  S->incoming.push_back(r.get());
  S->reqs[r->ticket] = std::move(r);
  S->st.submitted++;
  S->cv_work.notify_one();
  return SATT_OK;
}

extern "C" int satt_server_test(satt_server *S, long long ticket) {
  if (!S) return SATT_E_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return SATT_E_ARG;
  return it->second->done ? 1 : 0;
}

extern "C" int satt_server_wait(satt_server *S, long long ticket) {
  if (!S) return SATT_E_ARG;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return SATT_E_ARG;
  Req *r = it->second.get();
  S->cv_done.wait(lk, [&] { return r->done; });
  const int st = r->resp->status;
  S->reqs.erase(it);
  return st;
}

extern "C" int satt_server_drain(satt_server *S) {
  if (!S) return SATT_E_ARG;
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_done.wait(lk, [&] { return S->st.completed == S->st.submitted; });
  for (auto it = S->reqs.begin(); it != S->reqs.end();)
    it = it->second->done ? S->reqs.erase(it) : std::next(it);
  return S->hip_error;
}

extern "C" int satt_server_get_stats(satt_server *S, satt_server_stats *out) {
  if (!S || !out) return SATT_E_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->st;
  return SATT_OK;
}

extern "C" void satt_server_destroy(satt_server *S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv_work.notify_all();
  }
  if (S->thr.joinable()) S->thr.join();
  (void)hipSetDevice(S->device);
  for (auto st : S->streams) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
  for (auto e : S->spare_events) (void)hipEventDestroy(e);
  if (S->dev) (void)hipFree(S->dev);
  if (S->pin) (void)hipHostFree(S->pin);
  delete S;
}
