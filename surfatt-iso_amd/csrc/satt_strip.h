// Strip-pipeline kernels (satt_strip.hip): geometry of the skewed layouts
// and the launch entry points used by the batched plan.
#pragma once
#include "../../include/satt.h"
#include "satt_internal.h"

namespace satt {

constexpr int kPF = 6;  // per-period skew fields: a, b, c, fun, dis_x, dis_y
constexpr int kFF = 3;  // per-item forward skew fields: T0, T0x, T0y
constexpr int kAF = 6;  // per-item adjoint skew fields: a1, a2, b1, b2, d, delta
// The FP32 strip passes address a set of up to kPF fields of one skew with
// 32-bit byte offsets: plans in mixed precision keep a skew within this.
constexpr size_t kMaxSkewElems32 = 0xFFFFFFFFull / (kPF * sizeof(float));

// Per-slot scratch of the batched path, in units of one skewed field (SE
// doubles) and one natural field (N doubles).
//   skew fields 0..11 : adjoint coefficients (a1,a2,b1,b2,d,delta) x 2 skews;
//                       the forward uses them first for T0,T0x,T0y x 2 skews
//                       (0..5), tau plus (6), tau minus (7), snapshot (8)
//   skew fields 12,13 : travel time T, plus and minus skew (forward output)
//   skew fields 14..16: Ta plus, Ta minus, snapshot (adjoint)
//   natural 0..3      : T, Ta, Td (density adjoint), delta source
// The mixed mode keeps the forward's float fields from double skew 9 (free
// until the output T skews are written).
constexpr int kSlotSkews = 17;
constexpr int kSlotNat = 4;
enum : int { kSkTauP = 6, kSkTauM = 7, kSkSnapF = 8, kSkTP = 12, kSkTM = 13,
             kSkTaP = 14, kSkTaM = 15, kSkSnapA = 16 };
enum : int { kNatT = 0, kNatTa = 1, kNatTd = 2, kNatDelta = 3 };
// The mixed mode adds 6 skewed fields per slot: the adjoint's float fields
// (coefficients k0..kN in both skews, which both its FP32 phase and its FP64
// polish read, Ta plus and minus, snapshot) from double skew 16.  Of the
// forward's FP64 fields it keeps T0 (0, 3), tau and the snapshot: its passes
// read the period fields and T0x, T0y in float.
constexpr int kSlotSkewsMixed = 23;
constexpr int kSkA32 = 16;

// Per-slot forward tables (aux): sin and cos of (xx - x0) per column, the
// frozen flags of the columns and the rows, then fun0, sin(y0), cos(y0).
inline int aux_len(int nx, int ny) { return (3 * nx + ny + 4 + 1) & ~1; }

// Doubles of one slot's scratch: its skewed fields, kSlotNat natural fields
// and the forward tables, rounded up to 256 bytes.
inline size_t slot_doubles(int nx, int ny, int slot_skews) {
  const size_t n = (size_t)slot_skews * make_skew_geom(nx, ny).skew_elems() +
                   (size_t)kSlotNat * nx * ny + (size_t)aux_len(nx, ny);
  return (n + 31) & ~(size_t)31;
}

// Sweep stages of the strip kernels; each has its own team size: the FP32
// forward, the FP64 forward (polish, or the whole FP64 solve), the FP32
// adjoint and the FP64 adjoint.
enum { kStageFwd32 = 0, kStageFwd64 = 1, kStageAdj32 = 2, kStageAdj64 = 3, kStages = 4 };

// Sizes of a launch group, for the grids of the element-wise kernels, and
// the waves per item of its strip kernels in each stage.
struct GroupDims {
  int count;        // slots
  size_t max_se;    // largest skewed field
  size_t max_n;     // largest grid
  int max_nxy;      // largest nx or ny
  int max_ny = 0, max_ns = 0;  // largest ny and strip count
  int waves[kStages] = {1, 1, 1, 1};  // waves per item (a team; at most kMaxTeam)
  int depth[kStages] = {2, 2, 2, 2};  // steps the sweeps fetch their operands ahead
};

// The deeper of the two prefetch depths the sweep kernels are built with
// (the other is two steps).  Fetching further ahead hides more memory
// latency where few waves share a SIMD, at the price of one operand set of
// registers per step.
constexpr int kDeepPrefetch = 4;

// XCDs (chiplets, each with its own L2) of the device: workgroup w of a
// launch runs on XCD w mod kXcds.
constexpr int kXcds = 8;

// Waves per item of the strip kernels: each strip of a sweep goes to one
// wave of the item's team, the next strip's wave following at least one
// strip width behind.  More waves cut an item's latency on wide grids; one
// wave per item is the throughput form when items fill the device.
constexpr int kMaxTeam = 8;
constexpr size_t kTeamLdsMax = 160 * 1024;  // a gfx950 workgroup's LDS
// one wave's two FP64 edge channels of SATT_MAX_NY rows fill it
static_assert(2 * 8 * (size_t)SATT_MAX_NY <= kTeamLdsMax, "SATT_MAX_NY exceeds the LDS");
// The largest team (<= want) of a sweep stage whose LDS (edge channels of
// the stage's value type) fits for grids up to ny rows, ns strips.
int team_waves(int want, int ny, int ns, int stage);
// Measures each sweep stage's fastest team size and prefetch depth on the
// plan's own inputs and leaves the plan set to them (satt_api.hip).
int tune_plan_shapes(satt_plan *P, satt_tune_result &r);

// Batched kernels over the slots of a launch group (satt_strip.hip); all
// asynchronous on `s`.  Forward: setup (per-item tables, then T0 skews and
// the tau init, one thread per element), then the sweeps: k_fwd_strip
// (FP64), or k_fwd_mixed32 (FP32 phase) followed by k_fwd_polish (FP64 to
// the reference's stop rule).
void launch_fwd_setup(hipStream_t s, const SlotDesc *sd, const GroupDims &gd, bool mixed);
void launch_fwd_strip(hipStream_t s, const SlotDesc *sd, const GroupDims &gd);
void launch_fwd_mixed32(hipStream_t s, const SlotDesc *sd, const GroupDims &gd);
void launch_fwd_polish(hipStream_t s, const SlotDesc *sd, const GroupDims &gd);
// Adjoint (which 0) or density adjoint (which 1): setup (the coefficients
// without the source -- FP64 with the Ta init, and in the mixed mode the
// float set too -- then the delta source from the rhs, accumulated in
// natural layout and scattered into the coefficient fields), then the
// sweeps: k_adj_strip (FP64), or k_adj_mixed32 followed by k_adj_polish
// (one FP64 iteration).  Output Ta (which 0) or Td (which 1), masked.
void launch_adj_setup(hipStream_t s, const SlotDesc *sd, const GroupDims &gd, int which,
                      bool mixed);
void launch_adj_strip(hipStream_t s, const SlotDesc *sd, const GroupDims &gd, int which);
void launch_adj_mixed32(hipStream_t s, const SlotDesc *sd, const GroupDims &gd, int which);
void launch_adj_polish(hipStream_t s, const SlotDesc *sd, const GroupDims &gd, int which);
// Each slot's own kernel Ta/svel**2 into its `kern` field (when set).
void launch_item_kernel(hipStream_t s, const SlotDesc *sd, const GroupDims &gd);
// Per-period skewed fields of one grid (out32: the same in float, or null).
void launch_period_skew(hipStream_t s, const Grid &g, const SkewGeom &sg, int np,
                        const PeriodFields &pf, double *out, float *out32);

}  // namespace satt
