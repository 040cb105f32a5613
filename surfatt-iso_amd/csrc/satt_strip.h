// Strip-pipeline kernels (satt_strip.hip): geometry of the skewed layouts
// and the launch entry points used by the batched plan.
#pragma once
#include "satt_internal.h"

namespace satt {

constexpr int kPF = 6;  // per-period skew fields: a, b, c, fun, dis_x, dis_y
constexpr int kFF = 3;  // per-item forward skew fields: T0, T0x, T0y
constexpr int kAF = 6;  // per-item adjoint skew fields: a1, a2, b1, b2, d, delta

// Skewed layout of one field: ns strips of 64 columns, R = ny+63 rows of 64
// lanes.  Plus skew: cell (64c+l, j) at row j+l; minus skew: row j-l+63.
struct SkewGeom {
  int ns, R;
  __host__ __device__ size_t skew_elems() const { return (size_t)ns * R * 64; }
};

inline SkewGeom make_skew_geom(int nx, int ny) {
  SkewGeom s;
  s.ns = (nx + 63) / 64;
  s.R = ny + 63;
  return s;
}

// Per-slot scratch of the batched path, in units of one skewed field (SE
// doubles) and one natural field (N doubles).
//   skew fields 0..11 : adjoint coefficients (a1,a2,b1,b2,d,delta) x 2 skews;
//                       the forward uses them first for T0,T0x,T0y x 2 skews
//                       (0..5), tau plus (6), tau minus (7), snapshot (8)
//   skew fields 12,13 : travel time T, plus and minus skew (forward output)
//   skew fields 14..16: Ta plus, Ta minus, snapshot (adjoint)
//   natural 0..3      : T, Ta, Td (density adjoint), delta source
// The mixed mode keeps its float fields inside the same slot: the forward's
// from double skew 9 (free until the output T skews are written), the
// adjoint's from skew 0 (its FP64 coefficients are built after the FP32
// phase has ended).
constexpr int kSlotSkews = 17;
constexpr int kSlotNat = 4;
enum : int { kSkTauP = 6, kSkTauM = 7, kSkSnapF = 8, kSkTP = 12, kSkTM = 13,
             kSkTaP = 14, kSkTaM = 15, kSkSnapA = 16 };
enum : int { kNatT = 0, kNatTa = 1, kNatTd = 2, kNatDelta = 3 };
// The mixed mode adds 6 skewed fields per slot: the adjoint's float fields
// (coefficients k0..kN in both skews, Ta plus and minus, snapshot) from
// double skew 16, beside the FP64 coefficients of its polish (0..11).
constexpr int kSlotSkewsMixed = 23;
constexpr int kSkA32 = 16;

// Per-slot forward tables (aux): sin and cos of (xx - x0) per column, the
// frozen flags of the columns and the rows, then fun0, sin(y0), cos(y0).
inline int aux_len(int nx, int ny) { return (3 * nx + ny + 4 + 1) & ~1; }

struct StripWork {
  double *skew;                // [slot][slot_skews][SE]
  double *T, *Ta, *Td, *delta; // natural layout, [slot][N] each
  double *aux;                 // [slot][aux_len]
  int aux_len, slot_skews;
};

// out32 (may be null): the same fields rounded to float, same layout.
void launch_period_skew(hipStream_t s, const Grid &g, const SkewGeom &sg, int np,
                        const PeriodFields &pf, double *out, float *out32);

// Forward: setup (per-item tables, then T0 skews and the tau init, one
// thread per element), then the
// sweeps: k_fwd_strip (FP64), or k_fwd_mixed32 (FP32 phase, iteration count
// to iters32) followed by k_fwd_polish (FP64 to the reference's stop rule).
void launch_fwd_setup(hipStream_t s, const Grid &g, const SkewGeom &sg, const PeriodFields &pf,
                      const Items &it, const StripWork &w, int *status, bool mixed);
void launch_fwd_strip(hipStream_t s, const Grid &g, const SkewGeom &sg, const double *pskew,
                      const PeriodFields &pf, const Items &it, const StripWork &w,
                      int *iters, int *status, int iter_stride);
void launch_fwd_mixed32(hipStream_t s, const Grid &g, const SkewGeom &sg, const float *pskew32,
                        const PeriodFields &pf, const Items &it, const StripWork &w,
                        int *iters32, int *status, int iter_stride);
void launch_fwd_polish(hipStream_t s, const Grid &g, const SkewGeom &sg, const double *pskew,
                       const PeriodFields &pf, const Items &it, const StripWork &w, int *iters,
                       int *status, int iter_stride);

// Adjoint: setup (the coefficients without the source -- FP64 with the Ta
// init, and in the mixed mode the float set too -- then the delta source
// from rhs per receiver, accumulated in natural layout and scattered into
// the coefficient fields), then the sweeps: k_adj_strip (FP64), or
// k_adj_mixed32 followed by k_adj_polish (one FP64 iteration).  out_nat:
// kNatTa (adjoint) or kNatTd (density adjoint).
void launch_adj_setup(hipStream_t s, const Grid &g, const SkewGeom &sg, const double *pskew,
                      const Items &it, const StripWork &w, const double *rhs, int *iters,
                      int *status, int iter_stride, bool mixed);
void launch_adj_strip(hipStream_t s, const Grid &g, const SkewGeom &sg, const Items &it,
                      const StripWork &w, int out_nat, int *iters, int *status, int iter_stride,
                      int apply_mask);
void launch_adj_mixed32(hipStream_t s, const Grid &g, const SkewGeom &sg, const Items &it,
                        const StripWork &w, int *iters32, int *status, int iter_stride);
void launch_adj_polish(hipStream_t s, const Grid &g, const SkewGeom &sg, const Items &it,
                       const StripWork &w, int out_nat, int *iters, int *status, int iter_stride,
                       int apply_mask);

}  // namespace satt
