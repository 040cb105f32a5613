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

struct StripFwdWork {
  double *iskew;  // [slot][2 skews][kFF][SE]  (shares the adjoint's skew scratch)
  double *tau;    // [slot][N] natural
  double *snap;   // [slot][N] tau at the start of the iteration
  double *T;      // [slot][N] output travel times
};

struct StripAdjWork {
  const double *T;  // [slot][N] forward travel times
  double *iskew;    // [slot][2 skews][kAF][SE]
  double *Ta;       // [slot][N] output
  double *snap;     // [slot][N]
  double *delta;    // [slot][N] natural delta source
  const double *rhs;  // per receiver (global index)
};

void launch_period_skew(hipStream_t s, const Grid &g, const SkewGeom &sg, int np,
                        const PeriodFields &pf, double *out);
void launch_fwd_strip(hipStream_t s, const Grid &g, const SkewGeom &sg, const double *pskew,
                      const PeriodFields &pf, const Items &it, const StripFwdWork &w,
                      int *iters, int *status, int iter_stride);
void launch_adj_strip(hipStream_t s, const Grid &g, const SkewGeom &sg, const PeriodFields &pf,
                      const Items &it, const StripAdjWork &w, int *iters, int *status,
                      int iter_stride, int apply_mask);

}  // namespace satt
