// Surface-wave dispersion of every column of a 3-D shear-velocity model: the
// phase or group velocity of the fundamental Love or Rayleigh mode at each
// period, the step that turns the 3-D model into the per-period velocity
// maps the solver path consumes.
//
//   fwdsurf3d_decomp / fwdsurf1d  src/surfker/surfker.f90:200-254
//   get_vprho (Brocher 2005)      src/surfker/surfker.f90:306-320
//   surfdisp96 (Herrmann's CPS)   src/surfker/surfdisp96.f:57-1098
//
// One wave per column.  A column's periods are searched in order (each
// search starts from the previous period's root) and each step of a search
// evaluates the period equation, a chain of per-layer propagators from the
// half-space up.  The per-layer terms (the trigonometric and exponential
// functions of each layer, Dunkin's 5x5 compound matrix for Rayleigh waves)
// are independent, so the lanes compute one layer each into LDS; then every
// lane runs the short chain itself, in the reference's order, so that all
// lanes hold the same value and the search runs wave-uniformly.  The column's
// layered model sits in LDS beside the terms.
//
// The arithmetic keeps the reference's types variable by variable: the model
// (thickness, Vp, Vs, density, the earth-flattening factors), the starting
// estimate and the group-velocity difference quotient are REAL(4), the period
// equation and the root search REAL(8).  Literals that the reference writes
// in default real (0.01, 0.95, 0.90, 1.5, 0.005, ...) are float here too.
#include <hip/hip_runtime.h>

#include <cmath>

#include "satt_internal.h"

namespace satt {
namespace {

constexpr double kTwoPiD = 2.0 * 3.141592653589793;  // 2.d0*3.141592653589793d0
constexpr int kSearchCap = 1 << 16;  // bound on bracketing steps (the search ends long before)

// One column's layered model in LDS (layer m of field f at w[f * n + m])
// and the per-layer terms of the period equation (kTermWords doubles per
// layer).
constexpr int kTermWords = 25;
struct Layers {
  float *w;
  double *terms;
  int n;    // layers in the model (mmax)
  int llw;  // first layer below the water (1-based, as the reference's llw)
  __device__ float &at(int f, int m) const { return w[f * n + m]; }
  __device__ float &d(int m) const { return at(0, m); }
  __device__ float &a(int m) const { return at(1, m); }
  __device__ float &b(int m) const { return at(2, m); }
  __device__ float &rho(int m) const { return at(3, m); }
  __device__ float &rtp(int m) const { return at(4, m); }
  __device__ float &dtp(int m) const { return at(5, m); }
  __device__ float &btp(int m) const { return at(6, m); }
};
constexpr int kLayerFields = 7;

__device__ inline double sgn(double x) { return copysign(1.0, x); }

// SH period equation, Haskell-Thomson from the half-space up
// (surfdisp96.f:737-799).  Lanes fill one layer's (cosq, y, z, mu) each;
// every lane then runs the chain.
__device__ double love_det(const Layers &L, double wvno, double omega) {
  const int n = L.n;
  for (int m = L.llw - 1 + (int)threadIdx.x; m <= n - 2; m += (int)blockDim.x) {
    const double beta1 = (double)L.b(m), rho1 = (double)L.rho(m);
    const double xmu = rho1 * beta1 * beta1;
    const double xkb = omega / beta1;
    const double rb = sqrt((wvno + xkb) * fabs(wvno - xkb));
    const double q = (double)L.d(m) * rb;
    double cosq, y, z;
    if (wvno < xkb) {
      const double sinq = sin(q);
      y = sinq / rb;
      z = -rb * sinq;
      cosq = cos(q);
    } else if (wvno == xkb) {
      cosq = 1.0;
      y = (double)L.d(m);
      z = 0.0;
    } else {
      const double fac = q < 16 ? exp(-2.0 * q) : 0.0;
      cosq = (1.0 + fac) * 0.5;
      const double sinq = (1.0 - fac) * 0.5;
      y = sinq / rb;
      z = rb * sinq;
    }
    double *t = L.terms + (size_t)m * kTermWords;
    t[0] = cosq;
    t[1] = y;
    t[2] = z;
    t[3] = xmu;
  }
  __syncthreads();
  const double beta1 = (double)L.b(n - 1), rho1 = (double)L.rho(n - 1);
  const double xkb = omega / beta1;
  const double rb = sqrt((wvno + xkb) * fabs(wvno - xkb));
  double e1 = rho1 * rb, e2 = 1.0 / (beta1 * beta1);
  for (int m = n - 2; m >= L.llw - 1; --m) {
    const double *t = L.terms + (size_t)m * kTermWords;
    const double cosq = t[0], y = t[1], z = t[2], xmu = t[3];
    const double e10 = e1 * cosq + e2 * xmu * z;
    const double e20 = e1 * y / xmu + e2 * cosq;
    double xnor = fabs(e10);
    const double ynor = fabs(e20);
    if (ynor > xnor) xnor = ynor;
    if (xnor < 1.0e-40) xnor = 1.0;
    e1 = e10 / xnor;
    e2 = e20 / xnor;
  }
  __syncthreads();  // the terms are rewritten by the next evaluation
  return e1;
}

// Layer terms of the compound (Dunkin) matrix with the exponents of the
// evanescent parts taken out (surfdisp96.f:904-1023).
struct LayerTerms {
  double w, cosp, a0, cpcq, cpy, cpz, cqw, cqx, xy, xz, wy, wz;
};

__device__ LayerTerms layer_terms(double p, double q, double ra, double rb, double wvno,
                                  double xka, double xkb, double dpth) {
  LayerTerms t;
  double pex = 0.0, sex = 0.0, x = 0.0, sinp;
  if (wvno < xka) {
    sinp = sin(p);
    t.w = sinp / ra;
    x = -ra * sinp;
    t.cosp = cos(p);
  } else if (wvno == xka) {
    t.cosp = 1.0;
    t.w = dpth;
    x = 0.0;
  } else {
    pex = p;
    const double fac = p < 16 ? exp(-2.0 * p) : 0.0;
    t.cosp = (1.0 + fac) * 0.5;
    sinp = (1.0 - fac) * 0.5;
    t.w = sinp / ra;
    x = ra * sinp;
  }
  double cosq, y, z;
  if (wvno < xkb) {
    const double sinq = sin(q);
    y = sinq / rb;
    z = -rb * sinq;
    cosq = cos(q);
  } else if (wvno == xkb) {
    cosq = 1.0;
    y = dpth;
    z = 0.0;
  } else {
    sex = q;
    const double fac = q < 16 ? exp(-2.0 * q) : 0.0;
    cosq = (1.0 + fac) * 0.5;
    const double sinq = (1.0 - fac) * 0.5;
    y = sinq / rb;
    z = rb * sinq;
  }
  const double exa = pex + sex;
  t.a0 = exa < 60.0 ? exp(-exa) : 0.0;
  t.cpcq = t.cosp * cosq;
  t.cpy = t.cosp * y;
  t.cpz = t.cosp * z;
  t.cqw = cosq * t.w;
  t.cqx = cosq * x;
  t.xy = x * y;
  t.xz = x * z;
  t.wy = t.w * y;
  t.wz = t.w * z;
  return t;
}

// Dunkin's compound matrix of one layer (surfdisp96.f:1054-1098), ca[i][j]
// for the reference's ca(i+1, j+1), into 25 words.
__device__ void dunkin_matrix(double *out, const LayerTerms &t, double wvno2, double gam,
                              double gammk, double rho) {
  const double gamm1 = gam - 1.0, twgm1 = gam + gamm1, gmgmk = gam * gammk;
  const double gmgm1 = gam * gamm1, gm1sq = gamm1 * gamm1, rho2 = rho * rho;
  const double a0pq = t.a0 - t.cpcq;
  double ca[5][5];
  ca[0][0] = t.cpcq - 2.0 * gmgm1 * a0pq - gmgmk * t.xz - wvno2 * gm1sq * t.wy;
  ca[0][1] = (wvno2 * t.cpy - t.cqx) / rho;
  ca[0][2] = -(twgm1 * a0pq + gammk * t.xz + wvno2 * gamm1 * t.wy) / rho;
  ca[0][3] = (t.cpz - wvno2 * t.cqw) / rho;
  ca[0][4] = -(2.0 * wvno2 * a0pq + t.xz + wvno2 * wvno2 * t.wy) / rho2;
  ca[1][0] = (gmgmk * t.cpz - gm1sq * t.cqw) * rho;
  ca[1][1] = t.cpcq;
  ca[1][2] = gammk * t.cpz - gamm1 * t.cqw;
  ca[1][3] = -t.wz;
  ca[1][4] = ca[0][3];
  ca[3][0] = (gm1sq * t.cpy - gmgmk * t.cqx) * rho;
  ca[3][1] = -t.xy;
  ca[3][2] = gamm1 * t.cpy - gammk * t.cqx;
  ca[3][3] = ca[1][1];
  ca[3][4] = ca[0][1];
  ca[4][0] = -(2.0 * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * t.xz + gm1sq * gm1sq * t.wy) * rho2;
  ca[4][1] = ca[3][0];
  ca[4][2] = -(gammk * gamm1 * twgm1 * a0pq + gam * gammk * gammk * t.xz + gamm1 * gm1sq * t.wy) *
             rho;
  ca[4][3] = ca[1][0];
  ca[4][4] = ca[0][0];
  const double tt = -2.0 * wvno2;
  ca[2][0] = tt * ca[4][2];
  ca[2][1] = tt * ca[3][2];
  ca[2][2] = t.a0 + 2.0 * (t.cpcq - ca[0][0]);
  ca[2][3] = tt * ca[1][2];
  ca[2][4] = tt * ca[0][2];
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) out[5 * i + j] = ca[i][j];
}

// e <- e * ca, then normalised by its largest element (surfdisp96.f:
// 858-872, 1025-1052).
__device__ void chain_step(double e[5], const double *ca) {
  double ee[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    double cr = 0.0;
#pragma unroll
    for (int j = 0; j < 5; ++j) cr = cr + e[j] * ca[5 * j + i];
    ee[i] = cr;
  }
  double big = 0.0;
#pragma unroll
  for (int i = 0; i < 5; ++i)
    if (fabs(ee[i]) > big) big = fabs(ee[i]);
  if (big < 1.0e-40) big = 1.0;
#pragma unroll
  for (int i = 0; i < 5; ++i) e[i] = ee[i] / big;
}

__device__ inline double radical(double wvno, double k) { return sqrt((wvno + k) * fabs(wvno - k)); }

// P-SV period equation, compound matrices from the half-space up
// (surfdisp96.f:801-902).
__device__ double rayleigh_det(const Layers &L, double wvno, double omga) {
  const double omega = omga < 1.0e-4 ? 1.0e-4 : omga;
  const double wvno2 = wvno * wvno;
  const int n = L.n;
  double xka = omega / (double)L.a(n - 1), xkb = omega / (double)L.b(n - 1);
  double ra = radical(wvno, xka), rb = radical(wvno, xkb);
  double t = (double)L.b(n - 1) / omega;
  double gammk = 2.0 * t * t, gam = gammk * wvno2;
  const double gamm1 = gam - 1.0;
  double rho1 = (double)L.rho(n - 1);
  double e[5];
  e[0] = rho1 * rho1 * (gamm1 * gamm1 - gam * gammk * ra * rb);
  e[1] = -rho1 * ra;
  e[2] = rho1 * (gamm1 - gammk * ra * rb);
  e[3] = rho1 * rb;
  e[4] = wvno2 - ra * rb;
  for (int m = L.llw - 1 + (int)threadIdx.x; m <= n - 2; m += (int)blockDim.x) {
    const double xkam = omega / (double)L.a(m), xkbm = omega / (double)L.b(m);
    const double tm = (double)L.b(m) / omega;
    const double gammkm = 2.0 * tm * tm, gamm = gammkm * wvno2;
    const double ram = radical(wvno, xkam), rbm = radical(wvno, xkbm);
    const double dpth = (double)L.d(m);
    const LayerTerms lt = layer_terms(ram * dpth, rbm * dpth, ram, rbm, wvno, xkam, xkbm, dpth);
    dunkin_matrix(L.terms + (size_t)m * kTermWords, lt, wvno2, gamm, gammkm, (double)L.rho(m));
  }
  __syncthreads();
  for (int m = n - 2; m >= L.llw - 1; --m) chain_step(e, L.terms + (size_t)m * kTermWords);
  __syncthreads();  // the terms are rewritten by the next evaluation
  if (L.llw != 1) {  // the water layer on top
    xka = omega / (double)L.a(0);
    ra = radical(wvno, xka);
    const double dpth = (double)L.d(0);
    rho1 = (double)L.rho(0);
    const double znul = 1.0e-05;
    const LayerTerms lt = layer_terms(ra * dpth, znul, ra, znul, wvno, xka, znul, dpth);
    const double w0 = -rho1 * lt.w;
    return lt.cosp * e[0] + w0 * e[1];
  }
  return e[0];
}

__device__ inline double period_eq(const Layers &L, int ifunc, double wvno, double omega) {
  return ifunc == 1 ? love_det(L, wvno, omega) : rayleigh_det(L, wvno, omega);
}

// Refine a bracketed root: interval halving and Neville iteration on c(del),
// alternating as the reference does (surfdisp96.f:578-713).
__device__ double refine_root(const Layers &L, int ifunc, double t, double c1, double c2,
                             double del1, double del2) {
  const double omega = kTwoPiD / t;
  auto halve = [&](double &c3, double &del3) {
    c3 = 0.5 * (c1 + c2);
    del3 = period_eq(L, ifunc, omega / c3, omega);
  };
  double c3, del3;
  halve(c3, del3);
  int nev = 1, m = 1;
  double x[20], y[20];
  const double s_frac = (double)0.01f;  // a default-real literal
  for (int nctrl = 2; nctrl < 100; ++nctrl) {
    if (c3 < fmin(c1, c2) || c3 > fmax(c1, c2)) {
      nev = 0;
      halve(c3, del3);
    }
    const double s13 = del1 - del3, s32 = del3 - del2;
    if (sgn(del3) * sgn(del1) < 0.0) {
      c2 = c3;
      del2 = del3;
    } else {
      c1 = c3;
      del1 = del3;
    }
    if (fabs(c1 - c2) <= 1.0e-6 * c1) break;
    if (sgn(s13) != sgn(s32)) nev = 0;
    const double ss1 = fabs(del1), s1 = s_frac * ss1, ss2 = fabs(del2), s2 = s_frac * ss2;
    if (s1 > ss2 || s2 > ss1 || nev == 0) {
      halve(c3, del3);
      nev = 1;
      m = 1;
    } else {
      if (nev == 2) {
        x[m] = c3;
        y[m] = del3;
      } else {
        x[0] = c1;
        y[0] = del1;
        x[1] = c2;
        y[1] = del2;
        m = 1;
      }
      // Neville's scheme on x(y), evaluated at y = 0
      bool flat = false;
      for (int kk = 1; kk <= m; ++kk) {
        const int j = m - kk;
        const double denom = y[m] - y[j];
        if (fabs(denom) < 1.0e-10 * fabs(y[m])) {
          flat = true;
          break;
        }
        x[j] = (-y[j] * x[j + 1] + y[m] * x[j]) / denom;
      }
      if (!flat) {
        c3 = x[0];
        del3 = period_eq(L, ifunc, omega / c3, omega);
        nev = 2;
        m = m + 1 > 10 ? 10 : m + 1;
      } else {
        halve(c3, del3);
        nev = 1;
        m = 1;
      }
    }
  }
  return c3;
}

// The search state the reference keeps between its calls (del1st).
struct SearchState {
  double del1st = 0.0;
};

// Step from c1 in increments of dc until the period equation changes sign,
// then refine (surfdisp96.f:403-500).  Returns false where the reference
// reports no root (iret = -1).
__device__ bool find_root(const Layers &L, int ifunc, double t1, double &c1, double clow,
                          double dc, double cm, float betmx, bool first, SearchState &S) {
  double omega = kTwoPiD / t1;
  double del1 = period_eq(L, ifunc, omega / c1, omega);
  if (first) S.del1st = del1;
  const double plmn = sgn(S.del1st) * sgn(del1);
  int idir = first || plmn >= 0.0 ? 1 : -1;
  for (int step = 0; step < kSearchCap; ++step) {
    const double c2 = idir > 0 ? c1 + dc : c1 - dc;
    if (c2 <= clow) {
      idir = 1;
      c1 = clow;
      continue;
    }
    omega = kTwoPiD / t1;
    const double del2 = period_eq(L, ifunc, omega / c2, omega);
    if (sgn(del1) != sgn(del2)) {
      c1 = refine_root(L, ifunc, t1, c1, c2, del1, del2);
      return !(c1 > (double)betmx);
    }
    c1 = c2;
    del1 = del2;
    if (c1 < cm) return false;
    if (c1 >= (double)betmx + dc) return false;
  }
  return false;
}

// Rayleigh-wave half-space velocity by Newton steps on the Rayleigh
// function, in single precision (surfdisp96.f:380-401).
__device__ float halfspace_rayleigh(float a, float b) {
  float c = 0.95f * b;
  for (int it = 0; it < 5; ++it) {
    const float gamma = b / a, kappa = c / b;
    const float k2 = kappa * kappa;
    const float gk = gamma * kappa;
    const float gk2 = gk * gk;
    const float fac1 = sqrtf(1.0f - gk2), fac2 = sqrtf(1.0f - k2);
    const float tk = 2.0f - k2;
    const float fr = tk * tk - 4.0f * fac1 * fac2;
    float frp = -4.0f * tk * kappa + 4.0f * fac2 * gamma * gamma * kappa / fac1 +
                4.0f * fac1 * kappa / fac2;
    frp = frp / b;
    c = c - fr / frp;
  }
  return c;
}

// Earth flattening (surfdisp96.f:502-576): thicknesses and velocities to the
// flat-earth equivalents, the factors kept for the density correction.
__device__ void flatten(const Layers &L) {
  const int n = L.n;
  const double ar = 6370.0;
  double dr = 0.0, r0 = ar;
  L.d(n - 1) = 1.0f;
  for (int i = 0; i < n; ++i) {
    L.dtp(i) = L.d(i);
    L.rtp(i) = L.rho(i);
  }
  for (int i = 0; i < n; ++i) {
    dr = dr + (double)L.d(i);
    const double r1 = ar - dr;
    const double z0 = ar * log(ar / r0), z1 = ar * log(ar / r1);
    L.d(i) = (float)(z1 - z0);
    const double tmp = (ar + ar) / (r0 + r1);
    L.a(i) = (float)((double)L.a(i) * tmp);
    L.b(i) = (float)((double)L.b(i) * tmp);
    L.btp(i) = (float)tmp;
    r0 = r1;
  }
  L.d(n - 1) = 0.0f;
}

// The density factors of the flattening, one layer per lane (the half-space
// thickness the reference restores around them is 0 again afterwards).
__device__ void flatten_density(const Layers &L, int ifunc) {
  const int n = L.n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float bt = L.btp(i);
    if (ifunc == 1) {
      const float b2 = bt * bt;
      L.rho(i) = L.rtp(i) * (1.0f / (bt * (b2 * b2)));
    } else {
      L.rho(i) = L.rtp(i) * powf(bt, -2.275f);
    }
  }
}

// Vp and density from Vs (Brocher 2005), single precision
// (src/surfker/surfker.f90:306-320).
__device__ inline void brocher(float vs, float &vp, float &rho) {
  const float v2 = vs * vs, v3 = vs * v2, v4 = v2 * v2;
  vp = 0.9409f + 2.0947f * vs - 0.8206f * v2 + 0.2683f * v3 - 0.0251f * v4;
  const float p2 = vp * vp, p3 = vp * p2, p4 = p2 * p2, p5 = vp * p4;
  rho = 1.6612f * vp - 0.4721f * p2 + 0.0671f * p3 - 0.0043f * p4 + 0.000106f * p5;
}

__global__ __launch_bounds__(64) void k_fwdsurf(DispJob J) {
  extern __shared__ double lds_terms[];
  const int col = blockIdx.x, lane = threadIdx.x;
  const int nz = J.nz, kmax = J.kmax, n = nz + (J.refine ? 1 : 0);
  Layers L{reinterpret_cast<float *>(lds_terms + (size_t)n * kTermWords), lds_terms, n, 1};
  // fwdsurf1d: Vs to REAL(4), Vp and density, layer thicknesses
  // depz(k+1)-depz(k) in double over a half-space at the bottom; refinegrid:
  // the same layers with thicknesses from the REAL(4) depths, the last one
  // repeated, over a half-space copy of the last layer
  for (int k = lane; k < n; k += 64) {
    const int kk = k < nz ? k : nz - 1;
    const float vs = (float)J.vs[col + (size_t)J.ncol * kk];
    float vp, rho;
    brocher(vs, vp, rho);
    L.b(k) = vs;
    L.a(k) = vp;
    L.rho(k) = rho;
    if (!J.refine) {
      L.d(k) = k < nz - 1 ? (float)(J.depz[k + 1] - J.depz[k]) : 0.0f;
    } else {
      const int kt = k < nz - 1 ? k : nz - 2;
      L.d(k) = k < nz ? (float)J.depz[kt + 1] - (float)J.depz[kt] : 0.0f;
    }
  }
  __syncthreads();
  float vmin = L.b(0);
  for (int k = 1; k < n; ++k) vmin = L.b(k) < vmin ? L.b(k) : vmin;
  // surfdisp96 with a spherical earth (iflsph 1) and the fundamental mode
  const float ddc = vmin < 1.0f ? 0.001f : (vmin < 2.0f ? 0.002f : 0.005f);
  const float sone = 1.5f, h = 0.005f;
  L.llw = L.b(0) <= 0.0f ? 2 : 1;
  const double one = 1.0e-2;
  __syncthreads();
  if (lane == 0) flatten(L);  // a running depth: one lane
  __syncthreads();
  float betmx = -1.0e20f, betmn = 1.0e20f;
  int jmn = 0, jsol = 1;
  for (int i = 0; i < n; ++i) {
    const float bi = L.b(i), ai = L.a(i);
    if (bi > 0.01f && bi < betmn) {
      betmn = bi;
      jmn = i;
      jsol = 1;
    } else if (bi <= 0.01f && ai < betmn) {
      betmn = ai;
      jmn = i;
      jsol = 0;
    }
    if (bi > betmx) betmx = bi;
  }
  const int ifunc = J.iwave;
  flatten_density(L, ifunc);
  __syncthreads();
  const double onea = (double)sone;
  float cc1 = jsol == 0 ? betmn : halfspace_rayleigh(L.a(jmn), L.b(jmn));
  cc1 = 0.95f * cc1;
  cc1 = 0.90f * cc1;
  const double cc = (double)cc1, dc = fabs((double)ddc), cm = cc;
  double c1 = cc;
  SearchState S;
  double cprev = 0.0;
  int status = kOk;
  for (int k = 0; k < kmax; ++k) {
    double t1 = J.periods[k];
    float t1a, t1b = 0.0f;
    if (J.igr > 0) {
      t1a = (float)(t1 / (double)(1.0f + h));
      t1b = (float)(t1 / (double)(1.0f - h));
      t1 = (double)t1a;
    } else {
      t1a = (float)t1;
    }
    double clow;
    if (k == 0) {
      c1 = cc;
      clow = cc;
    } else {
      c1 = cprev - onea * dc;
      clow = cm;
    }
    if (!find_root(L, ifunc, t1, c1, clow, dc, cm, betmx, k == 0, S)) {
      status = kEDispersion;  // the reference stops: 'Dispersion calculation failed'
      if (lane == 0)
        for (int q = k; q < kmax; ++q) J.out[q + (size_t)kmax * col] = 0.0;
      break;
    }
    const double ck = c1;
    cprev = ck;
    double cgrp = 0.0;
    if (J.igr > 0) {
      t1 = (double)t1b;
      const double clow2 = 0.0 + one * dc;  // cb(k) is still 0 for the fundamental mode
      c1 = c1 - onea * dc;
      if (!find_root(L, ifunc, t1, c1, clow2, dc, cm, betmx, false, S)) c1 = ck;
      cgrp = c1;
    }
    const float cc0 = (float)ck, cc1k = (float)cgrp;
    double out;
    if (J.igr == 0) {
      out = (double)cc0;
    } else {
      const float gvel = (1.0f / t1a - 1.0f / t1b) / (1.0f / (t1a * cc0) - 1.0f / (t1b * cc1k));
      out = (double)gvel;
    }
    if (lane == 0) J.out[k + (size_t)kmax * col] = out;
  }
  if (lane == 0) J.status[col] = status;
}

}  // namespace

size_t fwdsurf_lds_bytes(int n) {
  return (size_t)n * (kTermWords * sizeof(double) + kLayerFields * sizeof(float));
}

void launch_fwdsurf(hipStream_t s, const DispJob &J) {
  hipLaunchKernelGGL(k_fwdsurf, dim3((unsigned)J.ncol), dim3(64),
                     fwdsurf_lds_bytes(J.nz + (J.refine ? 1 : 0)), s, J);
}

}  // namespace satt
