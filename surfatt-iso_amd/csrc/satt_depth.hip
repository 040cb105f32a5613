// Rayleigh-wave depth sensitivity kernels of every column of a 3-D model:
// the phase-velocity branch of depthkernel1d (src/surfker/surfker.f90:70-103)
// over CPS sregn96 (src/surfker/sregn96.f90), the step before
// combine_kernels in the 3-D inversion (src/acqui.f90:163-195).
//
// The phase velocities come first, from the dispersion kernel on the
// refinegrid layering (satt_disp.hip); then one thread per (column, period)
// runs sregn96 at that velocity: the Dunkin compound matrices from the
// half-space up and the Haskell matrices from the surface down give the
// eigenfunctions, the energy integrals give the group velocity and the
// partial derivatives, and the sphericity correction maps them to the
// spherical model.  A thread's per-layer arrays (the model, the propagated
// vectors, the eigenfunctions, the partials) live in a global work space laid
// out [field][layer][thread], so the layer loops read and write coalesced.
//
// The arithmetic keeps the reference's kinds: the layered model is REAL(4)
// widened into REAL(8) arrays; dnka and hska take density and Vs through
// sngl(), so rho*rho and 2.0*b*b are float products; complex arithmetic uses
// the Fortran runtime's formulas (products (ac-bd, ad+bc), Smith's quotient);
// the vertical wavenumbers cdsqrt(dcmplx(x, 0)) are real or imaginary.
#include <hip/hip_runtime.h>

#include <cmath>

#include "satt_internal.h"

namespace satt {
namespace {

struct cpx {
  double re, im;
};
__device__ __forceinline__ cpx cx(double r, double i) { return cpx{r, i}; }
__device__ __forceinline__ cpx cadd(cpx a, cpx b) { return cx(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cpx csub(cpx a, cpx b) { return cx(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cpx cmul(cpx a, cpx b) {
  return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ cpx crmul(double x, cpx b) { return cmul(cx(x, 0.0), b); }
__device__ __forceinline__ cpx cdiv(cpx a, cpx b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, den = b.re + b.im * r;
    return cx((a.re + a.im * r) / den, (a.im - a.re * r) / den);
  }
  const double r = b.re / b.im, den = b.re * r + b.im;
  return cx((a.re * r + a.im) / den, (a.im * r - a.re) / den);
}
__device__ __forceinline__ cpx cneg(cpx a) { return cx(-a.re, -a.im); }
__device__ __forceinline__ double cabsd(cpx a) { return hypot(a.re, a.im); }
__device__ __forceinline__ cpx cexpd(cpx a) {
  const double e = exp(a.re);
  return cx(e * cos(a.im), e * sin(a.im));
}
__device__ __forceinline__ cpx csqrt_real(double x) {
  return x >= 0 ? cx(sqrt(x), 0.0) : cx(0.0, sqrt(-x));
}

constexpr double kPiD = (double)3.14159265359f;  // the constants module's pi

// Per-layer fields of one thread's work space.
enum : int {
  kZD, kZRHO, kZA, kZB, kXMU, kXLAM, kUR, kUZ, kTZ, kTR, kDCDA, kDCDB, kDCDR,
  kVTP, kDTP, kRTP, kEXE, kEXA, kCD0, kVV0 = kCD0 + 5, kIWAT = kVV0 + 4, kDepthFields
};

struct WS {
  double *w;
  size_t T;   // threads
  size_t t;   // this thread
  int mmax;
  __device__ double &at(int f, int m) const { return w[((size_t)f * mmax + m) * T + t]; }
  __device__ double &cd(int m, int j) const { return at(kCD0 + j, m); }
  __device__ double &vv(int m, int j) const { return at(kVV0 + j, m); }
  __device__ int iwat(int m) const { return (int)at(kIWAT, m); }
};

struct RState {
  int allfluid;
  double sumi0, sumi1, sumi2, sumi3, ugr;
  cpx ra, rb, e[4][4], einv[4][4];
};

// bldsph (sregn96.f90:128-182)
__device__ void bldsph(const WS &W) {
  const double ar = 6371.0;
  double dr = 0.0, r0 = ar;
  const int mmax = W.mmax;
  for (int i = 0; i < mmax; ++i) W.at(kDTP, i) = 0.0;
  for (int i = 0; i < mmax; ++i) {
    if (i == mmax - 1)
      dr = dr + 1.0;
    else
      dr = dr + W.at(kZD, i);
    const double r1 = ar - dr;
    const double z0 = ar * log(ar / r0);
    const double z1 = ar * log(ar / r1);
    const double tmp = (2.0 * ar) / (r0 + r1);
    W.at(kVTP, i) = tmp;
    const double rtp = pow(tmp, (double)(-2.275f));
    W.at(kRTP, i) = rtp;
    W.at(kDTP, i) = ar / r0;
    W.at(kZA, i) = W.at(kZA, i) * tmp;
    W.at(kZB, i) = W.at(kZB, i) * tmp;
    W.at(kZRHO, i) = W.at(kZRHO, i) * rtp;
    W.at(kZD, i) = z1 - z0;
    r0 = r1;
  }
  W.at(kZD, mmax - 1) = 0.0;
}

struct Trig {
  double cosp, cosq, rsinp, rsinq, sinpr, sinqr, pex, svex;
};

// varsv (sregn96.f90:826-910)
__device__ Trig varsv(cpx p, cpx q, cpx rp, cpx rsv, int iwat, double zd) {
  Trig o;
  o.pex = 0.0;
  o.svex = 0.0;
  const double pr = p.re, pim = p.im, qr = q.re, qim = q.im;
  o.pex = pr;
  const cpx epp = cdiv(cx(cos(pim), sin(pim)), cx(2.0, 0.0));
  const cpx epm = cx(epp.re, -epp.im);
  const double pfac = pr < 30.0 ? exp(-2.0 * pr) : 0.0;
  o.cosp = cadd(epp, crmul(pfac, epm)).re;
  const cpx sinp = csub(epp, crmul(pfac, epm));
  o.rsinp = cmul(rp, sinp).re;
  if (fabs(pr) < (double)1.0e-5f && cabsd(rp) < (double)1.0e-5f)
    o.sinpr = zd;
  else
    o.sinpr = cdiv(sinp, rp).re;
  if (iwat == 1) {
    o.cosq = 1.0;
    o.rsinq = 0.0;
    o.sinqr = 0.0;
  } else {
    o.svex = qr;
    const cpx eqp = cdiv(cx(cos(qim), sin(qim)), cx(2.0, 0.0));
    const cpx eqm = cx(eqp.re, -eqp.im);
    const double svfac = qr < 30.0 ? exp(-2.0 * qr) : 0.0;
    o.cosq = cadd(eqp, crmul(svfac, eqm)).re;
    const cpx sinq = csub(eqp, crmul(svfac, eqm));
    o.rsinq = cmul(rsv, sinq).re;
    if (fabs(qr) < (double)1.0e-5f && cabsd(rsv) < (double)1.0e-5f)
      o.sinqr = zd;
    else
      o.sinqr = cdiv(sinq, rsv).re;
  }
  return o;
}

// dnka (sregn96.f90:490-646); rho and b REAL(4)
__device__ void dnka(double ca[5][5], double cosp, double rsinp, double sinpr, double cossv,
                     double rsinsv, double sinsvr, float rho, float b, int iwat, double ex,
                     double exa, double wvno, double wvno2, double om2) {
  if (iwat == 1) {
    for (int j = 0; j < 5; ++j)
      for (int i = 0; i < 5; ++i) ca[i][j] = 0.0;
    const double dfac = ex > 35.0 ? 0.0 : exp(-ex);
    ca[2][2] = dfac;
    ca[0][0] = cosp;
    ca[4][4] = cosp;
    ca[0][1] = -rsinp / ((double)rho * om2);
    ca[1][0] = -(double)rho * sinpr * om2;
    ca[1][1] = cosp;
    ca[3][3] = cosp;
    ca[3][4] = ca[0][1];
    ca[4][3] = ca[1][0];
    return;
  }
  const double a0 = exa < 60.0 ? exp(-exa) : 0.0;
  const double cpcq = cosp * cossv, cpy = cosp * sinsvr, cpz = cosp * rsinsv;
  const double cqw = cossv * sinpr, cqx = cossv * rsinp, xy = rsinp * sinsvr;
  const double xz = rsinp * rsinsv, wy = sinpr * sinsvr, wz = sinpr * rsinsv;
  const float rho2 = rho * rho;
  const double gam = (double)(2.0f * b * b) * wvno2 / om2;
  const double gam2 = gam * gam, gamm1 = gam - 1.0, gamm2 = gamm1 * gamm1;
  const double cqww2 = cqw * wvno2, cqxw2 = cqx / wvno2, gg1 = gam * gamm1;
  const double a0c = cmul(cx(2.0, 0.0), csub(cx(a0, 0.0), cx(cpcq, 0.0))).re;
  const double xz2 = xz / wvno2, gxz2 = gam * xz2, g2xz2 = gam2 * xz2;
  const double a0cgg1 = a0c * (gam + gamm1);
  const double wy2 = wy * wvno2, g2wy2 = gamm2 * wy2, g1wy2 = gamm1 * wy2;
  const double rom2 = (double)rho * om2;
  double temp = a0c * gg1 + g2xz2 + g2wy2;
  ca[2][2] = a0 + temp + temp;
  ca[0][0] = cpcq - temp;
  ca[0][1] = (-cqx + wvno2 * cpy) / rom2;
  temp = cadd(cadd(cmul(cx(0.5, 0.0), cx(a0cgg1, 0.0)), cx(gxz2, 0.0)), cx(g1wy2, 0.0)).re;
  ca[0][2] = wvno * temp / rom2;
  ca[0][3] = (-cqww2 + cpz) / rom2;
  temp = wvno2 * (a0c + wy2) + xz;
  ca[0][4] = -temp / ((double)rho2 * om2 * om2);
  ca[1][0] = (-gamm2 * cqw + gam2 * cpz / wvno2) * (double)rho * om2;
  ca[1][1] = cpcq;
  ca[1][2] = (gamm1 * cqww2 - gam * cpz) / wvno;
  ca[1][3] = -wz;
  ca[1][4] = ca[0][3];
  temp = cadd(cadd(cmul(cmul(cx(0.5, 0.0), cx(a0cgg1, 0.0)), cx(gg1, 0.0)), cx(gam2 * gxz2, 0.0)),
              cx(gamm2 * g1wy2, 0.0)).re;
  {
    cpx t = cmul(cneg(cx(2.0, 0.0)), cx(temp, 0.0));
    t = cmul(t, cx((double)rho, 0.0));
    t = cmul(t, cx(om2, 0.0));
    ca[2][0] = cdiv(t, cx(wvno, 0.0)).re;
  }
  {
    const double inner = gam * cqxw2 - gamm1 * cpy;
    ca[2][1] = cmul(cx(-wvno * inner, 0.0), cx(2.0, 0.0)).re;
  }
  ca[2][3] = -2.0 * ca[1][2];
  ca[2][4] = -2.0 * ca[0][2];
  ca[3][0] = (-gam2 * cqxw2 + gamm2 * cpy) * (double)rho * om2;
  ca[3][1] = -xy;
  ca[3][2] = -ca[2][1] / 2.0;
  ca[3][3] = ca[1][1];
  ca[3][4] = ca[0][1];
  temp = gamm2 * (a0c * gam2 + g2wy2) + gam2 * g2xz2;
  ca[4][0] = -(double)rho2 * om2 * om2 * temp / wvno2;
  ca[4][1] = ca[3][0];
  ca[4][2] = -ca[2][0] / 2.0;
  ca[4][3] = ca[1][0];
  ca[4][4] = ca[0][0];
}

// evalg with jbdry = 0 (sregn96.f90:648-824) for layer m: E, EINV, ra, rb,
// and gbr(1,1:5) when gbr is given.
__device__ void evalg(const WS &W, RState &S, int m, cpx *gbr, double wvno, double om,
                      double om2, double wvno2) {
  const double za = W.at(kZA, m), zb = W.at(kZB, m), rho = W.at(kZRHO, m);
  const double xka = om / za;
  const double xkb = zb > 0.0 ? om / zb : 0.0;
  const cpx ra = csqrt_real(wvno2 - xka * xka), rb = csqrt_real(wvno2 - xkb * xkb);
  S.ra = ra;
  S.rb = rb;
  double gam = zb * wvno / om;
  gam = 2.0 * (gam * gam);
  const double gamm1 = gam - 1.0;
  cpx(*E)[4] = S.e;
  cpx(*EI)[4] = S.einv;
  if (W.iwat(m) == 0) {
    E[0][0] = cx(wvno, 0); E[0][1] = rb; E[0][2] = cx(wvno, 0); E[0][3] = cneg(rb);
    E[1][0] = ra; E[1][1] = cx(wvno, 0); E[1][2] = cneg(ra); E[1][3] = cx(wvno, 0);
    E[2][0] = cx(rho * om2 * gamm1, 0);
    E[2][1] = cdiv(crmul(rho * om2 * gam, rb), cx(wvno, 0));
    E[2][2] = cx(rho * om2 * gamm1, 0);
    E[2][3] = cdiv(crmul(-rho * om2 * gam, rb), cx(wvno, 0));
    E[3][0] = cdiv(crmul(rho * om2 * gam, ra), cx(wvno, 0));
    E[3][1] = cx(rho * om2 * gamm1, 0);
    E[3][2] = cdiv(crmul(-rho * om2 * gam, ra), cx(wvno, 0));
    E[3][3] = cx(rho * om2 * gamm1, 0);
    EI[0][0] = cx(0.5 * gam / wvno, 0);
    EI[0][1] = cdiv(cx(-0.5 * gamm1, 0), ra);
    EI[0][2] = cx(-0.5 / (rho * om2), 0);
    EI[0][3] = cdiv(cx(0.5 * wvno, 0), crmul(rho * om2, ra));
    EI[1][0] = cdiv(cx(-0.5 * gamm1, 0), rb);
    EI[1][1] = cx(0.5 * gam / wvno, 0);
    EI[1][2] = cdiv(cx(0.5 * wvno, 0), crmul(rho * om2, rb));
    EI[1][3] = cx(-0.5 / (rho * om2), 0);
    EI[2][0] = cx(0.5 * gam / wvno, 0);
    EI[2][1] = cdiv(cx(0.5 * gamm1, 0), ra);
    EI[2][2] = cx(-0.5 / (rho * om2), 0);
    EI[2][3] = cdiv(cx(-0.5 * wvno, 0), crmul(rho * om2, ra));
    EI[3][0] = cdiv(cx(0.5 * gamm1, 0), rb);
    EI[3][1] = cx(0.5 * gam / wvno, 0);
    EI[3][2] = cdiv(cx(-0.5 * wvno, 0), crmul(rho * om2, rb));
    EI[3][3] = cx(-0.5 / (rho * om2), 0);
    if (gbr) {
      gbr[0] = crmul(rho * rho * om2 * om2,
                     cadd(cneg(cmul(crmul(gam * gam, ra), rb)), cx(wvno2 * gamm1 * gamm1, 0)));
      gbr[1] = cneg(cmul(crmul(rho, crmul(wvno2, ra)), cx(om2, 0)));
      gbr[2] = cneg(cmul(cmul(crmul(rho, cadd(cneg(cmul(crmul(gam, ra), rb)),
                                               cx(wvno2 * gamm1, 0))),
                              cx(om2, 0)),
                         cx(wvno, 0)));
      gbr[3] = cmul(crmul(rho, crmul(wvno2, rb)), cx(om2, 0));
      gbr[4] = crmul(wvno2, csub(cx(wvno2, 0), cmul(ra, rb)));
      const cpx den = cneg(cmul(crmul(rho * rho * om2 * om2 * wvno2, ra), rb));
      for (int k = 0; k < 5; ++k) gbr[k] = cdiv(crmul(0.25, gbr[k]), den);
    }
  } else {
    if (gbr) {
      if (S.allfluid) {
        gbr[0] = cdiv(cx(0.5, 0), ra);
        gbr[1] = cdiv(cx(0.5, 0), cx(-rho * om2, 0));
        gbr[2] = gbr[3] = gbr[4] = cx(0, 0);
      } else {
        gbr[0] = gbr[1] = gbr[2] = cx(0, 0);
        gbr[3] = cdiv(cx(0.5 * rho * om2, 0), ra);
        gbr[4] = cx(-0.5, 0);
      }
    }
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) E[i][j] = EI[i][j] = cx(0, 0);
    E[0][0] = ra;
    E[0][1] = cneg(ra);
    E[1][0] = cx(-rho * om2, 0);
    E[1][1] = cx(-rho * om2, 0);
    EI[0][0] = cdiv(cx(0.5, 0), ra);
    EI[0][1] = cx(-0.5 / (rho * om2), 0);
    EI[1][0] = cdiv(cx(-0.5, 0), ra);
    EI[1][1] = cx(-0.5 / (rho * om2), 0);
  }
}

// hska (sregn96.f90:912-986)
__device__ void hska(double AA[4][4], double cosp, double rsinp, double sinpr, double tcossv,
                     double trsinsv, double tsinsvr, float rho, float b, int iwat, double pex,
                     double svex, double wvno, double wvno2, double om2) {
  if (iwat == 1) {
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) AA[i][j] = 0.0;
    const double dfac = pex > 35.0 ? 0.0 : exp(-pex);
    AA[0][0] = dfac;
    AA[3][3] = dfac;
    AA[1][1] = cosp;
    AA[2][2] = cosp;
    AA[1][2] = -rsinp / ((double)rho * om2);
    AA[2][1] = -(double)rho * om2 * sinpr;
    return;
  }
  const double dfac = (pex - svex) > 70.0 ? 0.0 : exp(svex - pex);
  const double cossv = dfac * tcossv, rsinsv = dfac * trsinsv, sinsvr = dfac * tsinsvr;
  const double gam = (double)(2.0f * b * b) * wvno2 / om2;
  const double gamm1 = gam - 1.0;
  const double rom2 = (double)rho * om2;
  AA[0][0] = cossv + gam * (cosp - cossv);
  AA[0][1] = -wvno * gamm1 * sinpr + gam * rsinsv / wvno;
  AA[0][2] = -wvno * (cosp - cossv) / rom2;
  AA[0][3] = (wvno2 * sinpr - rsinsv) / rom2;
  AA[1][0] = gam * rsinp / wvno - wvno * gamm1 * sinsvr;
  AA[1][1] = cosp - gam * (cosp - cossv);
  AA[1][2] = (-rsinp + wvno2 * sinsvr) / rom2;
  AA[1][3] = -AA[0][2];
  AA[2][0] = rom2 * gam * gamm1 * (cosp - cossv) / wvno;
  AA[2][1] = rom2 * (-gamm1 * gamm1 * sinpr + gam * gam * rsinsv / wvno2);
  AA[2][2] = AA[1][1];
  AA[2][3] = -AA[0][1];
  AA[3][0] = rom2 * (gam * gam * rsinp / wvno2 - gamm1 * gamm1 * sinsvr);
  AA[3][1] = -AA[2][0];
  AA[3][2] = -AA[1][0];
  AA[3][3] = AA[0][0];
}

// normc (sregn96.f90:1398-1427)
template <int N>
__device__ double normc(double *ee) {
  double t1 = 0.0;
  for (int i = 0; i < N; ++i)
    if (fabs(ee[i]) > t1) t1 = fabs(ee[i]);
  if (t1 < 1.0e-40) t1 = 1.0;
  for (int i = 0; i < N; ++i) ee[i] = ee[i] / t1;
  return log(t1);
}

// up (sregn96.f90:400-488)
__device__ double up(const WS &W, RState &S, double omega, double wvno) {
  const int mmax = W.mmax;
  const double wvno2 = wvno * wvno, om2 = omega * omega;
  cpx gbr[5];
  evalg(W, S, mmax - 1, gbr, wvno, omega, om2, wvno2);
  for (int j = 0; j < 5; ++j) W.cd(mmax - 1, j) = gbr[j].re;
  W.at(kEXE, mmax - 1) = 0.0;
  double exsum = 0.0;
  double prev[5];
  for (int j = 0; j < 5; ++j) prev[j] = gbr[j].re;
  for (int m = mmax - 2; m >= 0; --m) {
    const double za = W.at(kZA, m), zb = W.at(kZB, m), zd = W.at(kZD, m);
    const double xka = omega / za;
    const double xkb = zb > 0.0 ? omega / zb : 0.0;
    const cpx rp = csqrt_real(wvno2 - xka * xka), rsv = csqrt_real(wvno2 - xkb * xkb);
    const int iw = W.iwat(m);
    const Trig g = varsv(crmul(zd, rp), crmul(zd, rsv), rp, rsv, iw, zd);
    double ca[5][5];
    dnka(ca, g.cosp, g.rsinp, g.sinpr, g.cosq, g.rsinq, g.sinqr, (float)W.at(kZRHO, m), (float)zb,
         iw, g.pex, g.pex + g.svex, wvno, wvno2, om2);
    double ee[5];
    for (int i = 0; i < 5; ++i) {
      double cr = 0.0;
      for (int j = 0; j < 5; ++j) cr = cr + prev[j] * ca[j][i];
      ee[i] = cr;
    }
    const double exn = normc<5>(ee);
    exsum = exsum + g.pex + g.svex + exn;
    W.at(kEXE, m) = exsum;
    for (int i = 0; i < 5; ++i) {
      W.cd(m, i) = ee[i];
      prev[i] = ee[i];
    }
  }
  return W.cd(0, 0);
}

// down (sregn96.f90:988-1058)
__device__ void down(const WS &W, double omega, double wvno) {
  const int mmax = W.mmax;
  const double om2 = omega * omega, wvno2 = wvno * wvno;
  double v[4] = {1.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < 4; ++i) W.vv(0, i) = v[i];
  W.at(kEXA, 0) = 0.0;
  double exsum = 0.0;
  for (int m = 0; m < mmax - 1; ++m) {
    const double za = W.at(kZA, m), zb = W.at(kZB, m), zd = W.at(kZD, m);
    const double xka = omega / za;
    const double xkb = zb > 0.0 ? omega / zb : 0.0;
    const cpx rp = csqrt_real(wvno2 - xka * xka), rsv = csqrt_real(wvno2 - xkb * xkb);
    const int iw = W.iwat(m);
    const Trig g = varsv(crmul(zd, rp), crmul(zd, rsv), rp, rsv, iw, zd);
    double aa[4][4];
    hska(aa, g.cosp, g.rsinp, g.sinpr, g.cosq, g.rsinq, g.sinqr, (float)W.at(kZRHO, m),
         (float)zb, iw, g.pex, g.svex, wvno, wvno2, om2);
    double aa0[4];
    for (int i = 0; i < 4; ++i) {
      double cc = 0.0;
      for (int j = 0; j < 4; ++j) cc = cc + aa[i][j] * v[j];
      aa0[i] = cc;
    }
    const double ex2 = normc<4>(aa0);
    exsum = exsum + g.pex + ex2;
    W.at(kEXA, m + 1) = exsum;
    for (int i = 0; i < 4; ++i) {
      W.vv(m + 1, i) = aa0[i];
      v[i] = aa0[i];
    }
  }
}

// svfunc (sregn96.f90:192-398)
__device__ void svfunc(const WS &W, RState &S, double omega, double wvno) {
  const int mmax = W.mmax;
  up(W, S, omega, wvno);
  down(W, omega, wvno);
  const double f1213 = -W.cd(0, 1);
  W.at(kUR, 0) = W.cd(0, 2) / W.cd(0, 1);
  W.at(kUZ, 0) = 1.0;
  W.at(kTZ, 0) = 0.0;
  W.at(kTR, 0) = 0.0;
  const double exe0 = W.at(kEXE, 0);
  for (int i = 1; i < mmax; ++i) {
    const double cd1 = W.cd(i, 0), cd2 = W.cd(i, 1), cd3 = W.cd(i, 2), cd4 = -W.cd(i, 2);
    const double cd5 = W.cd(i, 3), cd6 = W.cd(i, 4);
    const double tz1 = -W.vv(i, 3), tz2 = -W.vv(i, 2), tz3 = W.vv(i, 1), tz4 = W.vv(i, 0);
    const double uu1 = tz2 * cd6 - tz3 * cd5 + tz4 * cd4;
    const double uu2 = -tz1 * cd6 + tz3 * cd3 - tz4 * cd2;
    const double uu3 = tz1 * cd5 - tz2 * cd3 + tz4 * cd1;
    const double uu4 = -tz1 * cd4 + tz2 * cd2 - tz3 * cd1;
    const double ext = W.at(kEXA, i) + W.at(kEXE, i) - exe0;
    if (ext > -80.0 && ext < 80.0) {
      const double fact = exp(ext);
      W.at(kUR, i) = uu1 * fact / f1213;
      W.at(kUZ, i) = uu2 * fact / f1213;
      W.at(kTZ, i) = uu3 * fact / f1213;
      W.at(kTR, i) = uu4 * fact / f1213;
    } else {
      W.at(kUR, i) = W.at(kUZ, i) = W.at(kTZ, i) = W.at(kTR, i) = 0.0;
    }
  }
  if (!S.allfluid) {
    int jwat = 0;
    for (int i = 0; i < mmax; ++i) {
      if (W.iwat(i) > 0)
        jwat = i + 1;
      else
        break;
    }
    for (int i = 0; i < jwat; ++i) W.at(kUR, i) = W.at(kTR, i) = 0.0;
  }
}

__device__ cpx ffunc(cpx nub, double dm) {
  if (cabsd(nub) < 1.0e-08) return cx(dm, 0);
  const cpx argcd = crmul(dm, nub);
  const cpx exqq = argcd.re < 40.0 ? cexpd(crmul(-2.0, argcd)) : cx(0, 0);
  return cdiv(csub(cx(1.0, 0), exqq), crmul(2.0, nub));
}
__device__ cpx gfunc(cpx nub, double dm) {
  const cpx argcd = crmul(dm, nub);
  if (argcd.re < 75.0) return crmul(dm, cexpd(cneg(argcd)));
  return cx(0, 0);
}
__device__ cpx h1func(cpx nua, cpx nub, double dm) {
  if (cabsd(cadd(nub, nua)) < 1.0e-08) return cx(dm, 0);
  const cpx argcd = crmul(dm, cadd(nua, nub));
  const cpx exqq = argcd.re < 40.0 ? cexpd(cneg(argcd)) : cx(0, 0);
  return cdiv(csub(cx(1.0, 0), exqq), cadd(nub, nua));
}
__device__ cpx h2func(cpx nua, cpx nub, double dm) {
  if (cabsd(csub(nub, nua)) < 1.0e-08) return cx(dm, 0);
  cpx argcd = crmul(dm, nua);
  const cpx exqp = argcd.re < 40.0 ? cexpd(cneg(argcd)) : cx(0, 0);
  argcd = crmul(dm, nub);
  const cpx exqq = argcd.re < 40.0 ? cexpd(cneg(argcd)) : cx(0, 0);
  return cdiv(csub(exqq, exqp), csub(nua, nub));
}

__device__ __forceinline__ cpx esym(const RState &S, int i, int j, int a, int b) {
  return cadd(cmul(S.e[i - 1][a - 1], S.e[j - 1][b - 1]), cmul(S.e[i - 1][b - 1], S.e[j - 1][a - 1]));
}
__device__ __forceinline__ cpx eprod(const RState &S, int i, int j, int a) {
  return cmul(S.e[i - 1][a - 1], S.e[j - 1][a - 1]);
}

// intijr (sregn96.f90:1197-1316): the reference evaluates evalg afresh for
// each integral; E, EINV, ra and rb depend on the layer, wvno and om only,
// so energy() sets them once per layer and passes them in.
__device__ double intijr(const WS &W, const RState &S, int i, int j, int m, int typelyr) {
  const cpx ra = S.ra, rb = S.rb;
  const cpx(*EI)[4] = S.einv;
  cpx c;
  if (W.iwat(m) == 1) {
    if (typelyr == 0) {
      const cpx km1pd = cadd(crmul(W.at(kUZ, m), EI[1][0]), crmul(W.at(kTZ, m), EI[1][1]));
      const cpx kmpu =
          cadd(crmul(W.at(kUZ, m + 1), EI[0][0]), crmul(W.at(kTZ, m + 1), EI[0][1]));
      const double d = W.at(kZD, m);
      const cpx FA = ffunc(ra, d), GA = gfunc(ra, d);
      c = cmul(cmul(cmul(eprod(S, i, j, 1), kmpu), kmpu), FA);
      c = cadd(c, cmul(cmul(cmul(esym(S, i, j, 1, 2), kmpu), km1pd), GA));
      c = cadd(c, cmul(cmul(cmul(eprod(S, i, j, 2), km1pd), km1pd), FA));
    } else {
      const cpx km1pd = cadd(crmul(W.at(kUZ, m), EI[1][0]), crmul(W.at(kTZ, m), EI[1][1]));
      c = cdiv(cmul(cmul(eprod(S, i, j, 2), km1pd), km1pd), crmul(2.0, ra));
    }
  } else {
    auto kpot = [&](int r, int k) {
      return cadd(cadd(cadd(cmul(EI[r][0], cx(W.at(kUR, k), 0)), cmul(EI[r][1], cx(W.at(kUZ, k), 0))),
                       cmul(EI[r][2], cx(W.at(kTZ, k), 0))),
                  cmul(EI[r][3], cx(W.at(kTR, k), 0)));
    };
    if (typelyr == 0) {
      const cpx km1pd = kpot(2, m), km1sd = kpot(3, m);
      const cpx kmpu = kpot(0, m + 1), kmsu = kpot(1, m + 1);
      const double d = W.at(kZD, m);
      const cpx FA = ffunc(ra, d), GA = gfunc(ra, d), FB = ffunc(rb, d), GB = gfunc(rb, d);
      const cpx H1 = h1func(ra, rb, d), H2 = h2func(ra, rb, d);
      c = cmul(cmul(cmul(eprod(S, i, j, 1), kmpu), kmpu), FA);
      c = cadd(c, cmul(cmul(cmul(eprod(S, i, j, 3), km1pd), km1pd), FA));
      c = cadd(c, cmul(cmul(cmul(eprod(S, i, j, 2), kmsu), kmsu), FB));
      c = cadd(c, cmul(cmul(cmul(eprod(S, i, j, 4), km1sd), km1sd), FB));
      c = cadd(c, cmul(H1, cadd(cmul(cmul(esym(S, i, j, 1, 2), kmpu), kmsu),
                                cmul(cmul(esym(S, i, j, 3, 4), km1pd), km1sd))));
      c = cadd(c, cmul(H2, cadd(cmul(cmul(esym(S, i, j, 1, 4), kmpu), km1sd),
                                cmul(cmul(esym(S, i, j, 2, 3), km1pd), kmsu))));
      c = cadd(c, cmul(cmul(cmul(GA, esym(S, i, j, 1, 3)), kmpu), km1pd));
      c = cadd(c, cmul(cmul(cmul(GB, esym(S, i, j, 2, 4)), kmsu), km1sd));
    } else {
      const cpx km1pd = kpot(2, m), km1sd = kpot(3, m);
      c = cdiv(cmul(cmul(eprod(S, i, j, 3), km1pd), km1pd), crmul(2.0, ra));
      c = cadd(c, cdiv(cmul(cmul(esym(S, i, j, 3, 4), km1pd), km1sd), cadd(ra, rb)));
      c = cadd(c, cdiv(cmul(cmul(eprod(S, i, j, 4), km1sd), km1sd), crmul(2.0, rb)));
    }
  }
  return c.re;
}

// energy (sregn96.f90:1060-1195) with getmat (:1530-1582)
__device__ void energy(const WS &W, RState &S, double om, double wvno) {
  const int mmax = W.mmax;
  S.sumi0 = S.sumi1 = S.sumi2 = S.sumi3 = 0.0;
  const double c = om / wvno, om2 = om * om, wvno2 = wvno * wvno;
  for (int m = 0; m < mmax; ++m) {
    const double za = W.at(kZA, m), zb = W.at(kZB, m), rho = W.at(kZRHO, m);
    const int iw = W.iwat(m);
    const int typelyr = m == mmax - 1 ? 1 : 0;
    evalg(W, S, m, nullptr, wvno, om, om2, wvno2);
    const double INT11 = intijr(W, S, 1, 1, m, typelyr);
    const double INT13 = intijr(W, S, 1, 3, m, typelyr);
    const double INT22 = intijr(W, S, 2, 2, m, typelyr);
    const double INT24 = intijr(W, S, 2, 4, m, typelyr);
    const double INT33 = intijr(W, S, 3, 3, m, typelyr);
    const double INT44 = intijr(W, S, 4, 4, m, typelyr);
    if (iw == 1) {
      const double ah = za, av = za, eta = 1.0;
      const double TC = rho * za * za, TA = rho * za * za, TN = 0.0;
      const double TF = TA - 2.0 * TN;
      const double a12 = -(wvno * wvno - om * om / (ah * ah)) / (rho * om * om);
      const double w = wvno / (rho * om2);
      const double URUR = INT22 * (w * w), UZUZ = INT11;
      const double URDUZ = -(wvno / (rho * om2)) * a12 * INT22;
      const double DUZDUZ = a12 * a12 * INT22;
      S.sumi0 = S.sumi0 + rho * (URUR + UZUZ);
      S.sumi1 = S.sumi1 + TA * URUR;
      S.sumi2 = S.sumi2 - TF * URDUZ;
      S.sumi3 = S.sumi3 + TC * DUZDUZ;
      const double facah = rho * ah * (URUR - 2.0 * eta * URDUZ / wvno);
      const double facav = rho * av * DUZDUZ / wvno2;
      W.at(kDCDA, m) = facah + facav;
      const double facr = -0.5 * c * c * (URUR + UZUZ);
      W.at(kDCDR, m) = 0.5 * (av * facav + ah * facah) / rho + facr;
      W.at(kUR, m) = -wvno * W.at(kTZ, m) / (rho * om2);
    } else {
      const double ah = za, av = za, bv = zb, eta = 1.0;
      const double TL = rho * zb * zb, TN = rho * zb * zb;
      const double TC = rho * za * za, TA = rho * za * za;
      const double TF = TA - 2.0 * TN;
      const double a12 = -wvno, a14 = 1.0 / TL, a21 = wvno * TF / TC, a23 = 1.0 / TC;
      const double URUR = INT11, UZUZ = INT22;
      const double DURDUR = a12 * a12 * INT22 + 2.0 * a12 * a14 * INT24 + a14 * a14 * INT44;
      const double DUZDUZ = a21 * a21 * INT11 + 2.0 * a21 * a23 * INT13 + a23 * a23 * INT33;
      const double URDUZ = a21 * INT11 + a23 * INT13;
      const double UZDUR = a12 * INT22 + a14 * INT24;
      S.sumi0 = S.sumi0 + rho * (URUR + UZUZ);
      S.sumi1 = S.sumi1 + TL * UZUZ + TA * URUR;
      S.sumi2 = S.sumi2 + TL * UZDUR - TF * URDUZ;
      S.sumi3 = S.sumi3 + TL * DURDUR + TC * DUZDUZ;
      const double facah = rho * ah * (URUR - 2.0 * eta * URDUZ / wvno);
      const double facav = rho * av * DUZDUZ / wvno2;
      const double facbh = 0.0;
      const double facbv =
          rho * bv * (UZUZ + 2.0 * UZDUR / wvno + DURDUR / wvno2 + 4.0 * eta * URDUZ / wvno);
      W.at(kDCDA, m) = facah + facav;
      W.at(kDCDB, m) = facbv + facbh;
      const double facr = -0.5 * c * c * (URUR + UZUZ);
      W.at(kDCDR, m) = 0.5 * (av * facav + ah * facah + bv * facbv) / rho + facr;
    }
  }
  S.ugr = (wvno * S.sumi1 + S.sumi2) / (om * S.sumi0);
  for (int m = 0; m < mmax; ++m) {
    W.at(kDCDA, m) = W.at(kDCDA, m) / (S.ugr * S.sumi0);
    W.at(kDCDB, m) = W.at(kDCDB, m) / (S.ugr * S.sumi0);
    W.at(kDCDR, m) = W.at(kDCDR, m) / (S.ugr * S.sumi0);
  }
  // the thickness partials (getdcdh, :1429-1528) are not among
  // depthkernel1d's outputs
}

// Vp and density from Vs (Brocher 2005), REAL(4) (surfker.f90:309-321).
__device__ inline void vprho(float vs, float &vp, float &rho) {
  const float v2 = vs * vs, v3 = vs * v2, v4 = v2 * v2;
  vp = 0.9409f + 2.0947f * vs - 0.8206f * v2 + 0.2683f * v3 - 0.0251f * v4;
  const float p2 = vp * vp, p3 = vp * p2, p4 = p2 * p2, p5 = vp * p4;
  rho = 1.6612f * vp - 0.4721f * p2 + 0.0671f * p3 - 0.0043f * p4 + 0.000106f * p5;
}

// One (column, period) per thread: the refinegrid model, sregn96 at the
// column's phase velocity with the sphericity correction (iflsph 1), and
// sen_vs/vp/rho(period, column, 1:nz) = dcdb/dcda/dcdr of the first nz
// layers (depthkernel1d, surfker.f90:88-103).
__global__ __launch_bounds__(64) void k_sregn96(DepthJob J) {
  const size_t T = (size_t)J.ncol * J.kmax;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int col = (int)(t / J.kmax), p = (int)(t % J.kmax);
  const int nz = J.nz, mmax = nz + 1;
  const size_t out0 = (size_t)p + (size_t)J.kmax * col, ostride = (size_t)J.kmax * J.ncol;
  if (J.col_status[col] != kOk) {  // no phase velocity: the reference stops there
    for (int k = 0; k < nz; ++k) {
      J.sen_vs[out0 + ostride * k] = 0.0;
      J.sen_vp[out0 + ostride * k] = 0.0;
      J.sen_rho[out0 + ostride * k] = 0.0;
    }
    return;
  }
  WS W{J.work, T, t, mmax};
  RState S{};
  // refinegrid's REAL(4) model, widened (sregn96.f90:1649-1651)
  for (int k = 0; k < mmax; ++k) {
    const int kk = k < nz ? k : nz - 1;
    const float vs = (float)J.vs[col + (size_t)J.ncol * kk];
    float vp, rho;
    vprho(vs, vp, rho);
    const int kt = k < nz - 1 ? k : nz - 2;
    const float thk = k < nz ? (float)J.depz[kt + 1] - (float)J.depz[kt] : 0.0f;
    W.at(kZB, k) = vs;
    W.at(kZA, k) = vp;
    W.at(kZRHO, k) = rho;
    W.at(kZD, k) = thk;
  }
  S.allfluid = 1;
  for (int i = 0; i < mmax; ++i) {
    if (W.at(kZB, i) > 0.0) {
      S.allfluid = 0;
      W.at(kIWAT, i) = 0.0;
    } else {
      W.at(kIWAT, i) = 1.0;
    }
  }
  bldsph(W);
  for (int i = 0; i < mmax; ++i) {
    const double zb = W.at(kZB, i), zr = W.at(kZRHO, i), za = W.at(kZA, i);
    const double xmu = zr * (zb * zb);
    W.at(kXMU, i) = xmu;
    W.at(kXLAM, i) = zr * (za * za) - 2.0 * xmu;
  }
  const double twopi = 2.0 * kPiD;
  const double omega = twopi / J.periods[p];
  const double c = J.cp[(size_t)col * J.kmax + p];
  S.ugr = c;
  const double wvno = omega / c;
  svfunc(W, S, omega, wvno);
  energy(W, S, omega, wvno);
  // sprayl (sregn96.f90:1584-1628): the kernels of the spherical model
  const double ar = 6371.0;
  const double r = c / (2.0 * ar * omega);
  const double tm = sqrt(1.0 + r * r);
  const double tm3 = tm * tm * tm;
  for (int k = 0; k < nz; ++k) {
    J.sen_vp[out0 + ostride * k] = W.at(kDCDA, k) * W.at(kVTP, k) / tm3;
    J.sen_vs[out0 + ostride * k] = W.at(kDCDB, k) * W.at(kVTP, k) / tm3;
    J.sen_rho[out0 + ostride * k] = W.at(kDCDR, k) * W.at(kRTP, k) / tm3;
  }
}

}  // namespace

size_t depth_work_doubles(size_t nthreads, int mmax) {
  return (size_t)kDepthFields * (size_t)mmax * nthreads;
}

void launch_depthkernel(hipStream_t s, const DepthJob &J) {
  const size_t T = (size_t)J.ncol * J.kmax;
  hipLaunchKernelGGL(k_sregn96, dim3((unsigned)((T + 63) / 64)), dim3(64), 0, s, J);
}

}  // namespace satt
