/*
 * Test infrastructure only (the oracle): a CPU restatement of the reference's
 * Rayleigh-wave depth sensitivity kernels, the phase-velocity branch of
 * depthkernel1d (src/surfker/surfker.f90:70-103) over CPS sregn96
 * (src/surfker/sregn96.f90), with refinegrid (surfker.f90:285-307) and
 * get_vprho (:309-321).  Nothing in the product links or calls it.
 *
 * Each routine follows the reference's statements in order.  Variables keep
 * the reference's kinds: the layered model is REAL(4) (rthk, rvp, rvs, rrho)
 * widened into the REAL(8) module arrays; dnka and hska take density and Vs
 * through sngl(), so their products rho*rho and 2.0*b*b are evaluated in
 * float.  Complex arithmetic is written out with the Fortran runtime's
 * formulas: products (ac-bd, ad+bc), quotients by Smith's method, and the
 * vertical wavenumbers cdsqrt(dcmplx(x,0)), which are either real or
 * imaginary.
 *
 * Pinned by tests/test_oracle_depth.py: finite differences of the
 * reference's own surfdisp96 (oracle/_ref) for dc/dVs, dc/dVp and
 * dc/drho of single layers, the energy-integral group velocity against
 * surfdisp96's, and the kernels' closed forms on a half-space.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double re, im;
} cpx;

static cpx cx(double r, double i) {
  cpx z = {r, i};
  return z;
}
static cpx cadd(cpx a, cpx b) { return cx(a.re + b.re, a.im + b.im); }
static cpx csub(cpx a, cpx b) { return cx(a.re - b.re, a.im - b.im); }
static cpx cmul(cpx a, cpx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
/* a real operand of a complex operation is the complex (x, 0) */
static cpx crmul(double x, cpx b) { return cmul(cx(x, 0.0), b); }
/* Smith's quotient */
static cpx cdiv(cpx a, cpx b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, den = b.re + b.im * r;
    return cx((a.re + a.im * r) / den, (a.im - a.re * r) / den);
  } else {
    const double r = b.re / b.im, den = b.re * r + b.im;
    return cx((a.re * r + a.im) / den, (a.im * r - a.re) / den);
  }
}
static cpx cneg(cpx a) { return cx(-a.re, -a.im); }
static double cabs_(cpx a) { return hypot(a.re, a.im); }
static cpx cexp_(cpx a) {
  const double e = exp(a.re);
  return cx(e * cos(a.im), e * sin(a.im));
}
/* cdsqrt(dcmplx(x, 0)) */
static cpx csqrt_real(double x) { return x >= 0 ? cx(sqrt(x), 0.0) : cx(0.0, sqrt(-x)); }

static const double kPi = (double)3.14159265359f; /* constants module pi */

/* The module RayleighWaveModel's state for one sregn96 call. */
typedef struct {
  int mmax;
  double *zd, *zrho, *za, *zb, *xmu, *xlam;
  double uu0[4];
  double *ur, *uz, *tz, *tr, *dcda, *dcdb, *dcdr, *dcdh;
  int *iwat;
  double sumi0, sumi1, sumi2, sumi3, flagr, are, ugr;
  double *vtp, *dtp, *rtp;
  double *exe, *exa, *cd, *vv; /* cd(mmax,5), vv(mmax,4) column-major */
  int allfluid;
  cpx ra, rb, e[4][4], einv[4][4];
} RModel;

#define CD(M, m, j) ((M)->cd[(m) + (size_t)(M)->mmax * (j)])
#define VV(M, m, j) ((M)->vv[(m) + (size_t)(M)->mmax * (j)])

static void alloc_model(RModel *M, int n) {
  memset(M, 0, sizeof *M);
  M->mmax = n;
  double *p = (double *)calloc((size_t)n * 32, sizeof(double));
  double **f[] = {&M->zd,   &M->zrho, &M->za,   &M->zb,   &M->xmu,  &M->xlam, &M->ur,
                  &M->uz,   &M->tz,   &M->tr,   &M->dcda, &M->dcdb, &M->dcdr, &M->dcdh,
                  &M->vtp,  &M->dtp,  &M->rtp,  &M->exe,  &M->exa};
  for (size_t k = 0; k < sizeof f / sizeof f[0]; ++k) *f[k] = p + k * n;
  M->cd = p + 19 * (size_t)n;  /* 5 columns */
  M->vv = p + 24 * (size_t)n;  /* 4 columns */
  M->iwat = (int *)calloc((size_t)n, sizeof(int));
}
static void free_model(RModel *M) {
  free(M->zd);
  free(M->iwat);
}

/* bldsph (sregn96.f90:128-182) */
static void bldsph(RModel *M) {
  const double ar = 6371.0;
  double dr = 0.0, r0 = ar;
  for (int i = 0; i < M->mmax; ++i) M->dtp[i] = 0.0;
  for (int i = 0; i < M->mmax; ++i) {
    if (i == M->mmax - 1)
      dr = dr + 1.0;
    else
      dr = dr + M->zd[i];
    const double r1 = ar - dr;
    const double z0 = ar * log(ar / r0);
    const double z1 = ar * log(ar / r1);
    const double tmp = (2.0 * ar) / (r0 + r1);
    M->vtp[i] = tmp;
    M->rtp[i] = pow(tmp, (double)(-2.275f));
    M->dtp[i] = ar / r0;
    M->za[i] = M->za[i] * tmp;
    M->zb[i] = M->zb[i] * tmp;
    M->zrho[i] = M->zrho[i] * M->rtp[i];
    M->zd[i] = z1 - z0;
    r0 = r1;
  }
  M->zd[M->mmax - 1] = 0.0;
}

/* varsv (sregn96.f90:826-910) for p = rp*h, q = rsv*h */
static void varsv(cpx p, cpx q, cpx rp, cpx rsv, double *cosp, double *cosq, double *rsinp,
                  double *rsinq, double *sinpr, double *sinqr, double *pex, double *svex, int iwat,
                  double zd) {
  *pex = 0.0;
  *svex = 0.0;
  const double pr = p.re, pi = p.im, qr = q.re, qi = q.im;
  *pex = pr;
  const cpx epp = cdiv(cx(cos(pi), sin(pi)), cx(2.0, 0.0));
  const cpx epm = cx(epp.re, -epp.im);
  const double pfac = pr < 30.0 ? exp(-2.0 * pr) : 0.0;
  *cosp = cadd(epp, crmul(pfac, epm)).re;
  const cpx sinp = csub(epp, crmul(pfac, epm));
  *rsinp = cmul(rp, sinp).re;
  if (fabs(pr) < (double)1.0e-5f && cabs_(rp) < (double)1.0e-5f)
    *sinpr = zd;
  else
    *sinpr = cdiv(sinp, rp).re;
  if (iwat == 1) {
    *cosq = 1.0;
    *rsinq = 0.0;
    *sinqr = 0.0;
  } else {
    *svex = qr;
    const cpx eqp = cdiv(cx(cos(qi), sin(qi)), cx(2.0, 0.0));
    const cpx eqm = cx(eqp.re, -eqp.im);
    const double svfac = qr < 30.0 ? exp(-2.0 * qr) : 0.0;
    *cosq = cadd(eqp, crmul(svfac, eqm)).re;
    const cpx sinq = csub(eqp, crmul(svfac, eqm));
    *rsinq = cmul(rsv, sinq).re;
    if (fabs(qr) < (double)1.0e-5f && cabs_(rsv) < (double)1.0e-5f)
      *sinqr = zd;
    else
      *sinqr = cdiv(sinq, rsv).re;
  }
}

/* dnka (sregn96.f90:490-646); rho and b are REAL(4) */
static void dnka(double ca[5][5], double cosp, double rsinp, double sinpr, double cossv,
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
  temp = cadd(cadd(cmul(cmul(cx(0.5, 0.0), cx(a0cgg1, 0.0)), cx(gg1, 0.0)),
                   cx(gam2 * gxz2, 0.0)),
              cx(gamm2 * g1wy2, 0.0)).re;
  {
    cpx t = cmul(cneg(cx(2.0, 0.0)), cx(temp, 0.0));
    t = cmul(t, cx((double)rho, 0.0));
    t = cmul(t, cx(om2, 0.0));
    ca[2][0] = cdiv(t, cx(wvno, 0.0)).re;
  }
  {
    const double inner = gam * cqxw2 - gamm1 * cpy;
    const cpx t = cmul(cx(-wvno * inner, 0.0), cx(2.0, 0.0));
    ca[2][1] = t.re;
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

/* evalg with jbdry = 0 (sregn96.f90:648-824): half-space conditions of layer
 * m (0-based), the E and EINV matrices and gbr(1,1:5). */
static void evalg(RModel *M, int m, cpx gbr[5], double wvno, double om, double om2,
                  double wvno2) {
  const double xka = om / M->za[m];
  const double xkb = M->zb[m] > 0.0 ? om / M->zb[m] : 0.0;
  const cpx ra = csqrt_real(wvno2 - xka * xka), rb = csqrt_real(wvno2 - xkb * xkb);
  M->ra = ra;
  M->rb = rb;
  double gam = M->zb[m] * wvno / om;
  gam = 2.0 * (gam * gam);
  const double gamm1 = gam - 1.0;
  const double rho = M->zrho[m];
  cpx (*E)[4] = M->e, (*EI)[4] = M->einv;
  if (M->iwat[m] == 0) {
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
    /* gbr(inp,1..5) and their common divisor, in the reference's order of
     * evaluation (a unary minus applies to the whole product) */
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
  } else {
    if (M->allfluid) {
      gbr[0] = cdiv(cx(0.5, 0), ra);
      gbr[1] = cdiv(cx(0.5, 0), cx(-rho * om2, 0));
      gbr[2] = gbr[3] = gbr[4] = cx(0, 0);
    } else {
      gbr[0] = gbr[1] = gbr[2] = cx(0, 0);
      gbr[3] = cdiv(cx(0.5 * rho * om2, 0), ra);
      gbr[4] = cx(-0.5, 0);
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

/* hska (sregn96.f90:912-986) */
static void hska(double AA[4][4], double cosp, double rsinp, double sinpr, double tcossv,
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

/* normc (sregn96.f90:1398-1427) */
static double normc(double *ee, int n) {
  double t1 = 0.0;
  for (int i = 0; i < n; ++i)
    if (fabs(ee[i]) > t1) t1 = fabs(ee[i]);
  if (t1 < 1.0e-40) t1 = 1.0;
  for (int i = 0; i < n; ++i) ee[i] = ee[i] / t1;
  return log(t1);
}

/* up (sregn96.f90:400-488) */
static double up(RModel *M, double omega, double wvno) {
  const int mmax = M->mmax;
  const double wvno2 = wvno * wvno, om2 = omega * omega;
  cpx gbr[5];
  evalg(M, mmax - 1, gbr, wvno, omega, om2, wvno2);
  for (int j = 0; j < 5; ++j) CD(M, mmax - 1, j) = gbr[j].re;
  M->exe[mmax - 1] = 0.0;
  double exsum = 0.0;
  for (int m = mmax - 2; m >= 0; --m) {
    const double xka = omega / M->za[m];
    const double xkb = M->zb[m] > 0.0 ? omega / M->zb[m] : 0.0;
    const cpx rp = csqrt_real(wvno2 - xka * xka), rsv = csqrt_real(wvno2 - xkb * xkb);
    const cpx p = crmul(M->zd[m], rp), q = crmul(M->zd[m], rsv);
    double cosp, cossv, rsinp, rsinsv, sinpr, sinsvr, pex, svex;
    varsv(p, q, rp, rsv, &cosp, &cossv, &rsinp, &rsinsv, &sinpr, &sinsvr, &pex, &svex,
          M->iwat[m], M->zd[m]);
    double ca[5][5];
    dnka(ca, cosp, rsinp, sinpr, cossv, rsinsv, sinsvr, (float)M->zrho[m], (float)M->zb[m],
         M->iwat[m], pex, pex + svex, wvno, wvno2, om2);
    double ee[5];
    for (int i = 0; i < 5; ++i) {
      double cr = 0.0;
      for (int j = 0; j < 5; ++j) cr = cr + CD(M, m + 1, j) * ca[j][i];
      ee[i] = cr;
    }
    const double exn = normc(ee, 5);
    exsum = exsum + pex + svex + exn;
    M->exe[m] = exsum;
    for (int i = 0; i < 5; ++i) CD(M, m, i) = ee[i];
  }
  return CD(M, 0, 0);
}

/* down (sregn96.f90:988-1058) */
static void down(RModel *M, double omega, double wvno) {
  const int mmax = M->mmax;
  const double om2 = omega * omega, wvno2 = wvno * wvno;
  for (int i = 0; i < 4; ++i) VV(M, 0, i) = i == 0 ? 1.0 : 0.0;
  M->exa[0] = 0.0;
  double exsum = 0.0;
  for (int m = 0; m < mmax - 1; ++m) {
    const double xka = omega / M->za[m];
    const double xkb = M->zb[m] > 0.0 ? omega / M->zb[m] : 0.0;
    const cpx rp = csqrt_real(wvno2 - xka * xka), rsv = csqrt_real(wvno2 - xkb * xkb);
    const cpx p = crmul(M->zd[m], rp), q = crmul(M->zd[m], rsv);
    double cosp, cossv, rsinp, rsinsv, sinpr, sinsvr, pex, svex;
    varsv(p, q, rp, rsv, &cosp, &cossv, &rsinp, &rsinsv, &sinpr, &sinsvr, &pex, &svex,
          M->iwat[m], M->zd[m]);
    double aa[4][4];
    hska(aa, cosp, rsinp, sinpr, cossv, rsinsv, sinsvr, (float)M->zrho[m], (float)M->zb[m],
         M->iwat[m], pex, svex, wvno, wvno2, om2);
    double aa0[4];
    for (int i = 0; i < 4; ++i) {
      double cc = 0.0;
      for (int j = 0; j < 4; ++j) cc = cc + aa[i][j] * VV(M, m, j);
      aa0[i] = cc;
    }
    const double ex2 = normc(aa0, 4);
    exsum = exsum + pex + ex2;
    M->exa[m + 1] = exsum;
    for (int i = 0; i < 4; ++i) VV(M, m + 1, i) = aa0[i];
  }
}

/* svfunc (sregn96.f90:192-398) */
static void svfunc(RModel *M, double omega, double wvno) {
  const int mmax = M->mmax;
  const double fr = up(M, omega, wvno);
  down(M, omega, wvno);
  const double f1213 = -CD(M, 0, 1);
  M->ur[0] = CD(M, 0, 2) / CD(M, 0, 1);
  M->uz[0] = 1.0;
  M->tz[0] = 0.0;
  M->tr[0] = 0.0;
  M->uu0[0] = M->ur[0];
  M->uu0[1] = 1.0;
  M->uu0[2] = fr;
  M->uu0[3] = fr;
  for (int i = 1; i < mmax; ++i) {
    const double cd1 = CD(M, i, 0), cd2 = CD(M, i, 1), cd3 = CD(M, i, 2), cd4 = -CD(M, i, 2);
    const double cd5 = CD(M, i, 3), cd6 = CD(M, i, 4);
    const double tz1 = -VV(M, i, 3), tz2 = -VV(M, i, 2), tz3 = VV(M, i, 1), tz4 = VV(M, i, 0);
    const double uu1 = tz2 * cd6 - tz3 * cd5 + tz4 * cd4;
    const double uu2 = -tz1 * cd6 + tz3 * cd3 - tz4 * cd2;
    const double uu3 = tz1 * cd5 - tz2 * cd3 + tz4 * cd1;
    const double uu4 = -tz1 * cd4 + tz2 * cd2 - tz3 * cd1;
    const double ext = M->exa[i] + M->exe[i] - M->exe[0];
    if (ext > -80.0 && ext < 80.0) {
      const double fact = exp(ext);
      M->ur[i] = uu1 * fact / f1213;
      M->uz[i] = uu2 * fact / f1213;
      M->tz[i] = uu3 * fact / f1213;
      M->tr[i] = uu4 * fact / f1213;
    } else {
      M->ur[i] = M->uz[i] = M->tz[i] = M->tr[i] = 0.0;
    }
  }
  if (!M->allfluid) {
    int jwat = 0;
    for (int i = 0; i < mmax; ++i) {
      if (M->iwat[i] > 0)
        jwat = i + 1;
      else
        break;
    }
    for (int i = 0; i < jwat; ++i) M->ur[i] = M->tr[i] = 0.0;
  }
}

/* getmat (sregn96.f90:1530-1582) */
typedef struct {
  double a12, a14, a21, a23, ah, av, bh, bv, eta, rho, TA, TC, TF, TL, TN;
} Mat;
static Mat getmat(const RModel *M, int m, double wvno, double om) {
  Mat g;
  memset(&g, 0, sizeof g);
  if (M->iwat[m] == 1) {
    g.ah = M->za[m];
    g.av = M->za[m];
    g.bh = 0.0;
    g.bv = 0.0;
    g.rho = M->zrho[m];
    g.eta = 1.0;
    g.TL = 0.0;
    g.TN = 0.0;
    g.TC = M->zrho[m] * M->za[m] * M->za[m];
    g.TA = M->zrho[m] * M->za[m] * M->za[m];
    g.TF = g.TA - 2.0 * g.TN;
    g.a12 = -(wvno * wvno - om * om / (g.ah * g.ah)) / (g.rho * om * om);
  } else {
    g.ah = M->za[m];
    g.av = M->za[m];
    g.bh = M->zb[m];
    g.bv = M->zb[m];
    g.rho = M->zrho[m];
    g.eta = 1.0;
    g.TL = M->zrho[m] * M->zb[m] * M->zb[m];
    g.TN = M->zrho[m] * M->zb[m] * M->zb[m];
    g.TC = M->zrho[m] * M->za[m] * M->za[m];
    g.TA = M->zrho[m] * M->za[m] * M->za[m];
    g.TF = g.TA - 2.0 * g.TN;
    g.a12 = -wvno;
    g.a14 = 1.0 / g.TL;
    g.a21 = wvno * g.TF / g.TC;
    g.a23 = 1.0 / g.TC;
  }
  return g;
}

/* ffunc, gfunc, h1func, h2func (sregn96.f90:1318-1396) */
static cpx ffunc(cpx nub, double dm) {
  if (cabs_(nub) < 1.0e-08) return cx(dm, 0);
  const cpx argcd = crmul(dm, nub);
  const cpx exqq = argcd.re < 40.0 ? cexp_(crmul(-2.0, argcd)) : cx(0, 0);
  return cdiv(csub(cx(1.0, 0), exqq), crmul(2.0, nub));
}
static cpx gfunc(cpx nub, double dm) {
  const cpx argcd = crmul(dm, nub);
  if (argcd.re < 75.0) return crmul(dm, cexp_(cneg(argcd)));
  return cx(0, 0);
}
static cpx h1func(cpx nua, cpx nub, double dm) {
  const cpx s = cadd(nub, nua);
  if (cabs_(s) < 1.0e-08) return cx(dm, 0);
  const cpx argcd = crmul(dm, cadd(nua, nub));
  const cpx exqq = argcd.re < 40.0 ? cexp_(cneg(argcd)) : cx(0, 0);
  return cdiv(csub(cx(1.0, 0), exqq), cadd(nub, nua));
}
static cpx h2func(cpx nua, cpx nub, double dm) {
  if (cabs_(csub(nub, nua)) < 1.0e-08) return cx(dm, 0);
  cpx argcd = crmul(dm, nua);
  const cpx exqp = argcd.re < 40.0 ? cexp_(cneg(argcd)) : cx(0, 0);
  argcd = crmul(dm, nub);
  const cpx exqq = argcd.re < 40.0 ? cexp_(cneg(argcd)) : cx(0, 0);
  return cdiv(csub(exqq, exqp), csub(nua, nub));
}

/* e(i,a)*e(j,b) + e(i,b)*e(j,a) for the 1-based indices of the reference */
static cpx esym(const RModel *M, int i, int j, int a, int b) {
  return cadd(cmul(M->e[i - 1][a - 1], M->e[j - 1][b - 1]),
              cmul(M->e[i - 1][b - 1], M->e[j - 1][a - 1]));
}
static cpx eprod(const RModel *M, int i, int j, int a) {
  return cmul(M->e[i - 1][a - 1], M->e[j - 1][a - 1]);
}

/* intijr (sregn96.f90:1197-1316), i and j 1-based, m 0-based; evalg has set
 * E, EINV, ra, rb of layer m (they depend on m, wvno and om only). */
static double intijr(RModel *M, int i, int j, int m, int typelyr, double om, double om2,
                     double wvno, double wvno2) {
  cpx gbr[5];
  evalg(M, m, gbr, wvno, om, om2, wvno2);
  const cpx ra = M->ra, rb = M->rb;
  cpx (*EI)[4] = M->einv;
  cpx c;
  if (M->iwat[m] == 1) {
    if (typelyr < 0) {
      const cpx kmpu = cadd(crmul(M->uz[m + 1], EI[0][0]), crmul(M->tz[m + 1], EI[0][1]));
      c = cdiv(cmul(cmul(eprod(M, i, j, 1), kmpu), kmpu), crmul(2.0, ra));
    } else if (typelyr == 0) {
      const cpx km1pd = cadd(crmul(M->uz[m], EI[1][0]), crmul(M->tz[m], EI[1][1]));
      const cpx kmpu = cadd(crmul(M->uz[m + 1], EI[0][0]), crmul(M->tz[m + 1], EI[0][1]));
      const cpx FA = ffunc(ra, M->zd[m]), GA = gfunc(ra, M->zd[m]);
      c = cmul(cmul(cmul(eprod(M, i, j, 1), kmpu), kmpu), FA);
      c = cadd(c, cmul(cmul(cmul(esym(M, i, j, 1, 2), kmpu), km1pd), GA));
      c = cadd(c, cmul(cmul(cmul(eprod(M, i, j, 2), km1pd), km1pd), FA));
    } else {
      const cpx km1pd = cadd(crmul(M->uz[m], EI[1][0]), crmul(M->tz[m], EI[1][1]));
      c = cdiv(cmul(cmul(eprod(M, i, j, 2), km1pd), km1pd), crmul(2.0, ra));
    }
  } else {
    /* EINV(r,:) . (ur, uz, tz, tr) at layer index k */
#define KPOT(r, k)                                                                          \
  cadd(cadd(cadd(cmul(EI[r][0], cx(M->ur[k], 0)), cmul(EI[r][1], cx(M->uz[k], 0))),          \
            cmul(EI[r][2], cx(M->tz[k], 0))),                                               \
       cmul(EI[r][3], cx(M->tr[k], 0)))
    if (typelyr < 0) {
      const cpx kmpu = KPOT(0, m), kmsu = KPOT(1, m);
      c = cdiv(cmul(cmul(eprod(M, i, j, 1), kmpu), kmpu), crmul(2.0, ra));
      c = cadd(c, cdiv(cmul(cmul(esym(M, i, j, 1, 2), kmpu), kmsu), cadd(ra, rb)));
      c = cadd(c, cdiv(cmul(cmul(eprod(M, i, j, 2), kmsu), kmsu), crmul(2.0, rb)));
    } else if (typelyr == 0) {
      const cpx km1pd = KPOT(2, m), km1sd = KPOT(3, m);
      const cpx kmpu = KPOT(0, m + 1), kmsu = KPOT(1, m + 1);
      const double d = M->zd[m];
      const cpx FA = ffunc(ra, d), GA = gfunc(ra, d), FB = ffunc(rb, d), GB = gfunc(rb, d);
      const cpx H1 = h1func(ra, rb, d), H2 = h2func(ra, rb, d);
      c = cmul(cmul(cmul(eprod(M, i, j, 1), kmpu), kmpu), FA);
      c = cadd(c, cmul(cmul(cmul(eprod(M, i, j, 3), km1pd), km1pd), FA));
      c = cadd(c, cmul(cmul(cmul(eprod(M, i, j, 2), kmsu), kmsu), FB));
      c = cadd(c, cmul(cmul(cmul(eprod(M, i, j, 4), km1sd), km1sd), FB));
      c = cadd(c, cmul(H1, cadd(cmul(cmul(esym(M, i, j, 1, 2), kmpu), kmsu),
                                cmul(cmul(esym(M, i, j, 3, 4), km1pd), km1sd))));
      c = cadd(c, cmul(H2, cadd(cmul(cmul(esym(M, i, j, 1, 4), kmpu), km1sd),
                                cmul(cmul(esym(M, i, j, 2, 3), km1pd), kmsu))));
      c = cadd(c, cmul(cmul(cmul(GA, esym(M, i, j, 1, 3)), kmpu), km1pd));
      c = cadd(c, cmul(cmul(cmul(GB, esym(M, i, j, 2, 4)), kmsu), km1sd));
    } else {
      const cpx km1pd = KPOT(2, m), km1sd = KPOT(3, m);
      c = cdiv(cmul(cmul(eprod(M, i, j, 3), km1pd), km1pd), crmul(2.0, ra));
      c = cadd(c, cdiv(cmul(cmul(esym(M, i, j, 3, 4), km1pd), km1sd), cadd(ra, rb)));
      c = cadd(c, cdiv(cmul(cmul(eprod(M, i, j, 4), km1sd), km1sd), crmul(2.0, rb)));
    }
#undef KPOT
  }
  return c.re;
}

/* getdcdh (sregn96.f90:1429-1528) */
static void getdcdh(RModel *M, double om2, double wvno, double wvno2, double fac) {
  for (int m = 0; m < M->mmax; ++m) {
    const double tuz = M->uz[m], ttz = M->tz[m], ttr = M->tr[m];
    const double tur = M->iwat[m] == 1 ? -wvno * ttz / (M->zrho[m] * om2) : M->ur[m];
    double gfac1, gfac2, gfac3, gfac4, gfac5, gfac6;
    if (m == 0) {
      const double drho = M->zrho[0] - 0.0, dmu = M->xmu[0] - 0.0, dlm = M->xlam[0] - 0.0;
      const double dl2mu = dlm + dmu + dmu;
      const double xl2mp = M->xlam[m] + M->xmu[m] + M->xmu[m];
      const double duzdzp = (ttz + wvno * M->xlam[m] * tur) / xl2mp;
      const double durdzp = M->iwat[m] == 1 ? wvno * tuz : (ttr / M->xmu[m]) - wvno * tuz;
      const double drur2 = tur * tur * drho, dlur2 = tur * tur * dl2mu;
      gfac1 = om2 * drho * (tuz * tuz);
      gfac2 = om2 * drur2;
      gfac3 = -wvno2 * dmu * (tuz * tuz);
      gfac4 = -wvno2 * dlur2;
      gfac5 = xl2mp * (duzdzp * duzdzp);
      gfac6 = M->xmu[m] * (durdzp * durdzp);
    } else {
      const double drho = M->zrho[m] - M->zrho[m - 1], dmu = M->xmu[m] - M->xmu[m - 1];
      const double dlm = M->xlam[m] - M->xlam[m - 1];
      const double dl2mu = dlm + dmu + dmu;
      const double xl2mp = M->xlam[m] + M->xmu[m] + M->xmu[m];
      const double xl2mm = M->xlam[m - 1] + M->xmu[m - 1] + M->xmu[m - 1];
      const double duzdzp = (ttz + wvno * M->xlam[m] * tur) / xl2mp;
      const double durdzp = M->xmu[m] == 0.0 ? wvno * tuz : (ttr / M->xmu[m]) - wvno * tuz;
      const double durdzm =
          M->xmu[m - 1] == 0.0 ? wvno * tuz : (ttr / M->xmu[m - 1]) - wvno * tuz;
      double drur2, dlur2, duzdzm;
      if (M->iwat[m - 1] == 1 && M->iwat[m] == 0) {
        const double URB = -wvno * M->tz[m] / (M->zrho[m - 1] * om2);
        drur2 = tur * tur * M->zrho[m] - URB * URB * M->zrho[m - 1];
        dlur2 = tur * tur * xl2mp - URB * URB * xl2mm;
        duzdzm = (ttz + wvno * M->xlam[m - 1] * URB) / M->xlam[m - 1];
      } else if (M->iwat[m - 1] == 1 && M->iwat[m] == 1) {
        const double URB = -wvno * M->tz[m] / (M->zrho[m - 1] * om2);
        drur2 = tur * tur * M->zrho[m] - URB * URB * M->zrho[m - 1];
        dlur2 = tur * tur * xl2mp - URB * URB * xl2mm;
        duzdzm = (ttz + wvno * M->xlam[m - 1] * URB) / xl2mm;
      } else {
        drur2 = tur * tur * drho;
        dlur2 = tur * tur * dl2mu;
        duzdzm = (ttz + wvno * M->xlam[m - 1] * tur) / xl2mm;
      }
      gfac1 = om2 * drho * (tuz * tuz);
      gfac2 = om2 * drur2;
      gfac3 = -wvno2 * dmu * (tuz * tuz);
      gfac4 = -wvno2 * dlur2;
      gfac5 = xl2mp * (duzdzp * duzdzp) - xl2mm * (duzdzm * duzdzm);
      gfac6 = M->xmu[m] * (durdzp * durdzp) - M->xmu[m - 1] * (durdzm * durdzm);
    }
    double dfac = fac * (gfac1 + gfac2 + gfac3 + gfac4 + gfac5 + gfac6);
    if (fabs(dfac) < 1.0e-38) dfac = 0.0;
    M->dcdh[m] = dfac;
  }
}

/* energy (sregn96.f90:1060-1195) */
static void energy(RModel *M, double om, double wvno) {
  const int mmax = M->mmax;
  M->sumi0 = M->sumi1 = M->sumi2 = M->sumi3 = 0.0;
  const double c = om / wvno, om2 = om * om, wvno2 = wvno * wvno;
  for (int m = 0; m < mmax; ++m) {
    const Mat g = getmat(M, m, wvno, om);
    const int typelyr = m == mmax - 1 ? 1 : 0;
    const double INT11 = intijr(M, 1, 1, m, typelyr, om, om2, wvno, wvno2);
    const double INT13 = intijr(M, 1, 3, m, typelyr, om, om2, wvno, wvno2);
    const double INT22 = intijr(M, 2, 2, m, typelyr, om, om2, wvno, wvno2);
    const double INT24 = intijr(M, 2, 4, m, typelyr, om, om2, wvno, wvno2);
    const double INT33 = intijr(M, 3, 3, m, typelyr, om, om2, wvno, wvno2);
    const double INT44 = intijr(M, 4, 4, m, typelyr, om, om2, wvno, wvno2);
    (void)INT33;
    (void)INT44;
    if (M->iwat[m] == 1) {
      const double w = wvno / (g.rho * om2);
      const double URUR = INT22 * (w * w), UZUZ = INT11;
      const double URDUZ = -(wvno / (g.rho * om2)) * g.a12 * INT22;
      const double DUZDUZ = g.a12 * g.a12 * INT22;
      M->sumi0 = M->sumi0 + g.rho * (URUR + UZUZ);
      M->sumi1 = M->sumi1 + g.TA * URUR;
      M->sumi2 = M->sumi2 - g.TF * URDUZ;
      M->sumi3 = M->sumi3 + g.TC * DUZDUZ;
      const double facah = g.rho * g.ah * (URUR - 2.0 * g.eta * URDUZ / wvno);
      const double facav = g.rho * g.av * DUZDUZ / wvno2;
      M->dcda[m] = facah + facav;
      const double facr = -0.5 * c * c * (URUR + UZUZ);
      M->dcdr[m] = 0.5 * (g.av * facav + g.ah * facah) / g.rho + facr;
      M->ur[m] = -wvno * M->tz[m] / (g.rho * om2);
    } else {
      const double URUR = INT11, UZUZ = INT22;
      const double DURDUR = g.a12 * g.a12 * INT22 + 2.0 * g.a12 * g.a14 * INT24 +
                            g.a14 * g.a14 * INT44;
      const double DUZDUZ = g.a21 * g.a21 * INT11 + 2.0 * g.a21 * g.a23 * INT13 +
                            g.a23 * g.a23 * INT33;
      const double URDUZ = g.a21 * INT11 + g.a23 * INT13;
      const double UZDUR = g.a12 * INT22 + g.a14 * INT24;
      M->sumi0 = M->sumi0 + g.rho * (URUR + UZUZ);
      M->sumi1 = M->sumi1 + g.TL * UZUZ + g.TA * URUR;
      M->sumi2 = M->sumi2 + g.TL * UZDUR - g.TF * URDUZ;
      M->sumi3 = M->sumi3 + g.TL * DURDUR + g.TC * DUZDUZ;
      const double facah = g.rho * g.ah * (URUR - 2.0 * g.eta * URDUZ / wvno);
      const double facav = g.rho * g.av * DUZDUZ / wvno2;
      const double facbh = 0.0;
      const double facbv = g.rho * g.bv *
                           (UZUZ + 2.0 * UZDUR / wvno + DURDUR / wvno2 +
                            4.0 * g.eta * URDUZ / wvno);
      M->dcda[m] = facah + facav;
      M->dcdb[m] = facbv + facbh;
      const double facr = -0.5 * c * c * (URUR + UZUZ);
      M->dcdr[m] = 0.5 * (g.av * facav + g.ah * facah + g.bv * facbv) / g.rho + facr;
    }
  }
  M->flagr = om2 * M->sumi0 - wvno2 * M->sumi1 - 2.0 * wvno * M->sumi2 - M->sumi3;
  M->ugr = (wvno * M->sumi1 + M->sumi2) / (om * M->sumi0);
  M->are = wvno / (2.0 * om * M->ugr * M->sumi0);
  const double fac = M->are * c / wvno2;
  for (int m = 0; m < mmax; ++m) {
    M->dcda[m] = M->dcda[m] / (M->ugr * M->sumi0);
    M->dcdb[m] = M->dcdb[m] / (M->ugr * M->sumi0);
    M->dcdr[m] = M->dcdr[m] / (M->ugr * M->sumi0);
  }
  getdcdh(M, om2, wvno, wvno2, fac);
}

/* sregn96 (sregn96.f90:1630-1731) with iflsph as given; thk, vp, vs, rhom
 * REAL(4) (nlayer); outputs REAL(8) (nlayer).  *cp in: the phase velocity;
 * out: the spherical one (csph); *cg: usph. */
void orc_sregn96(const float *thk, const float *vp, const float *vs, const float *rhom,
                 int nlayer, double t, double *cp, double *cg, double *dispu, double *dispw,
                 double *stressu, double *stressw, double *dc2da, double *dc2db, double *dc2dh,
                 double *dc2dr, int iflsph) {
  RModel Mo;
  RModel *M = &Mo;
  alloc_model(M, nlayer);
  const int mmax = nlayer;
  for (int i = 0; i < mmax; ++i) {
    M->zb[i] = vs[i];
    M->za[i] = vp[i];
    M->zrho[i] = rhom[i];
    M->zd[i] = thk[i];
  }
  M->allfluid = 1;
  for (int i = 0; i < mmax; ++i) {
    if (M->zb[i] > 0.0) {
      M->allfluid = 0;
      M->iwat[i] = 0;
    } else {
      M->iwat[i] = 1;
    }
  }
  const int nsph = iflsph;
  if (nsph > 0) bldsph(M);
  for (int i = 0; i < mmax; ++i) {
    M->xmu[i] = M->zrho[i] * (M->zb[i] * M->zb[i]);
    M->xlam[i] = M->zrho[i] * (M->za[i] * M->za[i]) - 2.0 * M->xmu[i];
  }
  const double twopi = 2.0 * kPi;
  const double omega = twopi / t;
  double c = *cp;
  M->ugr = c;
  double wvno = omega / c;
  svfunc(M, omega, wvno);
  energy(M, omega, wvno);
  double csph, usph;
  if (nsph > 0) {
    /* sprayl (sregn96.f90:1584-1628) */
    const double ar = 6371.0;
    const double r = c / (2.0 * ar * omega);
    const double tm = sqrt(1.0 + r * r);
    const double tm3 = tm * tm * tm;
    for (int i = 0; i < mmax; ++i) {
      M->dcda[i] = M->dcda[i] * M->vtp[i] / tm3;
      M->dcdb[i] = M->dcdb[i] * M->vtp[i] / tm3;
      M->dcdh[i] = M->dcdh[i] * M->dtp[i] / tm3;
      M->dcdr[i] = M->dcdr[i] * M->rtp[i] / tm3;
    }
    usph = M->ugr * tm;
    csph = c / tm;
    wvno = omega / csph;
  } else {
    csph = c;
    usph = M->ugr;
  }
  for (int i = 0; i < mmax - 1; ++i) {
    double sums = 0.0;
    for (int k = i + 1; k < mmax; ++k) sums = sums + M->dcdh[k];
    M->dcdh[i] = sums;
  }
  M->dcdh[mmax - 1] = 0.0;
  for (int i = 0; i < mmax; ++i) {
    if (dc2da) dc2da[i] = M->dcda[i];
    if (dc2db) dc2db[i] = M->dcdb[i];
    if (dc2dr) dc2dr[i] = M->dcdr[i];
    if (dc2dh) dc2dh[i] = M->dcdh[i];
    if (dispu) dispu[i] = M->ur[i];
    if (dispw) dispw[i] = M->uz[i];
    if (stressu) stressu[i] = M->tr[i];
    if (stressw) stressw[i] = M->tz[i];
  }
  *cp = csph;
  *cg = usph;
  free_model(M);
}

/* get_vprho (surfker.f90:309-321), REAL(4) */
static void vprho(float vs, float *vp, float *rho) {
  const float v2 = vs * vs, v3 = vs * v2, v4 = v2 * v2;
  *vp = 0.9409f + 2.0947f * vs - 0.8206f * v2 + 0.2683f * v3 - 0.0251f * v4;
  const float p = *vp, p2 = p * p, p3 = p * p2, p4 = p2 * p2, p5 = p * p4;
  *rho = 1.6612f * p - 0.4721f * p2 + 0.0671f * p3 - 0.0043f * p4 + 0.000106f * p5;
}

/* depthkernel1d's model (surfker.f90:88-91): Vs to REAL(4), get_vprho and
 * refinegrid (:285-307): nz layers and a half-space copy of the last,
 * the last layer's thickness repeated.  Outputs have nz + 1 entries. */
void orc_refinegrid(const double *vel, const double *depz, int nz, float *rthk, float *rvp,
                    float *rvs, float *rrho) {
  for (int k = 0; k < nz; ++k) {
    const float vs = (float)vel[k];
    float vp, rho;
    vprho(vs, &vp, &rho);
    rvs[k] = vs;
    rvp[k] = vp;
    rrho[k] = rho;
  }
  for (int k = 0; k < nz; ++k) {
    if (k == nz - 1)
      rthk[k] = rthk[k - 1];
    else
      rthk[k] = (float)depz[k + 1] - (float)depz[k];
  }
  rthk[nz] = 0.0f;
  rvp[nz] = rvp[nz - 1];
  rvs[nz] = rvs[nz - 1];
  rrho[nz] = rrho[nz - 1];
}

/* The Rayleigh phase-velocity branch of depthkernel1d (surfker.f90:93-103)
 * for one column, given the phase velocities cp(kmax) that surfdisp96
 * finds on the refinegrid model: sen_* (kmax, nz), Fortran order. */
int orc_depthkernel_rayleigh(const double *vel, const double *depz, int nz, const double *tRc,
                             const double *cp, int kmax, double *sen_vs, double *sen_vp,
                             double *sen_rho) {
  if (nz < 2 || kmax < 1) return -1;
  const int mmax = nz + 1;
  float *f = (float *)malloc(sizeof(float) * 4 * (size_t)mmax);
  double *d = (double *)calloc(4 * (size_t)mmax, sizeof(double));
  float *rthk = f, *rvp = f + mmax, *rvs = f + 2 * mmax, *rrho = f + 3 * mmax;
  double *dcdar = d, *dcdbr = d + mmax, *dcdhr = d + 2 * mmax, *dcdrr = d + 3 * mmax;
  orc_refinegrid(vel, depz, nz, rthk, rvp, rvs, rrho);
  for (int i = 0; i < kmax; ++i) {
    double c = cp[i], cg = 0.0;
    orc_sregn96(rthk, rvp, rvs, rrho, mmax, tRc[i], &c, &cg, NULL, NULL, NULL, NULL, dcdar,
                dcdbr, dcdhr, dcdrr, 1);
    for (int k = 0; k < nz; ++k) {
      sen_vs[i + (size_t)kmax * k] = dcdbr[k];
      sen_vp[i + (size_t)kmax * k] = dcdar[k];
      sen_rho[i + (size_t)kmax * k] = dcdrr[k];
    }
  }
  free(f);
  free(d);
  return 0;
}
