/*
 * swd_oracle_ray.c -- CPU restatement of src/SWD/sregn96.f90 (Rayleigh
 * eigenfunctions, energy integrals, group velocity and layer kernels).
 * TEST INFRASTRUCTURE ONLY (the checker); nothing in the product links it.
 *
 * The Fortran module state (sregn96.f90:54-104) becomes a per-call struct.
 * Quirks kept: pi is the REAL literal 3.1415926535898 (float), dnka/hska get
 * sngl(rho), sngl(b), 2.0*b*b is formed in float, rtp = tmp**(-2.275) uses
 * the REAL exponent, intijr re-evaluates evalg for each integral.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "swd_oracle.h"

typedef double complex cplx;

typedef struct {
  int mmax;
  double *zd, *zrho, *za, *zb, *xmu, *xlam;
  double *ur, *uz, *tz, *tr, *dcda, *dcdb, *dcdr, *dcdh;
  int *iwat;
  double sumi0, sumi1, sumi2, sumi3, flagr, are, ugr;
  double *vtp, *dtp, *rtp;
  double *exe, *exa, *cd /* mmax x 5 */, *vv /* mmax x 4 */;
  int allfluid;
  cplx ra, rb, e[4][4], einv[4][4];
} rctx;

#define CD(i, j) x->cd[(i) * 5 + (j)]
#define VV(i, j) x->vv[(i) * 4 + (j)]

static void ralloc(rctx *x, int n) {
  memset(x, 0, sizeof(*x));
  x->mmax = n;
  double *p = (double *)calloc((size_t)n * 28, sizeof(double));
  x->zd = p; x->zrho = p + n; x->za = p + 2 * n; x->zb = p + 3 * n;
  x->xmu = p + 4 * n; x->xlam = p + 5 * n; x->ur = p + 6 * n; x->uz = p + 7 * n;
  x->tz = p + 8 * n; x->tr = p + 9 * n; x->dcda = p + 10 * n; x->dcdb = p + 11 * n;
  x->dcdr = p + 12 * n; x->dcdh = p + 13 * n; x->vtp = p + 14 * n;
  x->dtp = p + 15 * n; x->rtp = p + 16 * n; x->exe = p + 17 * n;
  x->exa = p + 18 * n; x->cd = p + 19 * n; /* 5n */
  x->vv = p + 24 * n;                      /* 4n */
  x->iwat = (int *)calloc((size_t)n, sizeof(int));
}
static void rfree(rctx *x) {
  free(x->zd);
  free(x->iwat);
}

/* sregn96.f90:130-181 bldsph, R = 6371 km */
static void bldsph_r(rctx *x) {
  const double ar = 6371.0;
  double dr = 0.0, r0 = ar;
  int mmax = x->mmax;
  for (int i = 0; i < mmax; ++i) x->dtp[i] = 0.0;
  for (int i = 0; i < mmax; ++i) {
    if (i == mmax - 1)
      dr = dr + 1.0;
    else
      dr = dr + x->zd[i];
    double r1 = ar - dr;
    double z0 = ar * log(ar / r0);
    double z1 = ar * log(ar / r1);
    double tmp = (2.0 * ar) / (r0 + r1);
    x->vtp[i] = tmp;
    x->rtp[i] = pow(tmp, (double)(-2.275f));
    x->dtp[i] = ar / r0;
    x->za[i] = x->za[i] * tmp;
    x->zb[i] = x->zb[i] * tmp;
    x->zrho[i] = x->zrho[i] * x->rtp[i];
    x->zd[i] = z1 - z0;
    r0 = r1;
  }
  x->zd[mmax - 1] = 0.0;
}

/* sregn96.f90:827-911 varsv */
static void varsv(cplx p, cplx q, cplx rp, cplx rsv, double *cosp,
                  double *cosq, double *rsinp, double *rsinq, double *sinpr,
                  double *sinqr, double *pex, double *svex, int iwat,
                  double zd) {
  double pr = creal(p), pi = cimag(p), qr = creal(q), qi = cimag(q);
  *pex = pr;
  *svex = 0.0;
  const double tiny = (double)1.0e-5f;
  cplx epp = CMPLX(cos(pi), sin(pi)) / 2.0;
  cplx epm = conj(epp);
  double pfac = (pr < 30.) ? exp(-2. * pr) : 0.0;
  *cosp = creal(epp + pfac * epm);
  cplx sinp = epp - pfac * epm;
  *rsinp = creal(rp * sinp);
  if (fabs(pr) < tiny && cabs(rp) < tiny)
    *sinpr = zd;
  else
    *sinpr = creal(sinp / rp);
  if (iwat == 1) {
    *cosq = 1.0;
    *rsinq = 0.0;
    *sinqr = 0.0;
    return;
  }
  *svex = qr;
  cplx eqp = CMPLX(cos(qi), sin(qi)) / 2.0;
  cplx eqm = conj(eqp);
  double svfac = (qr < 30.) ? exp(-2. * qr) : 0.0;
  *cosq = creal(eqp + svfac * eqm);
  cplx sinq = eqp - svfac * eqm;
  *rsinq = creal(rsv * sinq);
  if (fabs(qr) < tiny && cabs(rsv) < tiny)
    *sinqr = zd;
  else
    *sinqr = creal(sinq / rsv);
}

/* sregn96.f90:491-647 dnka (rho, b arrive as float) */
static void dnka_r(double ca[5][5], double cosp, double rsinp, double sinpr,
                   double cossv, double rsinsv, double sinsvr, float rho,
                   float b, int iwat, double ex, double exa, double wvno,
                   double wvno2, double om2) {
  if (iwat == 1) {
    for (int j = 0; j < 5; ++j)
      for (int i = 0; i < 5; ++i) ca[i][j] = 0.0;
    double dfac = (ex > 35.0) ? 0.0 : exp(-ex);
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
  double a0 = (exa < 60.0) ? exp(-exa) : 0.0;
  double cpcq = cosp * cossv;
  double cpy = cosp * sinsvr;
  double cpz = cosp * rsinsv;
  double cqw = cossv * sinpr;
  double cqx = cossv * rsinp;
  double xy = rsinp * sinsvr;
  double xz = rsinp * rsinsv;
  double wy = sinpr * sinsvr;
  double wz = sinpr * rsinsv;
  float rho2 = rho * rho;
  float bb2 = 2.0f * b * b;
  double gam = (double)bb2 * wvno2 / om2;
  double gam2 = gam * gam;
  double gamm1 = gam - 1.;
  double gamm2 = gamm1 * gamm1;
  double cqww2 = cqw * wvno2;
  double cqxw2 = cqx / wvno2;
  double gg1 = gam * gamm1;
  double a0c = 2.0 * (a0 - cpcq);
  double xz2 = xz / wvno2;
  double gxz2 = gam * xz2;
  double g2xz2 = gam2 * xz2;
  double a0cgg1 = a0c * (gam + gamm1);
  double wy2 = wy * wvno2;
  double g2wy2 = gamm2 * wy2;
  double g1wy2 = gamm1 * wy2;
  double rom = (double)rho * om2;
  double temp = a0c * gg1 + g2xz2 + g2wy2;
  ca[2][2] = a0 + temp + temp;
  ca[0][0] = cpcq - temp;
  ca[0][1] = (-cqx + wvno2 * cpy) / rom;
  temp = 0.5 * a0cgg1 + gxz2 + g1wy2;
  ca[0][2] = wvno * temp / rom;
  ca[0][3] = (-cqww2 + cpz) / rom;
  temp = wvno2 * (a0c + wy2) + xz;
  ca[0][4] = -temp / ((double)rho2 * om2 * om2);
  ca[1][0] = (-gamm2 * cqw + gam2 * cpz / wvno2) * (double)rho * om2;
  ca[1][1] = cpcq;
  ca[1][2] = (gamm1 * cqww2 - gam * cpz) / wvno;
  ca[1][3] = -wz;
  ca[1][4] = ca[0][3];
  temp = 0.5 * a0cgg1 * gg1 + gam2 * gxz2 + gamm2 * g1wy2;
  ca[2][0] = -2.0 * temp * (double)rho * om2 / wvno;
  ca[2][1] = -wvno * (gam * cqxw2 - gamm1 * cpy) * 2.0;
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

/* sregn96.f90:649-825 evalg, jbdry = 0 (the only form the path uses) */
static void evalg0(rctx *x, int m, cplx gbr[5], double wvno, double om,
                   double om2, double wvno2) {
  double xka = om / x->za[m];
  double xkb = (x->zb[m] > 0.0) ? om / x->zb[m] : 0.0;
  x->ra = csqrt(CMPLX(wvno2 - xka * xka, 0.0));
  x->rb = csqrt(CMPLX(wvno2 - xkb * xkb, 0.0));
  cplx ra = x->ra, rb = x->rb;
  double gam = x->zb[m] * wvno / om;
  gam = 2.0 * (gam * gam);
  double gamm1 = gam - 1.0;
  double zr = x->zrho[m];
  if (x->iwat[m] == 0) {
    cplx (*E)[4] = x->e, (*EI)[4] = x->einv;
    E[0][0] = wvno; E[0][1] = rb; E[0][2] = wvno; E[0][3] = -rb;
    E[1][0] = ra; E[1][1] = wvno; E[1][2] = -ra; E[1][3] = wvno;
    E[2][0] = zr * om2 * gamm1;
    E[2][1] = zr * om2 * gam * rb / wvno;
    E[2][2] = zr * om2 * gamm1;
    E[2][3] = -zr * om2 * gam * rb / wvno;
    E[3][0] = zr * om2 * gam * ra / wvno;
    E[3][1] = zr * om2 * gamm1;
    E[3][2] = -zr * om2 * gam * ra / wvno;
    E[3][3] = zr * om2 * gamm1;
    EI[0][0] = 0.5 * gam / wvno;
    EI[0][1] = -0.5 * gamm1 / ra;
    EI[0][2] = -0.5 / (zr * om2);
    EI[0][3] = 0.5 * wvno / (zr * om2 * ra);
    EI[1][0] = -0.5 * gamm1 / rb;
    EI[1][1] = 0.5 * gam / wvno;
    EI[1][2] = 0.5 * wvno / (zr * om2 * rb);
    EI[1][3] = -0.5 / (zr * om2);
    EI[2][0] = 0.5 * gam / wvno;
    EI[2][1] = 0.5 * gamm1 / ra;
    EI[2][2] = -0.5 / (zr * om2);
    EI[2][3] = -0.5 * wvno / (zr * om2 * ra);
    EI[3][0] = 0.5 * gamm1 / rb;
    EI[3][1] = 0.5 * gam / wvno;
    EI[3][2] = -0.5 * wvno / (zr * om2 * rb);
    EI[3][3] = -0.5 / (zr * om2);
    if (gbr) {
      gbr[0] = (zr * zr) * om2 * om2 * (-gam * gam * ra * rb + wvno2 * gamm1 * gamm1);
      gbr[1] = -zr * (wvno2 * ra) * om2;
      gbr[2] = -zr * (-gam * ra * rb + wvno2 * gamm1) * om2 * wvno;
      gbr[3] = zr * (wvno2 * rb) * om2;
      gbr[4] = wvno2 * (wvno2 - ra * rb);
      for (int i = 0; i < 5; ++i)
        gbr[i] = 0.25 * gbr[i] / (-zr * zr * om2 * om2 * wvno2 * ra * rb);
    }
  } else {
    if (gbr) {
      if (x->allfluid) {
        gbr[0] = 0.5 / ra;
        gbr[1] = CMPLX(0.5, 0.0) / (-zr * om2);
        gbr[2] = 0.0; gbr[3] = 0.0; gbr[4] = 0.0;
      } else {
        gbr[0] = 0.0; gbr[1] = 0.0; gbr[2] = 0.0;
        gbr[3] = (0.5 * zr * om2) / ra;
        gbr[4] = -0.5;
      }
    }
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) x->e[i][j] = x->einv[i][j] = 0.0;
    x->e[0][0] = ra;
    x->e[0][1] = -ra;
    x->e[1][0] = -zr * om2;
    x->e[1][1] = -zr * om2;
    x->einv[0][0] = 0.5 / ra;
    x->einv[0][1] = -0.5 / (zr * om2);
    x->einv[1][0] = -0.5 / ra;
    x->einv[1][1] = -0.5 / (zr * om2);
  }
}

static void normc_n(double *ee, double *ex, int n) {
  double t1 = 0.0;
  for (int i = 0; i < n; ++i)
    if (fabs(ee[i]) > t1) t1 = fabs(ee[i]);
  if (t1 < 1.0e-40) t1 = 1.0;
  for (int i = 0; i < n; ++i) ee[i] = ee[i] / t1;
  *ex = log(t1);
}

/* sregn96.f90:401-489 up: Dunkin vectors from the half-space upward */
static double up_r(rctx *x, double omega, double wvno) {
  int mmax = x->mmax;
  double wvno2 = wvno * wvno, om2 = omega * omega;
  cplx gbr[5];
  evalg0(x, mmax - 1, gbr, wvno, omega, om2, wvno2);
  for (int j = 0; j < 5; ++j) CD(mmax - 1, j) = creal(gbr[j]);
  x->exe[mmax - 1] = 0.0;
  double exsum = 0.0;
  for (int m = mmax - 2; m >= 0; --m) {
    double xka = omega / x->za[m];
    double xkb = (x->zb[m] > 0.0) ? omega / x->zb[m] : 0.0;
    cplx rp = csqrt(CMPLX(wvno2 - xka * xka, 0.0));
    cplx rsv = csqrt(CMPLX(wvno2 - xkb * xkb, 0.0));
    cplx p = rp * x->zd[m], q = rsv * x->zd[m];
    double cosp, cossv, rsinp, rsinsv, sinpr, sinsvr, pex, svex;
    varsv(p, q, rp, rsv, &cosp, &cossv, &rsinp, &rsinsv, &sinpr, &sinsvr,
          &pex, &svex, x->iwat[m], x->zd[m]);
    double ca[5][5], ee[5];
    dnka_r(ca, cosp, rsinp, sinpr, cossv, rsinsv, sinsvr, (float)x->zrho[m],
           (float)x->zb[m], x->iwat[m], pex, pex + svex, wvno, wvno2, om2);
    for (int i = 0; i < 5; ++i) {
      double cr = 0.0;
      for (int j = 0; j < 5; ++j) cr = cr + CD(m + 1, j) * ca[j][i];
      ee[i] = cr;
    }
    double exn = 0.0;
    normc_n(ee, &exn, 5);
    exsum = exsum + pex + svex + exn;
    x->exe[m] = exsum;
    for (int i = 0; i < 5; ++i) CD(m, i) = ee[i];
  }
  return CD(0, 0);
}

/* sregn96.f90:913-987 hska (rho, b arrive as float) */
static void hska(double AA[4][4], double cosp, double rsinp, double sinpr,
                 double tcossv, double trsinsv, double tsinsvr, float rho,
                 float b, int iwat, double pex, double svex, double wvno,
                 double wvno2, double om2) {
  if (iwat == 1) {
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) AA[i][j] = 0.0;
    double dfac = (pex > 35.0) ? 0.0 : exp(-pex);
    AA[0][0] = dfac;
    AA[3][3] = dfac;
    AA[1][1] = cosp;
    AA[2][2] = cosp;
    AA[1][2] = -rsinp / ((double)rho * om2);
    AA[2][1] = -(double)rho * om2 * sinpr;
    return;
  }
  double dfac = ((pex - svex) > 70.0) ? 0.0 : exp(svex - pex);
  double cossv = dfac * tcossv;
  double rsinsv = dfac * trsinsv;
  double sinsvr = dfac * tsinsvr;
  float bb2 = 2.0f * b * b;
  double gam = (double)bb2 * wvno2 / om2;
  double gamm1 = gam - 1.0;
  double rom = (double)rho * om2;
  AA[0][0] = cossv + gam * (cosp - cossv);
  AA[0][1] = -wvno * gamm1 * sinpr + gam * rsinsv / wvno;
  AA[0][2] = -wvno * (cosp - cossv) / rom;
  AA[0][3] = (wvno2 * sinpr - rsinsv) / rom;
  AA[1][0] = gam * rsinp / wvno - wvno * gamm1 * sinsvr;
  AA[1][1] = cosp - gam * (cosp - cossv);
  AA[1][2] = (-rsinp + wvno2 * sinsvr) / rom;
  AA[1][3] = -AA[0][2];
  AA[2][0] = rom * gam * gamm1 * (cosp - cossv) / wvno;
  AA[2][1] = rom * (-gamm1 * gamm1 * sinpr + gam * gam * rsinsv / wvno2);
  AA[2][2] = AA[1][1];
  AA[2][3] = -AA[0][1];
  AA[3][0] = rom * (gam * gam * rsinp / wvno2 - gamm1 * gamm1 * sinsvr);
  AA[3][1] = -AA[2][0];
  AA[3][2] = -AA[1][0];
  AA[3][3] = AA[0][0];
}

/* sregn96.f90:989-1059 down: Haskell first column from the surface down */
static void down_r(rctx *x, double omega, double wvno) {
  int mmax = x->mmax;
  double om2 = omega * omega, wvno2 = wvno * wvno;
  for (int i = 0; i < 4; ++i) VV(0, i) = (i == 0) ? 1.0 : 0.0;
  x->exa[0] = 0.0;
  double exsum = 0.0;
  for (int m = 0; m < mmax - 1; ++m) {
    double xka = omega / x->za[m];
    double xkb = (x->zb[m] > 0.0) ? omega / x->zb[m] : 0.0;
    cplx rp = csqrt(CMPLX(wvno2 - xka * xka, 0.0));
    cplx rsv = csqrt(CMPLX(wvno2 - xkb * xkb, 0.0));
    cplx p = rp * x->zd[m], q = rsv * x->zd[m];
    double cosp, cossv, rsinp, rsinsv, sinpr, sinsvr, pex, svex;
    varsv(p, q, rp, rsv, &cosp, &cossv, &rsinp, &rsinsv, &sinpr, &sinsvr,
          &pex, &svex, x->iwat[m], x->zd[m]);
    double AA[4][4], aa0[4];
    hska(AA, cosp, rsinp, sinpr, cossv, rsinsv, sinsvr, (float)x->zrho[m],
         (float)x->zb[m], x->iwat[m], pex, svex, wvno, wvno2, om2);
    for (int i = 0; i < 4; ++i) {
      double cc = 0.0;
      for (int j = 0; j < 4; ++j) cc = cc + AA[i][j] * VV(m, j);
      aa0[i] = cc;
    }
    double ex2 = 0.0;
    normc_n(aa0, &ex2, 4);
    exsum = exsum + pex + ex2;
    x->exa[m + 1] = exsum;
    for (int i = 0; i < 4; ++i) VV(m + 1, i) = aa0[i];
  }
}

/* sregn96.f90:193-399 svfunc */
static void svfunc(rctx *x, double omega, double wvno) {
  int mmax = x->mmax;
  up_r(x, omega, wvno);
  down_r(x, omega, wvno);
  double f1213 = -CD(0, 1);
  x->ur[0] = CD(0, 2) / CD(0, 1);
  x->uz[0] = 1.0;
  x->tz[0] = 0.0;
  x->tr[0] = 0.0;
  for (int i = 1; i < mmax; ++i) {
    double cd1 = CD(i, 0), cd2 = CD(i, 1), cd3 = CD(i, 2), cd4 = -CD(i, 2);
    double cd5 = CD(i, 3), cd6 = CD(i, 4);
    double tz1 = -VV(i, 3), tz2 = -VV(i, 2), tz3 = VV(i, 1), tz4 = VV(i, 0);
    double uu1 = tz2 * cd6 - tz3 * cd5 + tz4 * cd4;
    double uu2 = -tz1 * cd6 + tz3 * cd3 - tz4 * cd2;
    double uu3 = tz1 * cd5 - tz2 * cd3 + tz4 * cd1;
    double uu4 = -tz1 * cd4 + tz2 * cd2 - tz3 * cd1;
    double ext = x->exa[i] + x->exe[i] - x->exe[0];
    if (ext > -80.0 && ext < 80.0) {
      double fact = exp(ext);
      x->ur[i] = uu1 * fact / f1213;
      x->uz[i] = uu2 * fact / f1213;
      x->tz[i] = uu3 * fact / f1213;
      x->tr[i] = uu4 * fact / f1213;
    } else {
      x->ur[i] = x->uz[i] = x->tz[i] = x->tr[i] = 0.0;
    }
  }
  if (!x->allfluid) {
    int jwat = 0;
    for (int i = 0; i < mmax; ++i) {
      if (x->iwat[i] > 0)
        jwat = i + 1;
      else
        break;
    }
    for (int i = 0; i < jwat; ++i) {
      x->ur[i] = 0.0;
      x->tr[i] = 0.0;
    }
  }
}

/* sregn96.f90:1319-1397 ffunc/gfunc/h1func/h2func */
static cplx ffunc(cplx nub, double dm) {
  if (cabs(nub) < 1.0e-08) return dm;
  cplx argcd = nub * dm;
  cplx exqq = (creal(argcd) < 40.0) ? cexp(-2.0 * argcd) : 0.0;
  return (1.0 - exqq) / (2.0 * nub);
}
static cplx gfunc(cplx nub, double dm) {
  cplx argcd = nub * dm;
  if (creal(argcd) < 75) return cexp(-argcd) * dm;
  return 0.0;
}
static cplx h1func(cplx nua, cplx nub, double dm) {
  if (cabs(nub + nua) < 1.0e-08) return dm;
  cplx argcd = (nua + nub) * dm;
  cplx exqq = (creal(argcd) < 40.0) ? cexp(-argcd) : 0.0;
  return (1.0 - exqq) / (nub + nua);
}
static cplx h2func(cplx nua, cplx nub, double dm) {
  if (cabs(nub - nua) < 1.0e-08) return dm;
  cplx argcd = nua * dm;
  cplx exqp = (creal(argcd) < 40.0) ? cexp(-argcd) : 0.0;
  argcd = nub * dm;
  cplx exqq = (creal(argcd) < 40.0) ? cexp(-argcd) : 0.0;
  return (exqq - exqp) / (nua - nub);
}

/* sregn96.f90:1198-1317 intijr (typelyr 0: internal layer, 1: half-space);
 * i, j are 1-based as in the Fortran */
static double intijr(rctx *x, int i, int j, int m, int typelyr, double om,
                     double om2, double wvno, double wvno2) {
  evalg0(x, m, NULL, wvno, om, om2, wvno2);
  cplx (*e)[4] = x->e, (*ei)[4] = x->einv;
  cplx ra = x->ra, rb = x->rb, c;
  --i;
  --j;
  double *ur = x->ur, *uz = x->uz, *tz = x->tz, *tr = x->tr;
  if (x->iwat[m] == 1) {
    if (typelyr == 0) {
      cplx km1pd = ei[1][0] * uz[m] + ei[1][1] * tz[m];
      cplx kmpu = ei[0][0] * uz[m + 1] + ei[0][1] * tz[m + 1];
      cplx FA = ffunc(ra, x->zd[m]);
      cplx GA = gfunc(ra, x->zd[m]);
      c = e[i][0] * e[j][0] * kmpu * kmpu * FA +
          (e[i][0] * e[j][1] + e[i][1] * e[j][0]) * kmpu * km1pd * GA +
          e[i][1] * e[j][1] * km1pd * km1pd * FA;
    } else {
      cplx km1pd = ei[1][0] * uz[m] + ei[1][1] * tz[m];
      c = e[i][1] * e[j][1] * km1pd * km1pd / (2.0 * ra);
    }
  } else {
    if (typelyr == 0) {
      cplx km1pd = ei[2][0] * ur[m] + ei[2][1] * uz[m] + ei[2][2] * tz[m] + ei[2][3] * tr[m];
      cplx km1sd = ei[3][0] * ur[m] + ei[3][1] * uz[m] + ei[3][2] * tz[m] + ei[3][3] * tr[m];
      cplx kmpu = ei[0][0] * ur[m + 1] + ei[0][1] * uz[m + 1] + ei[0][2] * tz[m + 1] + ei[0][3] * tr[m + 1];
      cplx kmsu = ei[1][0] * ur[m + 1] + ei[1][1] * uz[m + 1] + ei[1][2] * tz[m + 1] + ei[1][3] * tr[m + 1];
      cplx FA = ffunc(ra, x->zd[m]);
      cplx GA = gfunc(ra, x->zd[m]);
      cplx FB = ffunc(rb, x->zd[m]);
      cplx GB = gfunc(rb, x->zd[m]);
      cplx H1 = h1func(ra, rb, x->zd[m]);
      cplx H2 = h2func(ra, rb, x->zd[m]);
      c = e[i][0] * e[j][0] * kmpu * kmpu * FA +
          e[i][2] * e[j][2] * km1pd * km1pd * FA +
          e[i][1] * e[j][1] * kmsu * kmsu * FB +
          e[i][3] * e[j][3] * km1sd * km1sd * FB +
          H1 * ((e[i][0] * e[j][1] + e[i][1] * e[j][0]) * kmpu * kmsu +
                (e[i][2] * e[j][3] + e[i][3] * e[j][2]) * km1pd * km1sd) +
          H2 * ((e[i][0] * e[j][3] + e[i][3] * e[j][0]) * kmpu * km1sd +
                (e[i][1] * e[j][2] + e[i][2] * e[j][1]) * km1pd * kmsu) +
          GA * (e[i][0] * e[j][2] + e[i][2] * e[j][0]) * kmpu * km1pd +
          GB * (e[i][1] * e[j][3] + e[i][3] * e[j][1]) * kmsu * km1sd;
    } else {
      cplx km1pd = ei[2][0] * ur[m] + ei[2][1] * uz[m] + ei[2][2] * tz[m] + ei[2][3] * tr[m];
      cplx km1sd = ei[3][0] * ur[m] + ei[3][1] * uz[m] + ei[3][2] * tz[m] + ei[3][3] * tr[m];
      c = e[i][2] * e[j][2] * km1pd * km1pd / (2.0 * ra) +
          (e[i][2] * e[j][3] + e[i][3] * e[j][2]) * km1pd * km1sd / (ra + rb) +
          e[i][3] * e[j][3] * km1sd * km1sd / (2.0 * rb);
    }
  }
  return creal(c);
}

/* sregn96.f90:1430-1529 getdcdh */
static void getdcdh(rctx *x, double om2, double wvno, double wvno2,
                    double fac) {
  double *zrho = x->zrho, *xmu = x->xmu, *xlam = x->xlam;
  for (int m = 0; m < x->mmax; ++m) {
    double tuz = x->uz[m], ttz = x->tz[m], ttr = x->tr[m], tur;
    if (x->iwat[m] == 1)
      tur = -wvno * ttz / (zrho[m] * om2);
    else
      tur = x->ur[m];
    double gfac1, gfac2, gfac3, gfac4, gfac5, gfac6;
    if (m == 0) {
      double drho = zrho[0] - 0.0;
      double dmu = xmu[0] - 0.0;
      double dlm = xlam[0] - 0.0;
      double dl2mu = dlm + dmu + dmu;
      double xl2mp = xlam[m] + xmu[m] + xmu[m];
      double duzdzp = (ttz + wvno * xlam[m] * tur) / xl2mp;
      double durdzp;
      if (x->iwat[m] == 1)
        durdzp = wvno * tuz;
      else
        durdzp = (ttr / xmu[m]) - wvno * tuz;
      double drur2 = tur * tur * drho;
      double dlur2 = tur * tur * dl2mu;
      gfac1 = om2 * drho * (tuz * tuz);
      gfac2 = om2 * drur2;
      gfac3 = -wvno2 * dmu * (tuz * tuz);
      gfac4 = -wvno2 * dlur2;
      gfac5 = (xl2mp * (duzdzp * duzdzp));
      gfac6 = (xmu[m] * (durdzp * durdzp));
    } else {
      double drho = zrho[m] - zrho[m - 1];
      double dmu = xmu[m] - xmu[m - 1];
      double dlm = xlam[m] - xlam[m - 1];
      double dl2mu = dlm + dmu + dmu;
      double xl2mp = xlam[m] + xmu[m] + xmu[m];
      double xl2mm = xlam[m - 1] + xmu[m - 1] + xmu[m - 1];
      double duzdzp = (ttz + wvno * xlam[m] * tur) / xl2mp;
      double durdzp, durdzm, drur2, dlur2, duzdzm;
      if (xmu[m] == 0.0)
        durdzp = wvno * tuz;
      else
        durdzp = (ttr / xmu[m]) - wvno * tuz;
      if (xmu[m - 1] == 0.0)
        durdzm = wvno * tuz;
      else
        durdzm = (ttr / xmu[m - 1]) - wvno * tuz;
      if (x->iwat[m - 1] == 1 && x->iwat[m] == 0) {
        double URB = -wvno * x->tz[m] / (zrho[m - 1] * om2);
        drur2 = tur * tur * zrho[m] - URB * URB * zrho[m - 1];
        dlur2 = tur * tur * xl2mp - URB * URB * xl2mm;
        duzdzm = (ttz + wvno * xlam[m - 1] * URB) / xlam[m - 1];
      } else if (x->iwat[m - 1] == 1 && x->iwat[m] == 1) {
        double URB = -wvno * x->tz[m] / (zrho[m - 1] * om2);
        drur2 = tur * tur * zrho[m] - URB * URB * zrho[m - 1];
        dlur2 = tur * tur * xl2mp - URB * URB * xl2mm;
        duzdzm = (ttz + wvno * xlam[m - 1] * URB) / xl2mm;
      } else {
        drur2 = tur * tur * drho;
        dlur2 = tur * tur * dl2mu;
        duzdzm = (ttz + wvno * xlam[m - 1] * tur) / xl2mm;
      }
      gfac1 = om2 * drho * (tuz * tuz);
      gfac2 = om2 * drur2;
      gfac3 = -wvno2 * dmu * (tuz * tuz);
      gfac4 = -wvno2 * dlur2;
      gfac5 = (xl2mp * (duzdzp * duzdzp) - xl2mm * (duzdzm * duzdzm));
      gfac6 = (xmu[m] * (durdzp * durdzp) - xmu[m - 1] * (durdzm * durdzm));
    }
    double dfac = fac * (gfac1 + gfac2 + gfac3 + gfac4 + gfac5 + gfac6);
    if (fabs(dfac) < 1.0e-38) dfac = 0.0;
    x->dcdh[m] = dfac;
  }
}

/* sregn96.f90:1531-1583 getmat + 1061-1196 energy */
static void energy_r(rctx *x, double om, double wvno) {
  int mmax = x->mmax;
  x->sumi0 = x->sumi1 = x->sumi2 = x->sumi3 = 0.0;
  double c = om / wvno, om2 = om * om, wvno2 = wvno * wvno;
  for (int m = 0; m < mmax; ++m) {
    double ah, av, bh, bv, eta, rho, TA, TC, TF, TL, TN;
    double a12, a14 = 0, a21 = 0, a23 = 0;
    if (x->iwat[m] == 1) {
      ah = x->za[m]; av = x->za[m]; bh = 0.0; bv = 0.0; rho = x->zrho[m];
      eta = 1.0; TL = 0.0; TN = 0.0;
      TC = x->zrho[m] * x->za[m] * x->za[m];
      TA = x->zrho[m] * x->za[m] * x->za[m];
      TF = TA - 2. * TN;
      a12 = -(wvno * wvno - om * om / (ah * ah)) / (rho * om * om);
    } else {
      ah = x->za[m]; av = x->za[m]; bh = x->zb[m]; bv = x->zb[m];
      rho = x->zrho[m]; eta = 1.0;
      TL = x->zrho[m] * x->zb[m] * x->zb[m];
      TN = x->zrho[m] * x->zb[m] * x->zb[m];
      TC = x->zrho[m] * x->za[m] * x->za[m];
      TA = x->zrho[m] * x->za[m] * x->za[m];
      TF = TA - 2. * TN;
      a12 = -wvno;
      a14 = 1.0 / TL;
      a21 = wvno * TF / TC;
      a23 = 1.0 / TC;
    }
    (void)bh;
    int typelyr = (m == mmax - 1) ? 1 : 0;
    double INT11 = intijr(x, 1, 1, m, typelyr, om, om2, wvno, wvno2);
    double INT13 = intijr(x, 1, 3, m, typelyr, om, om2, wvno, wvno2);
    double INT22 = intijr(x, 2, 2, m, typelyr, om, om2, wvno, wvno2);
    double INT24 = intijr(x, 2, 4, m, typelyr, om, om2, wvno, wvno2);
    double INT33 = intijr(x, 3, 3, m, typelyr, om, om2, wvno, wvno2);
    double INT44 = intijr(x, 4, 4, m, typelyr, om, om2, wvno, wvno2);
    (void)INT33; (void)INT44; (void)INT13; (void)INT24;
    if (x->iwat[m] == 1) {
      double wr = wvno / (rho * om2);
      double URUR = INT22 * (wr * wr);
      double UZUZ = INT11;
      double URDUZ = -(wvno / (rho * om2)) * a12 * INT22;
      double DUZDUZ = a12 * a12 * INT22;
      x->sumi0 = x->sumi0 + rho * (URUR + UZUZ);
      x->sumi1 = x->sumi1 + TA * URUR;
      x->sumi2 = x->sumi2 - TF * URDUZ;
      x->sumi3 = x->sumi3 + TC * DUZDUZ;
      double facah = rho * ah * (URUR - 2. * eta * URDUZ / wvno);
      double facav = rho * av * DUZDUZ / wvno2;
      x->dcda[m] = facah + facav;
      double facr = -0.5 * c * c * (URUR + UZUZ);
      x->dcdr[m] = 0.5 * (av * facav + ah * facah) / rho + facr;
      x->ur[m] = -wvno * x->tz[m] / (rho * om2);
    } else {
      double URUR = INT11;
      double UZUZ = INT22;
      double DURDUR = a12 * a12 * INT22 + 2. * a12 * a14 * INT24 + a14 * a14 * INT44;
      double DUZDUZ = a21 * a21 * INT11 + 2. * a21 * a23 * INT13 + a23 * a23 * INT33;
      double URDUZ = a21 * INT11 + a23 * INT13;
      double UZDUR = a12 * INT22 + a14 * INT24;
      x->sumi0 = x->sumi0 + rho * (URUR + UZUZ);
      x->sumi1 = x->sumi1 + TL * UZUZ + TA * URUR;
      x->sumi2 = x->sumi2 + TL * UZDUR - TF * URDUZ;
      x->sumi3 = x->sumi3 + TL * DURDUR + TC * DUZDUZ;
      double facah = rho * ah * (URUR - 2. * eta * URDUZ / wvno);
      double facav = rho * av * DUZDUZ / wvno2;
      double facbh = 0.0;
      double facbv = rho * bv * (UZUZ + 2. * UZDUR / wvno + DURDUR / wvno2 +
                                 4. * eta * URDUZ / wvno);
      x->dcda[m] = facah + facav;
      x->dcdb[m] = facbv + facbh;
      double facr = -0.5 * c * c * (URUR + UZUZ);
      x->dcdr[m] = 0.5 * (av * facav + ah * facah + bv * facbv) / rho + facr;
    }
  }
  x->flagr = om2 * x->sumi0 - wvno2 * x->sumi1 - 2.0 * wvno * x->sumi2 - x->sumi3;
  x->ugr = (wvno * x->sumi1 + x->sumi2) / (om * x->sumi0);
  x->are = wvno / (2.0 * om * x->ugr * x->sumi0);
  double fac = x->are * c / wvno2;
  for (int m = 0; m < mmax; ++m) {
    x->dcda[m] = x->dcda[m] / (x->ugr * x->sumi0);
    x->dcdb[m] = x->dcdb[m] / (x->ugr * x->sumi0);
    x->dcdr[m] = x->dcdr[m] / (x->ugr * x->sumi0);
  }
  getdcdh(x, om2, wvno, wvno2, fac);
}

static void rsetup(rctx *x, const float *thk, const float *vp, const float *vs,
                   const float *rhom, int nlayer, int iflsph) {
  ralloc(x, nlayer);
  for (int i = 0; i < nlayer; ++i) {
    x->zb[i] = vs[i];
    x->za[i] = vp[i];
    x->zrho[i] = rhom[i];
    x->zd[i] = thk[i];
  }
  x->allfluid = 1;
  for (int i = 0; i < nlayer; ++i) {
    if (x->zb[i] > 0.0) {
      x->allfluid = 0;
      x->iwat[i] = 0;
    } else {
      x->iwat[i] = 1;
    }
  }
  if (iflsph > 0) bldsph_r(x);
  for (int i = 0; i < nlayer; ++i) {
    x->xmu[i] = x->zrho[i] * (x->zb[i] * x->zb[i]);
    x->xlam[i] = x->zrho[i] * (x->za[i] * x->za[i]) - 2 * x->xmu[i];
  }
}

static double sumrange(const double *a, int i0, int i1) {
  double s = 0.0;
  for (int i = i0; i < i1; ++i) s += a[i];
  return s;
}

/* sregn96.f90:1631-1735 */
void or_sregn96(const float *thk, const float *vp, const float *vs,
                const float *rhom, int nlayer, const double *t, double *cp,
                double *cg, double *dispu, double *dispw, double *stressu,
                double *stressw, double *dc2da, double *dc2db, double *dc2dh,
                double *dc2dr, int iflsph) {
  rctx X, *x = &X;
  rsetup(x, thk, vp, vs, rhom, nlayer, iflsph);
  const double pi = (double)3.1415926535898f;
  double twopi = 2.0 * pi;
  double omega = twopi / *t;
  double c = *cp;
  x->ugr = c;
  double wvno = omega / c;
  svfunc(x, omega, wvno);
  energy_r(x, omega, wvno);
  if (fabs(c) < 1.0e-36) c = 0.0;
  if (fabs(x->ugr) < 1.0e-36) x->ugr = 0.0;
  double csph, usph;
  int mmax = x->mmax;
  if (iflsph > 0) {
    const double ar = 6371.0;
    double r = c / (2. * ar * omega);
    double tm = sqrt(1. + r * r);
    double tm3 = tm * tm * tm;
    for (int i = 0; i < mmax; ++i) {
      x->dcda[i] = x->dcda[i] * x->vtp[i] / tm3;
      x->dcdb[i] = x->dcdb[i] * x->vtp[i] / tm3;
      x->dcdh[i] = x->dcdh[i] * x->dtp[i] / tm3;
      x->dcdr[i] = x->dcdr[i] * x->rtp[i] / tm3;
    }
    usph = x->ugr * tm;
    csph = c / tm;
  } else {
    csph = c;
    usph = x->ugr;
  }
  for (int i = 0; i < mmax - 1; ++i) x->dcdh[i] = sumrange(x->dcdh, i + 1, mmax);
  x->dcdh[mmax - 1] = 0.0;
  for (int i = 0; i < mmax; ++i) {
    dc2da[i] = x->dcda[i]; dc2db[i] = x->dcdb[i];
    dc2dr[i] = x->dcdr[i]; dc2dh[i] = x->dcdh[i];
    dispu[i] = x->ur[i]; dispw[i] = x->uz[i];
    stressu[i] = x->tr[i]; stressw[i] = x->tz[i];
  }
  *cp = csph;
  *cg = usph;
  rfree(x);
}

/* sregn96.f90:1737-1878 */
void or_sregnpu(const float *thk, const float *vp, const float *vs,
                const float *rhom, int nlayer, const double *t, double *cp,
                double *cg, double *dispu, double *dispw, double *stressu,
                double *stressw, double *t1, double *cp1, double *t2,
                double *cp2, double *dc2da, double *dc2db, double *dc2dh,
                double *dc2dr, double *du2da, double *du2db, double *du2dh,
                double *du2dr, int iflsph) {
  rctx X, *x = &X;
  rsetup(x, thk, vp, vs, rhom, nlayer, iflsph);
  int mmax = nlayer;
  double *k1 = (double *)malloc(sizeof(double) * 8 * (size_t)mmax);
  double *dc2da1 = k1, *dc2db1 = k1 + mmax, *dc2dh1 = k1 + 2 * mmax,
         *dc2dr1 = k1 + 3 * mmax, *dc2da2 = k1 + 4 * mmax,
         *dc2db2 = k1 + 5 * mmax, *dc2dh2 = k1 + 6 * mmax,
         *dc2dr2 = k1 + 7 * mmax;
  const float pif = atanf(1.0f) * 4.0f;
  const double pi = (double)pif;
  double twopi = 2.0 * pi;
  double omega = twopi / *t;
  double c = *cp;
  double wvno = omega / c;
  svfunc(x, omega, wvno);
  energy_r(x, omega, wvno);
  *cg = x->ugr;
  for (int i = 0; i < mmax; ++i) {
    dc2da[i] = x->dcda[i]; dc2db[i] = x->dcdb[i];
    dc2dr[i] = x->dcdr[i]; dc2dh[i] = x->dcdh[i];
    dispu[i] = x->ur[i]; dispw[i] = x->uz[i];
    stressu[i] = x->tr[i]; stressw[i] = x->tz[i];
  }
  omega = twopi / *t1;
  c = *cp1;
  wvno = omega / c;
  svfunc(x, omega, wvno);
  energy_r(x, omega, wvno);
  for (int i = 0; i < mmax; ++i) {
    dc2da1[i] = x->dcda[i]; dc2db1[i] = x->dcdb[i];
    dc2dr1[i] = x->dcdr[i]; dc2dh1[i] = x->dcdh[i];
  }
  omega = twopi / *t2;
  c = *cp2;
  wvno = omega / c;
  svfunc(x, omega, wvno);
  energy_r(x, omega, wvno);
  for (int i = 0; i < mmax; ++i) {
    dc2da2[i] = x->dcda[i]; dc2db2[i] = x->dcdb[i];
    dc2dr2[i] = x->dcdr[i]; dc2dh2[i] = x->dcdh[i];
  }
  /* sregn96.f90:1831-1834: the module arrays still hold the T2 kernels */
  double uc1 = *cg / *cp;
  double tt = *t, dt = *t2 - *t1;
  for (int i = 0; i < mmax; ++i) {
    du2da[i] = uc1 * (2.0 - uc1) * x->dcda[i] - uc1 * uc1 * tt * (dc2da2[i] - dc2da1[i]) / dt;
    du2db[i] = uc1 * (2.0 - uc1) * x->dcdb[i] - uc1 * uc1 * tt * (dc2db2[i] - dc2db1[i]) / dt;
    du2dr[i] = uc1 * (2.0 - uc1) * x->dcdr[i] - uc1 * uc1 * tt * (dc2dr2[i] - dc2dr1[i]) / dt;
    du2dh[i] = uc1 * (2.0 - uc1) * x->dcdh[i] - uc1 * uc1 * tt * (dc2dh2[i] - dc2dh1[i]) / dt;
  }
  if (iflsph > 0) {
    const double ar = 6371.0;
    omega = twopi / *t;
    double r = *cp / (2. * ar * omega);
    double tm = sqrt(1. + r * r);
    double q = 0.5 / (ar * omega);
    double tm1 = q * q / tm;
    double tm3 = tm * tm * tm;
    for (int i = 0; i < mmax; ++i) {
      du2da[i] = (tm * du2da[i] + *cg * *cp * dc2da[i] * tm1) * x->vtp[i];
      du2db[i] = (tm * du2db[i] + *cg * *cp * dc2db[i] * tm1) * x->vtp[i];
      du2dr[i] = (tm * du2dr[i] + *cg * *cp * dc2dr[i] * tm1) * x->rtp[i];
      du2dh[i] = (tm * du2dh[i] + *cg * *cp * dc2dh[i] * tm1) * x->dtp[i];
      dc2da[i] = dc2da[i] / tm3 * x->vtp[i];
      dc2db[i] = dc2db[i] / tm3 * x->vtp[i];
      dc2dr[i] = dc2dr[i] / tm3 * x->rtp[i];
      dc2dh[i] = dc2dh[i] / tm3 * x->dtp[i];
    }
    *cp = *cp / tm;
    *cg = *cg * tm;
  }
  for (int i = 0; i < mmax - 1; ++i) {
    dc2dh[i] = sumrange(dc2dh, i + 1, mmax);
    du2dh[i] = sumrange(du2dh, i + 1, mmax);
  }
  dc2dh[mmax - 1] = 0.0;
  du2dh[mmax - 1] = 0.0;
  free(k1);
  rfree(x);
}
