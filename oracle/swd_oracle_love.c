/*
 * swd_oracle_love.c -- CPU restatement of src/SWD/slegn96.f90 (Love
 * eigenfunctions, energy integrals, group velocity and layer kernels).
 * TEST INFRASTRUCTURE ONLY (the checker); nothing in the product links it.
 *
 * Quirks kept: pi is the REAL literal (float), rtp = tmp**(-5) is lowered
 * as 1/(t2*t2*t), the f2 term repeats el(1,1)*el(1,2) (slegn96.f90:548),
 * `stress` receives the displacement in slegn96_ (slegn96.f90:770).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "swd_oracle.h"

typedef double complex cplx;

typedef struct {
  int mmax;
  double *zd, *zb, *zrho, *xmu, *uu, *tt, *dcdb, *dcdh, *dcdr, *vtp, *dtp,
      *rtp, *exl;
  int *iwat;
  double cosq, sinq, yl, zl, mu;
  double sumi0, sumi1, sumi2, flagr, ale, ugr;
} lctx;

static void lalloc(lctx *x, int n) {
  memset(x, 0, sizeof(*x));
  x->mmax = n;
  double *p = (double *)calloc((size_t)n * 13, sizeof(double));
  x->zd = p; x->zb = p + n; x->zrho = p + 2 * n; x->xmu = p + 3 * n;
  x->uu = p + 4 * n; x->tt = p + 5 * n; x->dcdb = p + 6 * n;
  x->dcdh = p + 7 * n; x->dcdr = p + 8 * n; x->vtp = p + 9 * n;
  x->dtp = p + 10 * n; x->rtp = p + 11 * n; x->exl = p + 12 * n;
  x->iwat = (int *)calloc((size_t)n, sizeof(int));
}
static void lfree(lctx *x) {
  free(x->zd);
  free(x->iwat);
}

static double powm5d(double x) {
  double x2 = x * x, x4 = x2 * x2, x5 = x4 * x;
  return 1.0 / x5;
}

/* slegn96.f90:107-151 bldsph (Love), R = 6371 km */
static void bldsph_l(lctx *x) {
  const double ar = 6371.0;
  double dr = 0.0, r0 = ar;
  int mmax = x->mmax;
  x->zd[mmax - 1] = 1.0;
  for (int i = 0; i < mmax; ++i) {
    dr = dr + x->zd[i];
    double r1 = ar - dr;
    double z0 = ar * log(ar / r0);
    double z1 = ar * log(ar / r1);
    x->dtp[i] = ar / r0;
    double tmp = (2.0 * ar) / (r0 + r1);
    x->vtp[i] = tmp;
    x->rtp[i] = powm5d(tmp);
    x->zb[i] = x->zb[i] * tmp;
    x->zrho[i] = x->zrho[i] * x->rtp[i];
    x->zd[i] = z1 - z0;
    r0 = r1;
  }
  x->zd[mmax - 1] = 0.0;
}

/* slegn96.f90:248-328 varl; returns rb, xkb and the extracted exponent */
static void varl(lctx *x, int m, double *rb, double omega, double wvno,
                 double *xkb, double dpth, double *eexl) {
  *xkb = omega / x->zb[m];
  double wvnop = wvno + *xkb;
  double wvnom = fabs(wvno - *xkb);
  *rb = sqrt(wvnop * wvnom);
  double q = *rb * dpth;
  x->mu = x->zrho[m] * x->zb[m] * x->zb[m];
  *eexl = 0.0;
  if (wvno < *xkb) {
    x->sinq = sin(q);
    x->yl = x->sinq / *rb;
    x->zl = -*rb * x->sinq;
    x->cosq = cos(q);
  } else if (wvno == *xkb) {
    x->cosq = 1.0;
    x->yl = dpth;
    x->zl = 0.0;
  } else if (wvno > *xkb) {
    *eexl = q;
    double fac = 0.0;
    if (q < 18.0) fac = exp(-2.0 * q);
    x->cosq = (1.0 + fac) * 0.5;
    x->sinq = (1.0 - fac) * 0.5;
    x->yl = x->sinq / *rb;
    x->zl = *rb * x->sinq;
  }
}

/* slegn96.f90:372-445 up */
static double up_l(lctx *x, double omega, double wvno) {
  int mmax = x->mmax;
  double rb, xkb, eexl;
  if (x->zb[mmax - 1] > 0.01) {
    varl(x, mmax - 1, &rb, omega, wvno, &xkb, 0.0, &eexl);
    x->uu[mmax - 1] = 1.0;
    x->tt[mmax - 1] = -x->xmu[mmax - 1] * rb;
  } else {
    x->uu[mmax - 1] = 1.0;
    x->tt[mmax - 1] = 0.0;
  }
  x->exl[mmax - 1] = 0.0;
  double ttlast = 0.0;
  for (int k = mmax - 2; k >= 0; --k) {
    if (x->iwat[k] != 0) continue;
    varl(x, k, &rb, omega, wvno, &xkb, x->zd[k], &eexl);
    double a11 = x->cosq, a22 = x->cosq;
    double a12 = -(x->yl / x->mu);
    double a21 = -(x->zl * x->mu);
    int k1 = k + 1;
    double amp0 = a11 * x->uu[k1] + a12 * (x->tt[k1]);
    double str0 = a21 * x->uu[k1] + a22 * (x->tt[k1]);
    double rr = fabs(amp0), ss = fabs(str0);
    if (ss > rr) rr = ss;
    if (rr < 1.0e-30) rr = 1.0;
    x->exl[k] = log(rr) + eexl;
    x->uu[k] = amp0 / rr;
    x->tt[k] = str0 / rr;
    ttlast = x->tt[k];
  }
  return ttlast;
}

/* slegn96.f90:179-246 shfunc */
static void shfunc(lctx *x, double omega, double wvno) {
  int mmax = x->mmax;
  up_l(x, omega, wvno);
  double ext = 0.0;
  double umax = x->uu[0];
  x->tt[0] = 0.0;
  for (int k = 1; k < mmax; ++k) {
    if (x->iwat[k] == 0) {
      ext = ext + x->exl[k - 1];
      double fact = 0.0;
      if (ext < 80.0) fact = 1. / exp(ext);
      x->uu[k] = x->uu[k] * fact;
      x->tt[k] = x->tt[k] * fact;
    } else {
      x->uu[k] = 0.0;
      x->tt[k] = 0.0;
    }
    if (fabs(x->uu[k]) > fabs(umax)) umax = x->uu[k];
  }
  if (x->uu[0] != 0.0) umax = x->uu[0];
  if (fabs(umax) > 0.0) {
    for (int k = 0; k < mmax; ++k) {
      if (x->iwat[k] == 0) {
        x->uu[k] = x->uu[k] / umax;
        x->tt[k] = x->tt[k] / umax;
      }
    }
  }
}

/* slegn96.f90:447-629 energy */
static void energy_l(lctx *x, double omega, double wvno) {
  int mmax = x->mmax;
  double c = omega / wvno;
  double omega2 = omega * omega, wvno2 = wvno * wvno;
  x->sumi0 = x->sumi1 = x->sumi2 = 0.0;
  for (int k = 0; k < mmax; ++k) {
    if (x->iwat[k] != 0) {
      x->dcdb[k] = 0.0;
      x->dcdr[k] = 0.0;
      continue;
    }
    double TN = x->zrho[k] * x->zb[k] * x->zb[k];
    double TL = x->zrho[k] * x->zb[k] * x->zb[k];
    double VSHH = x->zb[k], VSHV = x->zb[k];
    int k1 = k + 1;
    double drho = x->zrho[k];
    double dpth = x->zd[k];
    double rb, xkb, eexl;
    varl(x, k, &rb, omega, wvno, &xkb, dpth, &eexl);
    double dmu = x->xmu[k];
    if (rb < 1.0e-10) rb = 1.0e-10;
    double upup, dupdup;
    if (k == mmax - 1) {
      upup = (0.5 / rb) * x->uu[mmax - 1] * x->uu[mmax - 1];
      dupdup = (0.5 * rb) * x->uu[mmax - 1] * x->uu[mmax - 1];
    } else {
      cplx nub = CMPLX(rb, 0.0);
      if (wvno < xkb) nub = CMPLX(0.0, rb);
      cplx xnub = dmu * nub;
      /* slegn96.f90:353-370 emat */
      cplx ei11 = 0.5 / wvno, ei12 = 0.5 / (wvno * xnub);
      cplx ei21 = 0.5 / wvno, ei22 = -0.5 / (wvno * xnub);
      cplx el11 = wvno, el12 = wvno;
      cplx km1dn = ei21 * x->uu[k] + ei22 * x->tt[k];
      cplx kmup = ei11 * x->uu[k1] + ei12 * x->tt[k1];
      cplx f3 = nub * dpth;
      cplx exqq = 0.0;
      if (creal(f3) < 40.0) exqq = cexp(-2.0 * f3);
      cplx f = (1.0 - exqq) / (2.0 * nub);
      exqq = 0.0;
      if (creal(f3) < 75.0) exqq = cexp(-f3);
      cplx g = dpth * exqq;
      cplx f1 = f * (el11 * el11 * kmup * kmup + el12 * el12 * km1dn * km1dn);
      cplx f2 = g * (el11 * el12 + el11 * el12) * kmup * km1dn;
      upup = creal(f1 + f2);
      dupdup = creal(nub * nub * (f1 - f2));
    }
    x->sumi0 = x->sumi0 + drho * upup;
    x->sumi1 = x->sumi1 + TN * upup;
    x->sumi2 = x->sumi2 + TL * dupdup;
    double DCDBH = c * drho * VSHH * upup;
    double DCDBV = c * drho * VSHV * dupdup / wvno2;
    x->dcdb[k] = DCDBH + DCDBV;
    x->dcdr[k] = 0.5 * c * (-c * c * upup + VSHH * VSHH * upup + VSHV * VSHV * dupdup / wvno2);
  }
  for (int k = 0; k < mmax; ++k) {
    if (x->iwat[k] == 0) {
      x->dcdb[k] = x->dcdb[k] / x->sumi1;
      x->dcdr[k] = x->dcdr[k] / x->sumi1;
    } else {
      x->dcdb[k] = 0.0;
      x->dcdr[k] = 0.0;
    }
  }
  x->flagr = omega2 * x->sumi0 - wvno2 * x->sumi1 - x->sumi2;
  x->ugr = x->sumi1 / (c * x->sumi0);
  x->ale = 0.5 / x->sumi1;
  double fac = x->ale * c / wvno2;
  int llflag = 0;
  for (int k = 0; k < mmax; ++k) {
    if (x->iwat[k] == 0) {
      double drho, dmu, dvdz;
      if (llflag == 0) {
        drho = x->zrho[k];
        dmu = x->xmu[k];
        dvdz = 0.0;
      } else {
        drho = x->zrho[k] - x->zrho[k - 1];
        dmu = x->xmu[k] - x->xmu[k - 1];
        dvdz = x->tt[k] * x->tt[k] * (1.0 / x->xmu[k] - 1.0 / x->xmu[k - 1]);
      }
      double dfac = fac * (x->uu[k] * x->uu[k] * (omega2 * drho - wvno2 * dmu) + dvdz);
      x->dcdh[k] = (fabs(dfac) < 1.0e-38) ? 0.0 : dfac;
      llflag = llflag + 1;
    } else {
      x->dcdh[k] = 0.0;
    }
  }
}

static void lsetup(lctx *x, const float *thk, const float *vs,
                   const float *rhom, int nlayer, int iflsph, int pu) {
  lalloc(x, nlayer);
  for (int i = 0; i < nlayer; ++i) {
    x->zb[i] = vs[i];
    x->zrho[i] = rhom[i];
    x->zd[i] = thk[i];
  }
  for (int i = 0; i < nlayer; ++i) {
    if (pu)
      x->iwat[i] = (x->zb[i] > 0.0) ? 0 : 1; /* slegn96.f90:831-837 */
    else
      x->iwat[i] = (x->zb[i] == 0.0) ? 1 : 0; /* slegn96.f90:712-718 */
  }
  if (iflsph > 0) bldsph_l(x);
  for (int i = 0; i < nlayer; ++i) x->xmu[i] = x->zrho[i] * (x->zb[i] * x->zb[i]);
}

static double sumrange(const double *a, int i0, int i1) {
  double s = 0.0;
  for (int i = i0; i < i1; ++i) s += a[i];
  return s;
}

/* slegn96.f90:672-783 */
void or_slegn96(const float *thk, const float *vs, const float *rhom,
                int nlayer, const double *t, double *cp, double *cg,
                double *disp, double *stress, double *dc2db, double *dc2dh,
                double *dc2dr, int iflsph) {
  lctx X, *x = &X;
  lsetup(x, thk, vs, rhom, nlayer, iflsph, 0);
  int mmax = nlayer;
  const double pi = (double)3.1415926535898f;
  double twopi = 2.0 * pi;
  double omega = twopi / *t;
  double c = *cp;
  double wvno = omega / c;
  shfunc(x, omega, wvno);
  energy_l(x, omega, wvno);
  double csph, usph;
  if (iflsph > 0) {
    const double a = 6371.0;
    double r = 3.0 * c / (2. * a * omega);
    double tm = sqrt(1. + r * r);
    double tm3 = tm * tm * tm;
    for (int i = 0; i < mmax; ++i) {
      x->dcdb[i] = x->dcdb[i] * x->vtp[i] / tm3;
      x->dcdh[i] = x->dcdh[i] * x->dtp[i] / tm3;
      x->dcdr[i] = x->dcdr[i] * x->rtp[i] / tm3;
    }
    csph = c / tm;
    usph = x->ugr * tm;
  } else {
    csph = c;
    usph = x->ugr;
  }
  /* the underflow guards at slegn96.f90:749-760 run after cg is formed and
   * change no output */
  for (int i = 0; i < mmax - 1; ++i) dc2dh[i] = sumrange(x->dcdh, i + 1, mmax);
  dc2dh[mmax - 1] = 0.0;
  for (int i = 0; i < mmax; ++i) {
    disp[i] = x->uu[i];
    stress[i] = x->uu[i];
    dc2db[i] = x->dcdb[i];
    dc2dr[i] = x->dcdr[i];
  }
  *cp = csph;
  *cg = usph;
  lfree(x);
}

/* slegn96.f90:785-918 */
void or_slegnpu(const float *thk, const float *vs, const float *rhom,
                int nlayer, const double *t, double *cp, double *cg,
                double *disp, double *stress, double *t1, double *cp1,
                double *t2, double *cp2, double *dc2db, double *dc2dh,
                double *dc2dr, double *du2db, double *du2dh, double *du2dr,
                int iflsph) {
  lctx X, *x = &X;
  lsetup(x, thk, vs, rhom, nlayer, iflsph, 1);
  int mmax = nlayer;
  double *k1 = (double *)malloc(sizeof(double) * 6 * (size_t)mmax);
  double *dc2db1 = k1, *dc2dh1 = k1 + mmax, *dc2dr1 = k1 + 2 * mmax,
         *dc2db2 = k1 + 3 * mmax, *dc2dh2 = k1 + 4 * mmax,
         *dc2dr2 = k1 + 5 * mmax;
  const float pif = atanf(1.0f) * 4.0f;
  double twopi = 2.0 * (double)pif;
  double omega = twopi / *t;
  double c = *cp;
  double wvno = omega / c;
  shfunc(x, omega, wvno);
  energy_l(x, omega, wvno);
  *cg = x->ugr;
  for (int i = 0; i < mmax; ++i) {
    dc2db[i] = x->dcdb[i]; stress[i] = x->tt[i];
    dc2dr[i] = x->dcdr[i]; disp[i] = x->uu[i];
    dc2dh[i] = x->dcdh[i];
  }
  omega = twopi / *t1;
  c = *cp1;
  wvno = omega / c;
  shfunc(x, omega, wvno);
  energy_l(x, omega, wvno);
  for (int i = 0; i < mmax; ++i) {
    dc2db1[i] = x->dcdb[i]; dc2dr1[i] = x->dcdr[i]; dc2dh1[i] = x->dcdh[i];
  }
  omega = twopi / *t2;
  c = *cp2;
  wvno = omega / c;
  shfunc(x, omega, wvno);
  energy_l(x, omega, wvno);
  for (int i = 0; i < mmax; ++i) {
    dc2db2[i] = x->dcdb[i]; dc2dr2[i] = x->dcdr[i]; dc2dh2[i] = x->dcdh[i];
  }
  double uc1 = *cg / *cp;
  double tt = *t, dt = *t2 - *t1;
  for (int i = 0; i < mmax; ++i) {
    du2db[i] = uc1 * (2.0 - uc1) * x->dcdb[i] - uc1 * uc1 * tt * (dc2db2[i] - dc2db1[i]) / dt;
    du2dr[i] = uc1 * (2.0 - uc1) * x->dcdr[i] - uc1 * uc1 * tt * (dc2dr2[i] - dc2dr1[i]) / dt;
    du2dh[i] = uc1 * (2.0 - uc1) * x->dcdh[i] - uc1 * uc1 * tt * (dc2dh2[i] - dc2dh1[i]) / dt;
  }
  if (iflsph > 0) {
    const double ar = 6371.0;
    omega = twopi / *t;
    double r = 3.0 * *cp / (2. * ar * omega);
    double tm = sqrt(1. + r * r);
    double q = 1.5 / (ar * omega);
    double tm1 = q * q / tm;
    double tm3 = tm * tm * tm;
    for (int i = 0; i < mmax; ++i) {
      du2db[i] = (tm * du2db[i] + *cg * *cp * dc2db[i] * tm1) * x->vtp[i];
      du2dr[i] = (tm * du2dr[i] + *cg * *cp * dc2dr[i] * tm1) * x->rtp[i];
      du2dh[i] = (tm * du2dh[i] + *cg * *cp * dc2dh[i] * tm1) * x->dtp[i];
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
  lfree(x);
}
