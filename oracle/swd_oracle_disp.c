/*
 * swd_oracle_disp.c -- CPU restatement of the phase-velocity root chain,
 * src/SWD/surfdisp96.f.  TEST INFRASTRUCTURE ONLY (the checker); nothing in
 * the product links it.
 *
 * Precision is kept as in the Fortran: the model arrays and the flattened
 * d, a, b, rho are float; gtsolh runs in float; cc1/betmn/betmx are REAL;
 * the returned velocity is sngl(c); everything else is double.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "swd_oracle.h"

/* Stop rule of nevill: |c1-c2| <= tol*c1 (surfdisp96.f:634 uses 1e-6).
 * tol < 1e-6 selects the exact-root mode: roots are returned unrounded. */
static double g_tol = 1.0e-6;
void or_set_tolerance(double tol) { g_tol = tol; }

typedef struct {
  int mmax, llw;
  float *d, *a, *b, *rho, *rtp, *dtp, *btp;
  double del1st; /* common /getsol_module/ (surfdisp96.f:424-427) */
  float dhalf;   /* common /sphere_module/ (surfdisp96.f:530-533) */
  long long nevals;
} chain_ctx;

static double fsign1(double x) { return copysign(1.0, x); }

/* x**(-5) for REAL x: flang lowers it as 1/(x2*x2*x) (checked by probe). */
static float powm5f(float x) {
  float x2 = x * x, x4 = x2 * x2, x5 = x4 * x;
  return 1.0f / x5;
}

/* surfdisp96.f:377-398 gtsolh: Rayleigh half-space speed, single precision */
static float gtsolh(float a, float b) {
  float c = 0.95f * b;
  for (int i = 0; i < 5; ++i) {
    float gamma = b / a;
    float kappa = c / b;
    float k2 = kappa * kappa;
    float gk = gamma * kappa;
    float gk2 = gk * gk;
    float fac1 = sqrtf(1.0f - gk2);
    float fac2 = sqrtf(1.0f - k2);
    float tk = 2.0f - k2;
    float fr = tk * tk - 4.0f * fac1 * fac2;
    float frp = -(4.0f * (2.0f - k2) * kappa) +
                4.0f * fac2 * gamma * gamma * kappa / fac1 +
                4.0f * fac1 * kappa / fac2;
    frp = frp / b;
    c = c - fr / frp;
  }
  return c;
}

/* surfdisp96.f:499-571 sphere: Earth flattening, R = 6370 km */
static void sphere(chain_ctx *x, int ifunc, int iflag) {
  const double ar = 6370.0;
  double dr = 0.0, r0 = ar;
  int mmax = x->mmax;
  x->d[mmax - 1] = 1.0f;
  if (iflag == 0) {
    for (int i = 0; i < mmax; ++i) {
      x->dtp[i] = x->d[i];
      x->rtp[i] = x->rho[i];
    }
    for (int i = 0; i < mmax; ++i) {
      dr = dr + (double)x->d[i];
      double r1 = ar - dr;
      double z0 = ar * log(ar / r0);
      double z1 = ar * log(ar / r1);
      x->d[i] = (float)(z1 - z0);
      double tmp = (ar + ar) / (r0 + r1);
      x->a[i] = (float)(x->a[i] * tmp);
      x->b[i] = (float)(x->b[i] * tmp);
      x->btp[i] = (float)tmp;
      r0 = r1;
    }
    x->dhalf = x->d[mmax - 1];
  } else {
    x->d[mmax - 1] = x->dhalf;
    for (int i = 0; i < mmax; ++i) {
      if (ifunc == 1)
        x->rho[i] = x->rtp[i] * powm5f(x->btp[i]);
      else if (ifunc == 2)
        x->rho[i] = x->rtp[i] * powf(x->btp[i], -2.275f);
    }
  }
  x->d[mmax - 1] = 0.0f;
}

/* surfdisp96.f:734-794 dltar1: Love period equation (Haskell SH) */
static double dltar1(const chain_ctx *x, double wvno, double omega) {
  int mmax = x->mmax;
  double beta1 = (double)x->b[mmax - 1];
  double rho1 = (double)x->rho[mmax - 1];
  double xkb = omega / beta1;
  double wvnop = wvno + xkb;
  double wvnom = fabs(wvno - xkb);
  double rb = sqrt(wvnop * wvnom);
  double e1 = rho1 * rb;
  double e2 = 1.0 / (beta1 * beta1);
  for (int m = mmax - 2; m >= x->llw - 1; --m) {
    beta1 = (double)x->b[m];
    rho1 = (double)x->rho[m];
    double xmu = rho1 * beta1 * beta1;
    xkb = omega / beta1;
    wvnop = wvno + xkb;
    wvnom = fabs(wvno - xkb);
    rb = sqrt(wvnop * wvnom);
    double q = (double)x->d[m] * rb;
    double sinq, cosq, y, z;
    if (wvno < xkb) {
      sinq = sin(q);
      y = sinq / rb;
      z = -rb * sinq;
      cosq = cos(q);
    } else if (wvno == xkb) {
      cosq = 1.0;
      y = (double)x->d[m];
      z = 0.0;
    } else {
      double fac = 0.0;
      if (q < 16) fac = exp(-2.0 * q);
      cosq = (1.0 + fac) * 0.5;
      sinq = (1.0 - fac) * 0.5;
      y = sinq / rb;
      z = rb * sinq;
    }
    double e10 = e1 * cosq + e2 * xmu * z;
    double e20 = e1 * y / xmu + e2 * cosq;
    double xnor = fabs(e10);
    double ynor = fabs(e20);
    if (ynor > xnor) xnor = ynor;
    if (xnor < 1.0e-40) xnor = 1.0;
    e1 = e10 / xnor;
    e2 = e20 / xnor;
  }
  return e1;
}

/* surfdisp96.f:901-1018 var: cos/sin or normalised cosh/sinh products.
 * Returns w, cosp and the ten products dnka needs. */
typedef struct {
  double a0, cpcq, cpy, cpz, cqw, cqx, xy, xz, wy, wz;
} var_out;

static void var(double p, double q, double ra, double rb, double wvno,
                double xka, double xkb, double dpth, double *w_out,
                double *cosp_out, var_out *o) {
  double pex = 0.0, sex = 0.0;
  double sinp, w = 0, x = 0, cosp = 0, sinq, y = 0, z = 0, cosq = 0;
  if (wvno < xka) {
    sinp = sin(p);
    w = sinp / ra;
    x = -ra * sinp;
    cosp = cos(p);
  } else if (wvno == xka) {
    cosp = 1.0;
    w = dpth;
    x = 0.0;
  } else if (wvno > xka) {
    pex = p;
    double fac = 0.0;
    if (p < 16) fac = exp(-2.0 * p);
    cosp = (1.0 + fac) * 0.5;
    sinp = (1.0 - fac) * 0.5;
    w = sinp / ra;
    x = ra * sinp;
  }
  if (wvno < xkb) {
    sinq = sin(q);
    y = sinq / rb;
    z = -rb * sinq;
    cosq = cos(q);
  } else if (wvno == xkb) {
    cosq = 1.0;
    y = dpth;
    z = 0.0;
  } else if (wvno > xkb) {
    sex = q;
    double fac = 0.0;
    if (q < 16) fac = exp(-2.0 * q);
    cosq = (1.0 + fac) * 0.5;
    sinq = (1.0 - fac) * 0.5;
    y = sinq / rb;
    z = rb * sinq;
  }
  double exa = pex + sex;
  o->a0 = 0.0;
  if (exa < 60.0) o->a0 = exp(-exa);
  o->cpcq = cosp * cosq;
  o->cpy = cosp * y;
  o->cpz = cosp * z;
  o->cqw = cosq * w;
  o->cqx = cosq * x;
  o->xy = x * y;
  o->xz = x * z;
  o->wy = w * y;
  o->wz = w * z;
  /* the exp(sex-pex) rescaling of cosq,y,z at :1010-1016 only touches
   * locals that dltar4 never reads */
  *w_out = w;
  *cosp_out = cosp;
}

/* surfdisp96.f:1051-1094 dnka: Dunkin 5x5 compound matrix, ca[row][col] */
static void dnka(double ca[5][5], double wvno2, double gam, double gammk,
                 double rho, const var_out *v) {
  const double one = 1.0, two = 2.0;
  double gamm1 = gam - one;
  double twgm1 = gam + gamm1;
  double gmgmk = gam * gammk;
  double gmgm1 = gam * gamm1;
  double gm1sq = gamm1 * gamm1;
  double rho2 = rho * rho;
  double a0pq = v->a0 - v->cpcq;
  ca[0][0] = v->cpcq - two * gmgm1 * a0pq - gmgmk * v->xz - wvno2 * gm1sq * v->wy;
  ca[0][1] = (wvno2 * v->cpy - v->cqx) / rho;
  ca[0][2] = -(twgm1 * a0pq + gammk * v->xz + wvno2 * gamm1 * v->wy) / rho;
  ca[0][3] = (v->cpz - wvno2 * v->cqw) / rho;
  ca[0][4] = -(two * wvno2 * a0pq + v->xz + wvno2 * wvno2 * v->wy) / rho2;
  ca[1][0] = (gmgmk * v->cpz - gm1sq * v->cqw) * rho;
  ca[1][1] = v->cpcq;
  ca[1][2] = gammk * v->cpz - gamm1 * v->cqw;
  ca[1][3] = -v->wz;
  ca[1][4] = ca[0][3];
  ca[3][0] = (gm1sq * v->cpy - gmgmk * v->cqx) * rho;
  ca[3][1] = -v->xy;
  ca[3][2] = gamm1 * v->cpy - gammk * v->cqx;
  ca[3][3] = ca[1][1];
  ca[3][4] = ca[0][1];
  ca[4][0] = -(two * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * v->xz +
               gm1sq * gm1sq * v->wy) * rho2;
  ca[4][1] = ca[3][0];
  ca[4][2] = -(gammk * gamm1 * twgm1 * a0pq + gam * gammk * gammk * v->xz +
               gamm1 * gm1sq * v->wy) * rho;
  ca[4][3] = ca[1][0];
  ca[4][4] = ca[0][0];
  double t = -two * wvno2;
  ca[2][0] = t * ca[4][2];
  ca[2][1] = t * ca[3][2];
  ca[2][2] = v->a0 + two * (v->cpcq - ca[0][0]);
  ca[2][3] = t * ca[1][2];
  ca[2][4] = t * ca[0][2];
}

/* surfdisp96.f:1022-1047 normc (the log of the norm is never read here) */
static void normc5(double ee[5]) {
  double t1 = 0.0;
  for (int i = 0; i < 5; ++i)
    if (fabs(ee[i]) > t1) t1 = fabs(ee[i]);
  if (t1 < 1.0e-40) t1 = 1.0;
  for (int i = 0; i < 5; ++i) ee[i] = ee[i] / t1;
}

/* surfdisp96.f:798-898 dltar4: Rayleigh period equation */
static double dltar4(const chain_ctx *x, double wvno, double omga) {
  int mmax = x->mmax;
  double omega = omga;
  if (omega < 1.0e-4) omega = 1.0e-4;
  double wvno2 = wvno * wvno;
  double xka = omega / (double)x->a[mmax - 1];
  double xkb = omega / (double)x->b[mmax - 1];
  double wvnop = wvno + xka;
  double wvnom = fabs(wvno - xka);
  double ra = sqrt(wvnop * wvnom);
  wvnop = wvno + xkb;
  wvnom = fabs(wvno - xkb);
  double rb = sqrt(wvnop * wvnom);
  double t = (double)x->b[mmax - 1] / omega;
  double gammk = 2.0 * t * t;
  double gam = gammk * wvno2;
  double gamm1 = gam - 1.0;
  double rho1 = (double)x->rho[mmax - 1];
  double e[5], ee[5], ca[5][5];
  e[0] = rho1 * rho1 * (gamm1 * gamm1 - gam * gammk * ra * rb);
  e[1] = -rho1 * ra;
  e[2] = rho1 * (gamm1 - gammk * ra * rb);
  e[3] = rho1 * rb;
  e[4] = wvno2 - ra * rb;
  for (int m = mmax - 2; m >= x->llw - 1; --m) {
    xka = omega / (double)x->a[m];
    xkb = omega / (double)x->b[m];
    t = (double)x->b[m] / omega;
    gammk = 2.0 * t * t;
    gam = gammk * wvno2;
    wvnop = wvno + xka;
    wvnom = fabs(wvno - xka);
    ra = sqrt(wvnop * wvnom);
    wvnop = wvno + xkb;
    wvnom = fabs(wvno - xkb);
    rb = sqrt(wvnop * wvnom);
    double dpth = (double)x->d[m];
    rho1 = (double)x->rho[m];
    double p = ra * dpth;
    double q = rb * dpth;
    double w, cosp;
    var_out v;
    var(p, q, ra, rb, wvno, xka, xkb, dpth, &w, &cosp, &v);
    dnka(ca, wvno2, gam, gammk, rho1, &v);
    for (int i = 0; i < 5; ++i) {
      double cr = 0.0;
      for (int j = 0; j < 5; ++j) cr = cr + e[j] * ca[j][i];
      ee[i] = cr;
    }
    normc5(ee);
    for (int i = 0; i < 5; ++i) e[i] = ee[i];
  }
  if (x->llw != 1) {
    xka = omega / (double)x->a[0];
    wvnop = wvno + xka;
    wvnom = fabs(wvno - xka);
    ra = sqrt(wvnop * wvnom);
    double dpth = (double)x->d[0];
    rho1 = (double)x->rho[0];
    double p = ra * dpth;
    const double znul = 1.0e-05;
    double w, cosp;
    var_out v;
    var(p, znul, ra, znul, wvno, xka, znul, dpth, &w, &cosp, &v);
    double w0 = -rho1 * w;
    return cosp * e[0] + w0 * e[1];
  }
  return e[0];
}

/* surfdisp96.f:712-732 dltar */
static double dltar(chain_ctx *x, double wvno, double omega, int kk) {
  x->nevals++;
  if (kk == 1) return dltar1(x, wvno, omega);
  return dltar4(x, wvno, omega);
}

/* surfdisp96.f:696-708 half */
static void half(chain_ctx *x, double c1, double c2, double *c3, double *del3,
                 double omega, int ifunc) {
  *c3 = 0.5 * (c1 + c2);
  double wvno = omega / *c3;
  *del3 = dltar(x, wvno, omega, ifunc);
}

/* surfdisp96.f:575-692 nevill: hybrid Neville / bisection polish */
static double nevill(chain_ctx *x, double t, double c1, double c2,
                     double del1, double del2, int ifunc) {
  double xx[20], yy[20];
  double omega = 2.0 * 3.141592653589793 / t;
  double c3, del3;
  int m = 1;
  half(x, c1, c2, &c3, &del3, omega, ifunc);
  int nev = 1;
  int nctrl = 1;
  for (;;) {
    nctrl = nctrl + 1;
    if (nctrl >= 100) break;
    if (c3 < fmin(c1, c2) || c3 > fmax(c1, c2)) {
      nev = 0;
      half(x, c1, c2, &c3, &del3, omega, ifunc);
    }
    double s13 = del1 - del3;
    double s32 = del3 - del2;
    if (fsign1(del3) * fsign1(del1) < 0.0) {
      c2 = c3;
      del2 = del3;
    } else {
      c1 = c3;
      del1 = del3;
    }
    if (fabs(c1 - c2) <= g_tol * c1) break;
    if (fsign1(s13) != fsign1(s32)) nev = 0;
    double ss1 = fabs(del1);
    double s1 = (double)0.01f * ss1;
    double ss2 = fabs(del2);
    double s2 = (double)0.01f * ss2;
    if (s1 > ss2 || s2 > ss1 || nev == 0) {
      half(x, c1, c2, &c3, &del3, omega, ifunc);
      nev = 1;
      m = 1;
    } else {
      if (nev == 2) {
        xx[m] = c3;
        yy[m] = del3;
      } else {
        xx[0] = c1;
        yy[0] = del1;
        xx[1] = c2;
        yy[1] = del2;
        m = 1;
      }
      int bad = 0;
      for (int kk = 1; kk <= m; ++kk) {
        int j = m - kk; /* 0-based j = (m-kk+1)-1 */
        double denom = yy[m] - yy[j];
        if (fabs(denom) < 1.0e-10 * fabs(yy[m])) {
          bad = 1;
          break;
        }
        xx[j] = (-yy[j] * xx[j + 1] + yy[m] * xx[j]) / denom;
      }
      if (!bad) {
        c3 = xx[0];
        double wvno = omega / c3;
        del3 = dltar(x, wvno, omega, ifunc);
        nev = 2;
        m = m + 1;
        if (m > 10) m = 10;
      } else {
        half(x, c1, c2, &c3, &del3, omega, ifunc);
        nev = 1;
        m = 1;
      }
    }
  }
  return c3;
}

/* surfdisp96.f:400-495 getsol: bracket then refine.  Returns iret. */
static int getsol(chain_ctx *x, double t1, double *c1p, double clow, double dc,
                  double cm, float betmx, int ifunc, int ifirst) {
  double twopi = 2.0 * 3.141592653589793;
  double c1 = *c1p, c2;
  double omega = twopi / t1;
  double wvno = omega / c1;
  double del1 = dltar(x, wvno, omega, ifunc), del2;
  if (ifirst == 1) x->del1st = del1;
  double plmn = fsign1(x->del1st) * fsign1(del1);
  int idir;
  if (ifirst == 1)
    idir = 1;
  else if (plmn >= 0.0)
    idir = 1;
  else
    idir = -1;
  for (;;) {
    if (idir > 0)
      c2 = c1 + dc;
    else
      c2 = c1 - dc;
    if (c2 <= clow) {
      idir = 1;
      c1 = clow;
      continue;
    }
    omega = twopi / t1;
    wvno = omega / c2;
    del2 = dltar(x, wvno, omega, ifunc);
    if (fsign1(del1) != fsign1(del2)) break;
    c1 = c2;
    del1 = del2;
    if (c1 < cm) {
      *c1p = c1;
      return -1;
    }
    if (c1 >= ((double)betmx + dc)) {
      *c1p = c1;
      return -1;
    }
  }
  double cn = nevill(x, t1, c1, c2, del1, del2, ifunc);
  c1 = cn;
  *c1p = c1;
  if (c1 > (double)betmx) return -1;
  return 1;
}

/* surfdisp96.f:57-370 */
void or_surfdisp96(const float *thkm, const float *vpm, const float *vsm,
                   const float *rhom, int nlayer, int iflsph, int iwave,
                   int mode, int igr, int kmax, const double *t, double *cg,
                   int *ierr, long long *nevals) {
  chain_ctx X;
  int mmax = nlayer;
  float *buf = (float *)malloc(sizeof(float) * 7 * (size_t)mmax);
  X.mmax = mmax;
  X.d = buf;
  X.a = buf + mmax;
  X.b = buf + 2 * mmax;
  X.rho = buf + 3 * mmax;
  X.rtp = buf + 4 * mmax;
  X.dtp = buf + 5 * mmax;
  X.btp = buf + 6 * mmax;
  X.del1st = 0.0;
  X.dhalf = 0.0f;
  X.nevals = 0;
  double *c = (double *)calloc((size_t)kmax, sizeof(double));
  double *cb = (double *)calloc((size_t)kmax, sizeof(double));
  int nsph = iflsph;
  *ierr = 0;
  for (int i = 0; i < mmax; ++i) {
    X.b[i] = vsm[i];
    X.a[i] = vpm[i];
    X.d[i] = thkm[i];
    X.rho[i] = rhom[i];
  }
  int idispl = 0, idispr = 0;
  if (iwave == 1) idispl = kmax;
  else if (iwave == 2) idispr = kmax;
  int iverb[2] = {0, 0};
  const float sone0 = 1.500f, ddc0 = 0.005f, h0 = 0.005f;
  X.llw = 1;
  if (X.b[0] <= 0.0f) X.llw = 2;
  const double one = 1.0e-2;
  if (nsph == 1) sphere(&X, 0, 0);
  int jmn = 0, jsol = 1;
  float betmx = -1.e20f, betmn = 1.e20f;
  for (int i = 0; i < mmax; ++i) {
    if (X.b[i] > 0.01f && X.b[i] < betmn) {
      betmn = X.b[i];
      jmn = i;
      jsol = 1;
    } else if (X.b[i] <= 0.01f && X.a[i] < betmn) {
      betmn = X.a[i];
      jmn = i;
      jsol = 0;
    }
    if (X.b[i] > betmx) betmx = X.b[i];
  }
  for (int ifunc = 1; ifunc <= 2; ++ifunc) {
    if (ifunc == 1 && idispl <= 0) continue;
    if (ifunc == 2 && idispr <= 0) continue;
    if (nsph == 1) sphere(&X, ifunc, 1);
    float ddc = ddc0, sone = sone0, h = h0;
    if (sone < 0.01f) sone = 2.0f;
    double onea = (double)sone;
    float cc1;
    if (jsol == 0)
      cc1 = betmn;
    else
      cc1 = gtsolh(X.a[jmn], X.b[jmn]);
    cc1 = 0.95f * cc1;
    cc1 = 0.90f * cc1;
    double cc = (double)cc1;
    double dc = fabs((double)ddc);
    double c1 = cc, clow = cc, cm = cc;
    for (int i = 0; i < kmax; ++i) {
      cb[i] = 0.0;
      c[i] = 0.0;
    }
    int ift = 999;
    for (int iq = 1; iq <= mode; ++iq) {
      int is = 1, ie = kmax, k;
      int failed = 0;
      for (k = is; k <= ie; ++k) {
        if (k >= ift) {
          failed = 1;
          break;
        }
        double t1 = t[k - 1];
        float t1a = 0.f, t1b = 0.f;
        if (igr > 0) {
          t1a = (float)(t1 / (double)(1.0f + h));
          t1b = (float)(t1 / (double)(1.0f - h));
          t1 = (double)t1a;
        } else {
          t1a = (float)t1;
        }
        int ifirst;
        if (k == is && iq == 1) {
          c1 = cc;
          clow = cc;
          ifirst = 1;
        } else if (k == is && iq > 1) {
          c1 = c[is - 1] + one * dc;
          clow = c1;
          ifirst = 1;
        } else if (k > is && iq > 1) {
          ifirst = 0;
          clow = c[k - 1] + one * dc;
          c1 = c[k - 2];
          if (c1 < clow) c1 = clow;
        } else {
          ifirst = 0;
          c1 = c[k - 2] - onea * dc;
          clow = cm;
        }
        int iret = getsol(&X, t1, &c1, clow, dc, cm, betmx, ifunc, ifirst);
        if (iret == -1) {
          failed = 1;
          break;
        }
        c[k - 1] = c1;
        if (igr > 0) {
          t1 = (double)t1b;
          ifirst = 0;
          clow = cb[k - 1] + one * dc;
          c1 = c1 - onea * dc;
          iret = getsol(&X, t1, &c1, clow, dc, cm, betmx, ifunc, ifirst);
          if (iret == -1) c1 = c[k - 1];
          cb[k - 1] = c1;
        } else {
          c1 = 0.0;
        }
        float cc0 = (float)c[k - 1];
        float ccg = (float)c1;
        if (igr == 0) {
          cg[k - 1] = (g_tol < 1.0e-6) ? c[k - 1] : (double)cc0;
        } else {
          float gvel = (1 / t1a - 1 / t1b) / (1 / (t1a * cc0) - 1 / (t1b * ccg));
          cg[k - 1] = gvel;
        }
      }
      if (!failed) continue;
      /* label 1700 / 1750 */
      if (iq <= 1) {
        if (iverb[ifunc - 1] == 0) {
          iverb[ifunc - 1] = 1;
          *ierr = 1;
        }
      }
      ift = k;
      for (int i = k; i <= ie; ++i) cg[i - 1] = 0.0;
    }
  }
  if (nevals) *nevals += X.nevals;
  free(c);
  free(cb);
  free(buf);
}

/* surfdisp.cpp:17-51 _flat2sphere, swdtp coded 0=Rc 1=Rg 2=Lc 3=Lg */
static double flat2sphere(double t, double c, int swdtp) {
  const double ar = 6371.0;
  double omega = 2.0 * M_PI / t;
  double tm;
  if (swdtp >= 2)
    tm = 1. + pow(1.5 * c / (ar * omega), 2);
  else
    tm = 1. + pow(0.5 * c / (ar * omega), 2);
  tm = sqrt(tm);
  if (swdtp == 0 || swdtp == 2) return c / tm;
  return c * tm;
}

/* surfdisp.cpp:64-116 */
int or_surfdisp(const float *thk, const float *vp, const float *vs,
                const float *rho, int nlayer, const double *t, double *cg,
                int kmax, int swdtp, int mode, int sphere_, int keep_flat) {
  int iwave = (swdtp >= 2) ? 1 : 2;
  int igr = (swdtp == 1 || swdtp == 3) ? 1 : 0;
  int ifsph = sphere_ ? 1 : 0;
  int ierr;
  or_surfdisp96(thk, vp, vs, rho, nlayer, ifsph, iwave, mode + 1, igr, kmax, t,
                cg, &ierr, NULL);
  if (ierr != 0) {
    for (int i = 0; i < kmax; ++i) {
      if (cg[i] == 0.0 || isnan(cg[i])) {
        or_surfdisp96(thk, vp, vs, rho, nlayer, ifsph, iwave, mode + 1, igr, 1,
                      &t[i], &cg[i], &ierr, NULL);
        if (ierr != 0 || cg[i] == 0.0 || isnan(cg[i])) return 1;
      }
    }
  }
  if (sphere_ && !keep_flat)
    for (int i = 0; i < kmax; ++i) cg[i] = flat2sphere(t[i], cg[i], swdtp);
  return 0;
}

/* empirical.cpp:9-18 */
void or_empirical_relation(float vsz, float *vpz, float *rhoz) {
  float vp = (float)(0.9409 + 2.0947 * vsz - 0.8206 * pow(vsz, 2) +
                     0.2683 * pow(vsz, 3) - 0.0251 * pow(vsz, 4));
  *vpz = vp;
  *rhoz = (float)(1.6612 * vp - 0.4721 * pow(vp, 2) + 0.0671 * pow(vp, 3) -
                  0.0043 * pow(vp, 4) + 0.000106 * pow(vp, 5));
}

/* empirical.cpp:26-34 */
void or_empirical_deriv(float vp, float vs, float *drda, float *dadb) {
  *drda = (float)(1.6612 - 0.4721 * 2 * vp + 0.0671 * 3 * pow(vp, 2) -
                  0.0043 * 4 * pow(vp, 3) + 0.000106 * 5 * pow(vp, 4));
  *dadb = (float)(2.0947 - 0.8206 * 2 * vs + 0.2683 * 3 * pow(vs, 2) -
                  0.0251 * 4 * pow(vs, 3));
}

/* Flattened root-search model as surfdisp96 builds it (sphere, then the
 * betmn/betmx scan and the starting velocity), for tests of the GPU prep. */
void or_chain_model(const float *thk, const float *vp, const float *vs, const float *rho, int n,
                    int iflsph, float *d, float *a, float *b, float *rhoR, float *rhoL,
                    float *cc, float *betmx_out) {
  chain_ctx X;
  float *buf = (float *)malloc(sizeof(float) * 7 * (size_t)n);
  X.mmax = n;
  X.d = buf; X.a = buf + n; X.b = buf + 2 * n; X.rho = buf + 3 * n;
  X.rtp = buf + 4 * n; X.dtp = buf + 5 * n; X.btp = buf + 6 * n;
  for (int i = 0; i < n; ++i) {
    X.d[i] = thk[i]; X.a[i] = vp[i]; X.b[i] = vs[i]; X.rho[i] = rho[i];
  }
  if (iflsph == 1) sphere(&X, 0, 0);
  int jmn = 0, jsol = 1;
  float bx = -1.e20f, bn = 1.e20f;
  for (int i = 0; i < n; ++i) {
    if (X.b[i] > 0.01f && X.b[i] < bn) { bn = X.b[i]; jmn = i; jsol = 1; }
    else if (X.b[i] <= 0.01f && X.a[i] < bn) { bn = X.a[i]; jmn = i; jsol = 0; }
    if (X.b[i] > bx) bx = X.b[i];
  }
  float c1 = (jsol == 0) ? bn : gtsolh(X.a[jmn], X.b[jmn]);
  c1 = 0.95f * c1;
  c1 = 0.90f * c1;
  *cc = c1;
  *betmx_out = bx;
  for (int w = 1; w <= 2; ++w) {
    if (iflsph == 1) sphere(&X, w, 1);
    for (int i = 0; i < n; ++i) (w == 1 ? rhoL : rhoR)[i] = X.rho[i];
    for (int i = 0; i < n; ++i) X.rho[i] = rho[i];
  }
  for (int i = 0; i < n; ++i) { d[i] = X.d[i]; a[i] = X.a[i]; b[i] = X.b[i]; }
  free(buf);
}
