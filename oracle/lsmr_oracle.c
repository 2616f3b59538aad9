/* lsmr_oracle.c -- test infrastructure (the oracle; never part of the
 * product): a plain-C restatement of JDSurfG's LSMR on a CSR matrix,
 * src/shared/lsmrModule_csr.f90 (LSMR of Fong and Saunders with local
 * reorthogonalisation, real(dp) = c_float, :15-25) and its products
 * src/shared/parallel.cpp:30-100 (aprod1 y += A x, aprod2 x += A^T y, serial
 * row order).  Every vector and scalar is float, as in the reference build;
 * operations are written in the reference's order (no contraction). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "swd_oracle.h"

/* real_t is float (the reference build); lsmr_oracle_f64.c includes this
 * file with LSMR_F64 defined for a double twin, or_lsmr_csr_f64, which
 * measures how far float accumulation moves the reference's iterates. */
#ifndef LSMR_F64
typedef float real_t;
#define RABS fabsf
#define RSQRT sqrtf
#define RMAX fmaxf
#define RMIN fminf
#define LSMR_FN or_lsmr_csr
#else
typedef double real_t;
#define RABS fabs
#define RSQRT sqrt
#define RMAX fmax
#define RMIN fmin
#define LSMR_FN or_lsmr_csr_f64
#endif

/* parallel.cpp:30-43, serial branch: y += A x, row by row */
static void aprod1(const real_t *x, real_t *y, const int *ind, const int *ptr, int m,
                   const real_t *val) {
  for (int i = 0; i < m; i++)
    for (int j = ptr[i]; j < ptr[i + 1]; j++) y[i] += val[j] * x[ind[j]];
}
/* parallel.cpp:70-83, serial branch: x += A^T y, row by row */
static void aprod2(real_t *x, const real_t *y, const int *ind, const int *ptr, int m,
                   const real_t *val) {
  for (int i = 0; i < m; i++)
    for (int j = ptr[i]; j < ptr[i + 1]; j++) x[ind[j]] += val[j] * y[i];
}

/* lsmrModule_csr.f90:358-393 dnrm2: scaled sum of squares */
static real_t nrm2(int n, const real_t *x) {
  if (n < 1) return 0.0f;
  if (n == 1) return RABS(x[0]);
  real_t scale = 0.0f, ssq = 1.0f;
  for (int i = 0; i < n; i++) {
    if (x[i] != 0.0f) {
      real_t a = RABS(x[i]);
      if (scale < a) {
        real_t r = scale / a;
        ssq = 1.0f + ssq * (r * r);
        scale = a;
      } else {
        real_t r = a / scale;
        ssq = ssq + r * r;
      }
    }
  }
  return scale * RSQRT(ssq);
}

/* lsmrModule_csr.f90:1151-1163 d2norm */
static real_t d2norm(real_t a, real_t b) {
  real_t scale = RABS(a) + RABS(b);
  if (scale == 0.0f) return 0.0f;
  real_t p = a / scale, q = b / scale;
  return scale * RSQRT(p * p + q * q);
}

static void scal(int n, real_t s, real_t *x) {
  for (int i = 0; i < n; i++) x[i] = s * x[i];
}

/* lsmrModule_csr.f90:516-1148 LSMR (istop codes and the stopping tests as
 * there; verbose output left out) */
void LSMR_FN(int m, int n, const real_t *val, const int *ind, const int *ptr, const real_t *b,
                 real_t damp, real_t atol, real_t btol, real_t conlim, int itnlim, int localSize,
                 real_t *x, int *istop_out, int *itn_out, real_t *normA_out, real_t *condA_out,
                 real_t *normr_out, real_t *normAr_out, real_t *normx_out) {
  int localVecs = localSize < m ? localSize : m;
  if (n < localVecs) localVecs = n;
  real_t *u = calloc((size_t)m, sizeof(real_t));
  real_t *v = calloc((size_t)n, sizeof(real_t));
  real_t *h = calloc((size_t)n, sizeof(real_t));
  real_t *hbar = calloc((size_t)n, sizeof(real_t));
  real_t *lv = calloc((size_t)n * (localVecs > 0 ? localVecs : 1), sizeof(real_t));
  const int damped = damp > 0.0f;
  memcpy(u, b, sizeof(real_t) * (size_t)m);
  for (int i = 0; i < n; i++) x[i] = 0.0f;
  real_t alpha = 0.0f, beta = nrm2(m, u);
  if (beta > 0.0f) {
    scal(m, 1.0f / beta, u);
    aprod2(v, u, ind, ptr, m, val);
    alpha = nrm2(n, v);
  }
  if (alpha > 0.0f) scal(n, 1.0f / alpha, v);
  real_t normAr = alpha * beta;
  int istop = 0, itn = 0;
  real_t normA = 0.0f, condA = 0.0f, normr = beta, normx = 0.0f;
  if (normAr == 0.0f) goto done;
  /* :868-875 local reorthogonalisation queue */
  int localOrtho = 0, localPointer = 0, localFull = 0;
  if (localVecs > 0) {
    localPointer = 1;
    localOrtho = 1;
    memcpy(lv, v, sizeof(real_t) * (size_t)n);
  }
  /* :877-912 */
  real_t zetabar = alpha * beta, alphabar = alpha, rho = 1, rhobar = 1, cbar = 1, sbar = 0;
  memcpy(h, v, sizeof(real_t) * (size_t)n);
  real_t betadd = beta, betad = 0, rhodold = 1, tautildeold = 0, thetatilde = 0, zeta = 0, d = 0;
  real_t normA2 = alpha * alpha, maxrbar = 0.0f, minrbar = 1e30f;
  const real_t normb = beta;
  real_t ctol = 0.0f;
  if (conlim > 0.0f) ctol = 1.0f / conlim;
  normr = beta;
  for (;;) {
    itn = itn + 1;
    /* :932-955 bidiagonalisation */
    scal(m, -alpha, u);
    aprod1(v, u, ind, ptr, m, val);
    beta = nrm2(m, u);
    if (beta > 0.0f) {
      scal(m, 1.0f / beta, u);
      if (localOrtho) { /* :1166-1175 localVEnqueue */
        if (localPointer < localVecs) {
          localPointer = localPointer + 1;
        } else {
          localPointer = 1;
          localFull = 1;
        }
        memcpy(lv + (size_t)(localPointer - 1) * n, v, sizeof(real_t) * (size_t)n);
      }
      scal(n, -beta, v);
      aprod2(v, u, ind, ptr, m, val);
      if (localOrtho) { /* :1177-1197 localVOrtho */
        int lim = localFull ? localVecs : localPointer;
        for (int c = 0; c < lim; c++) {
          const real_t *q = lv + (size_t)c * n;
          real_t dd = 0.0f;
          for (int i = 0; i < n; i++) dd = dd + v[i] * q[i];
          for (int i = 0; i < n; i++) v[i] = v[i] - dd * q[i];
        }
      }
      alpha = nrm2(n, v);
      if (alpha > 0.0f) scal(n, 1.0f / alpha, v);
    }
    /* :958-1000 plane rotations */
    real_t alphahat = d2norm(alphabar, damp);
    real_t chat = alphabar / alphahat;
    real_t shat = damp / alphahat;
    real_t rhoold = rho;
    rho = d2norm(alphahat, beta);
    real_t c = alphahat / rho;
    real_t s = beta / rho;
    real_t thetanew = s * alpha;
    alphabar = c * alpha;
    real_t rhobarold = rhobar;
    real_t zetaold = zeta;
    real_t thetabar = sbar * rho;
    real_t rhotemp = cbar * rho;
    rhobar = d2norm(cbar * rho, thetanew);
    cbar = cbar * rho / rhobar;
    sbar = thetanew / rhobar;
    zeta = cbar * zetabar;
    zetabar = -sbar * zetabar;
    /* :1003-1006 h, hbar, x */
    const real_t fh = thetabar * rho / (rhoold * rhobarold);
    const real_t fx = zeta / (rho * rhobar);
    const real_t fn = thetanew / rho;
    for (int i = 0; i < n; i++) {
      hbar[i] = h[i] - fh * hbar[i];
      x[i] = x[i] + fx * hbar[i];
      h[i] = v[i] - fn * h[i];
    }
    /* :1009-1043 estimates of ||r||, ||A||, cond(A) */
    real_t betaacute = chat * betadd;
    real_t betacheck = -shat * betadd;
    real_t betahat = c * betaacute;
    betadd = -s * betaacute;
    real_t thetatildeold = thetatilde;
    real_t rhotildeold = d2norm(rhodold, thetabar);
    real_t ctildeold = rhodold / rhotildeold;
    real_t stildeold = thetabar / rhotildeold;
    thetatilde = stildeold * rhobar;
    rhodold = ctildeold * rhobar;
    betad = -stildeold * betad + ctildeold * betahat;
    tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold;
    real_t taud = (zeta - thetatilde * tautildeold) / rhodold;
    d = d + betacheck * betacheck;
    {
      real_t bd = betad - taud;
      normr = RSQRT(d + bd * bd + betadd * betadd);
    }
    normA2 = normA2 + beta * beta;
    normA = RSQRT(normA2);
    normA2 = normA2 + alpha * alpha;
    maxrbar = RMAX(maxrbar, rhobarold);
    if (itn > 1) minrbar = RMIN(minrbar, rhobarold);
    condA = RMAX(maxrbar, rhotemp) / RMIN(minrbar, rhotemp);
    /* :1046-1078 convergence tests */
    normAr = RABS(zetabar);
    normx = nrm2(n, x);
    real_t test1 = normr / normb;
    real_t test2 = normAr / (normA * normr);
    real_t test3 = 1.0f / condA;
    real_t t1 = test1 / (1.0f + normA * normx / normb);
    real_t rtol = btol + atol * normA * normx / normb;
    if (itn >= itnlim) istop = 7;
    if (1.0f + test3 <= 1.0f) istop = 6;
    if (1.0f + test2 <= 1.0f) istop = 5;
    if (1.0f + t1 <= 1.0f) istop = 4;
    if (test3 <= ctol) istop = 3;
    if (test2 <= atol) istop = 2;
    if (test1 <= rtol) istop = 1;
    if (istop != 0) break;
  }
done:
  if (damped && istop == 2) istop = 3; /* :1104 */
  *istop_out = istop;
  *itn_out = itn;
  *normA_out = normA;
  *condA_out = condA;
  *normr_out = normr;
  *normAr_out = normAr;
  *normx_out = normx;
  free(u);
  free(v);
  free(h);
  free(hbar);
  free(lv);
}
