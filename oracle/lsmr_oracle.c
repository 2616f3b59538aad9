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

/* parallel.cpp:30-43, serial branch: y += A x, row by row */
static void aprod1(const float *x, float *y, const int *ind, const int *ptr, int m,
                   const float *val) {
  for (int i = 0; i < m; i++)
    for (int j = ptr[i]; j < ptr[i + 1]; j++) y[i] += val[j] * x[ind[j]];
}
/* parallel.cpp:70-83, serial branch: x += A^T y, row by row */
static void aprod2(float *x, const float *y, const int *ind, const int *ptr, int m,
                   const float *val) {
  for (int i = 0; i < m; i++)
    for (int j = ptr[i]; j < ptr[i + 1]; j++) x[ind[j]] += val[j] * y[i];
}

/* lsmrModule_csr.f90:358-393 dnrm2: scaled sum of squares */
static float nrm2(int n, const float *x) {
  if (n < 1) return 0.0f;
  if (n == 1) return fabsf(x[0]);
  float scale = 0.0f, ssq = 1.0f;
  for (int i = 0; i < n; i++) {
    if (x[i] != 0.0f) {
      float a = fabsf(x[i]);
      if (scale < a) {
        float r = scale / a;
        ssq = 1.0f + ssq * (r * r);
        scale = a;
      } else {
        float r = a / scale;
        ssq = ssq + r * r;
      }
    }
  }
  return scale * sqrtf(ssq);
}

/* lsmrModule_csr.f90:1151-1163 d2norm */
static float d2norm(float a, float b) {
  float scale = fabsf(a) + fabsf(b);
  if (scale == 0.0f) return 0.0f;
  float p = a / scale, q = b / scale;
  return scale * sqrtf(p * p + q * q);
}

static void scal(int n, float s, float *x) {
  for (int i = 0; i < n; i++) x[i] = s * x[i];
}

/* lsmrModule_csr.f90:516-1148 LSMR (istop codes and the stopping tests as
 * there; verbose output left out) */
void or_lsmr_csr(int m, int n, const float *val, const int *ind, const int *ptr, const float *b,
                 float damp, float atol, float btol, float conlim, int itnlim, int localSize,
                 float *x, int *istop_out, int *itn_out, float *normA_out, float *condA_out,
                 float *normr_out, float *normAr_out, float *normx_out) {
  int localVecs = localSize < m ? localSize : m;
  if (n < localVecs) localVecs = n;
  float *u = calloc((size_t)m, sizeof(float));
  float *v = calloc((size_t)n, sizeof(float));
  float *h = calloc((size_t)n, sizeof(float));
  float *hbar = calloc((size_t)n, sizeof(float));
  float *lv = calloc((size_t)n * (localVecs > 0 ? localVecs : 1), sizeof(float));
  const int damped = damp > 0.0f;
  memcpy(u, b, sizeof(float) * (size_t)m);
  for (int i = 0; i < n; i++) x[i] = 0.0f;
  float alpha = 0.0f, beta = nrm2(m, u);
  if (beta > 0.0f) {
    scal(m, 1.0f / beta, u);
    aprod2(v, u, ind, ptr, m, val);
    alpha = nrm2(n, v);
  }
  if (alpha > 0.0f) scal(n, 1.0f / alpha, v);
  float normAr = alpha * beta;
  int istop = 0, itn = 0;
  float normA = 0.0f, condA = 0.0f, normr = beta, normx = 0.0f;
  if (normAr == 0.0f) goto done;
  /* :868-875 local reorthogonalisation queue */
  int localOrtho = 0, localPointer = 0, localFull = 0;
  if (localVecs > 0) {
    localPointer = 1;
    localOrtho = 1;
    memcpy(lv, v, sizeof(float) * (size_t)n);
  }
  /* :877-912 */
  float zetabar = alpha * beta, alphabar = alpha, rho = 1, rhobar = 1, cbar = 1, sbar = 0;
  memcpy(h, v, sizeof(float) * (size_t)n);
  float betadd = beta, betad = 0, rhodold = 1, tautildeold = 0, thetatilde = 0, zeta = 0, d = 0;
  float normA2 = alpha * alpha, maxrbar = 0.0f, minrbar = 1e30f;
  const float normb = beta;
  float ctol = 0.0f;
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
        memcpy(lv + (size_t)(localPointer - 1) * n, v, sizeof(float) * (size_t)n);
      }
      scal(n, -beta, v);
      aprod2(v, u, ind, ptr, m, val);
      if (localOrtho) { /* :1177-1197 localVOrtho */
        int lim = localFull ? localVecs : localPointer;
        for (int c = 0; c < lim; c++) {
          const float *q = lv + (size_t)c * n;
          float dd = 0.0f;
          for (int i = 0; i < n; i++) dd = dd + v[i] * q[i];
          for (int i = 0; i < n; i++) v[i] = v[i] - dd * q[i];
        }
      }
      alpha = nrm2(n, v);
      if (alpha > 0.0f) scal(n, 1.0f / alpha, v);
    }
    /* :958-1000 plane rotations */
    float alphahat = d2norm(alphabar, damp);
    float chat = alphabar / alphahat;
    float shat = damp / alphahat;
    float rhoold = rho;
    rho = d2norm(alphahat, beta);
    float c = alphahat / rho;
    float s = beta / rho;
    float thetanew = s * alpha;
    alphabar = c * alpha;
    float rhobarold = rhobar;
    float zetaold = zeta;
    float thetabar = sbar * rho;
    float rhotemp = cbar * rho;
    rhobar = d2norm(cbar * rho, thetanew);
    cbar = cbar * rho / rhobar;
    sbar = thetanew / rhobar;
    zeta = cbar * zetabar;
    zetabar = -sbar * zetabar;
    /* :1003-1006 h, hbar, x */
    const float fh = thetabar * rho / (rhoold * rhobarold);
    const float fx = zeta / (rho * rhobar);
    const float fn = thetanew / rho;
    for (int i = 0; i < n; i++) {
      hbar[i] = h[i] - fh * hbar[i];
      x[i] = x[i] + fx * hbar[i];
      h[i] = v[i] - fn * h[i];
    }
    /* :1009-1043 estimates of ||r||, ||A||, cond(A) */
    float betaacute = chat * betadd;
    float betacheck = -shat * betadd;
    float betahat = c * betaacute;
    betadd = -s * betaacute;
    float thetatildeold = thetatilde;
    float rhotildeold = d2norm(rhodold, thetabar);
    float ctildeold = rhodold / rhotildeold;
    float stildeold = thetabar / rhotildeold;
    thetatilde = stildeold * rhobar;
    rhodold = ctildeold * rhobar;
    betad = -stildeold * betad + ctildeold * betahat;
    tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold;
    float taud = (zeta - thetatilde * tautildeold) / rhodold;
    d = d + betacheck * betacheck;
    {
      float bd = betad - taud;
      normr = sqrtf(d + bd * bd + betadd * betadd);
    }
    normA2 = normA2 + beta * beta;
    normA = sqrtf(normA2);
    normA2 = normA2 + alpha * alpha;
    maxrbar = fmaxf(maxrbar, rhobarold);
    if (itn > 1) minrbar = fminf(minrbar, rhobarold);
    condA = fmaxf(maxrbar, rhotemp) / fminf(minrbar, rhotemp);
    /* :1046-1078 convergence tests */
    normAr = fabsf(zetabar);
    normx = nrm2(n, x);
    float test1 = normr / normb;
    float test2 = normAr / (normA * normr);
    float test3 = 1.0f / condA;
    float t1 = test1 / (1.0f + normA * normx / normb);
    float rtol = btol + atol * normA * normx / normb;
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
