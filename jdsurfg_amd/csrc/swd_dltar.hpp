// swd_dltar.hpp -- period equations of the root search (surfdisp96.f:712-898):
// dltar1 (Love, Haskell SH) and dltar4 (Rayleigh, Dunkin compound
// matrices), reading the float4 {d, a, b, rho} layer stack of a column.
#pragma once
#include "swd_common.hpp"

// The recursion rounds like the reference build: no FP contraction.
#pragma clang fp contract(off)

// ----------------------------------------------------------------------------
// Period equations (surfdisp96.f:734-898), model read from the float4 stack.

namespace jdsg_pe {

// dltar1: Love (Haskell SH) from the half-space to the surface
JDSG_DEV double dltar1(const float4 *__restrict__ st, int stride, int mmax, int llw,
                       double wvno, double omega) {
  float4 h = st[(size_t)(mmax - 1) * stride];
  double beta1 = (double)h.z;
  double rho1 = (double)h.w;
  double xkb = omega / beta1;
  double wvnop = wvno + xkb;
  double wvnom = fabs(wvno - xkb);
  double rb = sqrt(wvnop * wvnom);
  double e1 = rho1 * rb;
  double e2 = 1.0 / (beta1 * beta1);
  float4 Lnext = (mmax - 2 >= 0) ? st[(size_t)(mmax - 2) * stride] : h;
  for (int m = mmax - 2; m >= llw - 1; --m) {
    float4 L = Lnext;
    if (m > 0) Lnext = st[(size_t)(m - 1) * stride];  // prefetch the next layer up
    beta1 = (double)L.z;
    rho1 = (double)L.w;
    double xmu = rho1 * beta1 * beta1;
    xkb = omega / beta1;
    wvnop = wvno + xkb;
    wvnom = fabs(wvno - xkb);
    rb = sqrt(wvnop * wvnom);
    double dm = (double)L.x;
    double q = dm * rb;
    double sinq, cosq, y, z;
    if (wvno < xkb) {
      sincos(q, &sinq, &cosq);
      y = sinq / rb;
      z = -rb * sinq;
    } else if (wvno == xkb) {
      cosq = 1.0;
      y = dm;
      z = 0.0;
    } else {
      double fac = (q < 16) ? exp(-2.0 * q) : 0.0;
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

// var (surfdisp96.f:901-1018): one of sin/cos, (cosh,sinh)*exp(-p), or the
// coincident case for each of P and S
JDSG_DEV void var_ps(double p, double ra, double wvno, double xka, double dpth,
                     double &cosp, double &w, double &x, double &pex) {
  pex = 0.0;
  if (wvno < xka) {
    double sinp;
    sincos(p, &sinp, &cosp);
    w = sinp / ra;
    x = -ra * sinp;
  } else if (wvno == xka) {
    cosp = 1.0;
    w = dpth;
    x = 0.0;
  } else {
    pex = p;
    double fac = (p < 16) ? exp(-2.0 * p) : 0.0;
    cosp = (1.0 + fac) * 0.5;
    double sinp = (1.0 - fac) * 0.5;
    w = sinp / ra;
    x = ra * sinp;
  }
}

// dltar4: Rayleigh, Dunkin compound matrices from the half-space upward
JDSG_DEV double dltar4(const float4 *__restrict__ st, int stride, int mmax, int llw,
                       double wvno, double omga) {
  double omega = omga;
  if (omega < 1.0e-4) omega = 1.0e-4;
  double wvno2 = wvno * wvno;
  float4 h = st[(size_t)(mmax - 1) * stride];
  double xka = omega / (double)h.y;
  double xkb = omega / (double)h.z;
  double wvnop = wvno + xka;
  double wvnom = fabs(wvno - xka);
  double ra = sqrt(wvnop * wvnom);
  wvnop = wvno + xkb;
  wvnom = fabs(wvno - xkb);
  double rb = sqrt(wvnop * wvnom);
  double t = (double)h.z / omega;
  double gammk = 2.0 * t * t;
  double gam = gammk * wvno2;
  double gamm1 = gam - 1.0;
  double rho1 = (double)h.w;
  double e0 = rho1 * rho1 * (gamm1 * gamm1 - gam * gammk * ra * rb);
  double e1 = -rho1 * ra;
  double e2 = rho1 * (gamm1 - gammk * ra * rb);
  double e3 = rho1 * rb;
  double e4 = wvno2 - ra * rb;
  float4 Lnext = (mmax - 2 >= 0) ? st[(size_t)(mmax - 2) * stride] : h;
  for (int m = mmax - 2; m >= llw - 1; --m) {
    float4 L = Lnext;
    if (m > 0) Lnext = st[(size_t)(m - 1) * stride];  // prefetch the next layer up
    xka = omega / (double)L.y;
    xkb = omega / (double)L.z;
    t = (double)L.z / omega;
    gammk = 2.0 * t * t;
    gam = gammk * wvno2;
    wvnop = wvno + xka;
    wvnom = fabs(wvno - xka);
    ra = sqrt(wvnop * wvnom);
    wvnop = wvno + xkb;
    wvnom = fabs(wvno - xkb);
    rb = sqrt(wvnop * wvnom);
    double dpth = (double)L.x;
    rho1 = (double)L.w;
    double p = ra * dpth;
    double q = rb * dpth;
    double cosp, w, x, pex, cosq, y, z, sex;
    var_ps(p, ra, wvno, xka, dpth, cosp, w, x, pex);
    var_ps(q, rb, wvno, xkb, dpth, cosq, y, z, sex);
    double exa = pex + sex;
    double a0 = (exa < 60.0) ? exp(-exa) : 0.0;
    double cpcq = cosp * cosq, cpy = cosp * y, cpz = cosp * z, cqw = cosq * w,
           cqx = cosq * x, xy = x * y, xz = x * z, wy = w * y, wz = w * z;
    // dnka (surfdisp96.f:1051-1094)
    const double two = 2.0;
    double gamm1_ = gam - 1.0;
    double twgm1 = gam + gamm1_;
    double gmgmk = gam * gammk;
    double gmgm1 = gam * gamm1_;
    double gm1sq = gamm1_ * gamm1_;
    double rho2 = rho1 * rho1;
    double a0pq = a0 - cpcq;
    double c00 = cpcq - two * gmgm1 * a0pq - gmgmk * xz - wvno2 * gm1sq * wy;
    double c01 = (wvno2 * cpy - cqx) / rho1;
    double c02 = -(twgm1 * a0pq + gammk * xz + wvno2 * gamm1_ * wy) / rho1;
    double c03 = (cpz - wvno2 * cqw) / rho1;
    double c04 = -(two * wvno2 * a0pq + xz + wvno2 * wvno2 * wy) / rho2;
    double c10 = (gmgmk * cpz - gm1sq * cqw) * rho1;
    double c11 = cpcq;
    double c12 = gammk * cpz - gamm1_ * cqw;
    double c13 = -wz;
    double c14 = c03;
    double c30 = (gm1sq * cpy - gmgmk * cqx) * rho1;
    double c31 = -xy;
    double c32 = gamm1_ * cpy - gammk * cqx;
    double c33 = c11;
    double c34 = c01;
    double c40 = -(two * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * xz + gm1sq * gm1sq * wy) * rho2;
    double c41 = c30;
    double c42 = -(gammk * gamm1_ * twgm1 * a0pq + gam * gammk * gammk * xz +
                   gamm1_ * gm1sq * wy) * rho1;
    double c43 = c10;
    double c44 = c00;
    double tt = -two * wvno2;
    double c20 = tt * c42;
    double c21 = tt * c32;
    double c22 = a0 + two * (cpcq - c00);
    double c23 = tt * c12;
    double c24 = tt * c02;
    // ee(i) = sum_j e(j) ca(j,i), in j order
    double ee0 = 0.0, ee1 = 0.0, ee2 = 0.0, ee3 = 0.0, ee4 = 0.0;
    ee0 = ee0 + e0 * c00; ee0 = ee0 + e1 * c10; ee0 = ee0 + e2 * c20; ee0 = ee0 + e3 * c30; ee0 = ee0 + e4 * c40;
    ee1 = ee1 + e0 * c01; ee1 = ee1 + e1 * c11; ee1 = ee1 + e2 * c21; ee1 = ee1 + e3 * c31; ee1 = ee1 + e4 * c41;
    ee2 = ee2 + e0 * c02; ee2 = ee2 + e1 * c12; ee2 = ee2 + e2 * c22; ee2 = ee2 + e3 * c32; ee2 = ee2 + e4 * c42;
    ee3 = ee3 + e0 * c03; ee3 = ee3 + e1 * c13; ee3 = ee3 + e2 * c23; ee3 = ee3 + e3 * c33; ee3 = ee3 + e4 * c43;
    ee4 = ee4 + e0 * c04; ee4 = ee4 + e1 * c14; ee4 = ee4 + e2 * c24; ee4 = ee4 + e3 * c34; ee4 = ee4 + e4 * c44;
    // normc (surfdisp96.f:1022-1047)
    double t1 = 0.0;
    if (fabs(ee0) > t1) t1 = fabs(ee0);
    if (fabs(ee1) > t1) t1 = fabs(ee1);
    if (fabs(ee2) > t1) t1 = fabs(ee2);
    if (fabs(ee3) > t1) t1 = fabs(ee3);
    if (fabs(ee4) > t1) t1 = fabs(ee4);
    if (t1 < 1.0e-40) t1 = 1.0;
    e0 = ee0 / t1;
    e1 = ee1 / t1;
    e2 = ee2 / t1;
    e3 = ee3 / t1;
    e4 = ee4 / t1;
  }
  if (llw != 1) {
    float4 L = st[0];
    xka = omega / (double)L.y;
    wvnop = wvno + xka;
    wvnom = fabs(wvno - xka);
    ra = sqrt(wvnop * wvnom);
    double dpth = (double)L.x;
    rho1 = (double)L.w;
    double p = ra * dpth;
    double cosp, w, x, pex;
    var_ps(p, ra, wvno, xka, dpth, cosp, w, x, pex);
    double w0 = -rho1 * w;
    return cosp * e0 + w0 * e1;
  }
  return e0;
}


}  // namespace jdsg_pe
