// swd_dltar_fast.hpp -- the period equations of surfdisp96.f:712-898 for the
// exact-root polish of the learned-bracket path.  Same mathematics as
// swd_dltar.hpp, arranged for throughput instead of bit-faithful rounding:
// FMA contraction, reciprocals from the hardware estimate refined by two
// Newton steps (shared where one denominator serves several quotients), and
// the three exponentials of a layer reduced to two (exp(-2p) = exp(-p)^2,
// exp(-p-q) = exp(-p) exp(-q)).  The values differ from the faithful form in
// the last bits only; a root polished to 1e-10 relative does not move.
#pragma once
#include "swd_common.hpp"

namespace jdsg_pf {

#pragma clang fp contract(fast)

JDSG_DEV double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// one of sin/cos, (cosh, sinh) * exp(-p), or the coincident case; ep is
// exp(-p) in the evanescent case (the layer's share of the common factor)
JDSG_DEV void var_ps(double p, double ra, double ira, double wvno, double xka, double dpth,
                     double &cosp, double &w, double &x, double &ep) {
  ep = 1.0;
  if (wvno < xka) {
    double sinp;
    sincos(p, &sinp, &cosp);
    w = sinp * ira;
    x = -ra * sinp;
  } else if (wvno == xka) {
    cosp = 1.0;
    w = dpth;
    x = 0.0;
  } else {
    ep = (p < 40.0) ? exp(-p) : 0.0;
    double fac = ep * ep;
    cosp = (1.0 + fac) * 0.5;
    double sinp = (1.0 - fac) * 0.5;
    w = sinp * ira;
    x = ra * sinp;
  }
}

// Love (Haskell SH), half-space to surface
JDSG_DEV double dltar1(const float4 *__restrict__ st, int stride, int mmax, int llw, double wvno,
                       double omega) {
  float4 h = st[(size_t)(mmax - 1) * stride];
  double beta1 = (double)h.z;
  double rho1 = (double)h.w;
  double xkb = omega * rcp(beta1);
  double rb = sqrt((wvno + xkb) * fabs(wvno - xkb));
  double e1 = rho1 * rb;
  double ib = rcp(beta1);
  double e2 = ib * ib;
  float4 Lnext = (mmax - 2 >= 0) ? st[(size_t)(mmax - 2) * stride] : h;
  for (int m = mmax - 2; m >= llw - 1; --m) {
    float4 L = Lnext;
    if (m > 0) Lnext = st[(size_t)(m - 1) * stride];
    beta1 = (double)L.z;
    rho1 = (double)L.w;
    double xmu = rho1 * beta1 * beta1;
    xkb = omega * rcp(beta1);
    rb = sqrt((wvno + xkb) * fabs(wvno - xkb));
    double dm = (double)L.x;
    double q = dm * rb;
    double cosq, y, z;
    if (wvno < xkb) {
      double sinq;
      sincos(q, &sinq, &cosq);
      y = sinq * rcp(rb);
      z = -rb * sinq;
    } else if (wvno == xkb) {
      cosq = 1.0;
      y = dm;
      z = 0.0;
    } else {
      double fac = (q < 16) ? exp(-2.0 * q) : 0.0;
      cosq = (1.0 + fac) * 0.5;
      double sinq = (1.0 - fac) * 0.5;
      y = sinq * rcp(rb);
      z = rb * sinq;
    }
    double e10 = e1 * cosq + e2 * xmu * z;
    double e20 = e1 * y * rcp(xmu) + e2 * cosq;
    double xnor = fmax(fabs(e10), fabs(e20));
    if (xnor < 1.0e-40) xnor = 1.0;
    double r = rcp(xnor);
    e1 = e10 * r;
    e2 = e20 * r;
  }
  return e1;
}

// Rayleigh (Dunkin compound matrices), half-space upward
JDSG_DEV double dltar4(const float4 *__restrict__ st, int stride, int mmax, int llw, double wvno,
                       double omga) {
  double omega = omga < 1.0e-4 ? 1.0e-4 : omga;
  double iom = rcp(omega);
  double wvno2 = wvno * wvno;
  float4 h = st[(size_t)(mmax - 1) * stride];
  double xka = omega * rcp((double)h.y);
  double xkb = omega * rcp((double)h.z);
  double ra = sqrt((wvno + xka) * fabs(wvno - xka));
  double rb = sqrt((wvno + xkb) * fabs(wvno - xkb));
  double t = (double)h.z * iom;
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
    if (m > 0) Lnext = st[(size_t)(m - 1) * stride];
    double al = (double)L.y, be = (double)L.z;
    xka = omega * rcp(al);
    xkb = omega * rcp(be);
    t = be * iom;
    gammk = 2.0 * t * t;
    gam = gammk * wvno2;
    ra = sqrt((wvno + xka) * fabs(wvno - xka));
    rb = sqrt((wvno + xkb) * fabs(wvno - xkb));
    double dpth = (double)L.x;
    rho1 = (double)L.w;
    double irho = rcp(rho1);
    double cosp, w, x, ep, cosq, y, z, eq;
    var_ps(ra * dpth, ra, rcp(ra), wvno, xka, dpth, cosp, w, x, ep);
    var_ps(rb * dpth, rb, rcp(rb), wvno, xkb, dpth, cosq, y, z, eq);
    double a0 = ep * eq;
    double cpcq = cosp * cosq, cpy = cosp * y, cpz = cosp * z, cqw = cosq * w,
           cqx = cosq * x, xy = x * y, xz = x * z, wy = w * y, wz = w * z;
    double gamm1_ = gam - 1.0;
    double twgm1 = gam + gamm1_;
    double gmgmk = gam * gammk;
    double gmgm1 = gam * gamm1_;
    double gm1sq = gamm1_ * gamm1_;
    double rho2 = rho1 * rho1;
    double a0pq = a0 - cpcq;
    double c00 = cpcq - 2.0 * gmgm1 * a0pq - gmgmk * xz - wvno2 * gm1sq * wy;
    double c01 = (wvno2 * cpy - cqx) * irho;
    double c02 = -(twgm1 * a0pq + gammk * xz + wvno2 * gamm1_ * wy) * irho;
    double c03 = (cpz - wvno2 * cqw) * irho;
    double c04 = -(2.0 * wvno2 * a0pq + xz + wvno2 * wvno2 * wy) * (irho * irho);
    double c10 = (gmgmk * cpz - gm1sq * cqw) * rho1;
    double c11 = cpcq;
    double c12 = gammk * cpz - gamm1_ * cqw;
    double c13 = -wz;
    double c30 = (gm1sq * cpy - gmgmk * cqx) * rho1;
    double c31 = -xy;
    double c32 = gamm1_ * cpy - gammk * cqx;
    double c40 = -(2.0 * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * xz + gm1sq * gm1sq * wy) * rho2;
    double c42 = -(gammk * gamm1_ * twgm1 * a0pq + gam * gammk * gammk * xz +
                   gamm1_ * gm1sq * wy) * rho1;
    double tt = -2.0 * wvno2;
    double c22 = a0 + 2.0 * (cpcq - c00);
    // ca(1,*) ... ca(5,*) with the symmetric entries of dnka folded in
    double ee0 = e0 * c00 + e1 * c10 + e2 * (tt * c42) + e3 * c30 + e4 * c40;
    double ee1 = e0 * c01 + e1 * c11 + e2 * (tt * c32) + e3 * c31 + e4 * c30;
    double ee2 = e0 * c02 + e1 * c12 + e2 * c22 + e3 * c32 + e4 * c42;
    double ee3 = e0 * c03 + e1 * c13 + e2 * (tt * c12) + e3 * c11 + e4 * c10;
    double ee4 = e0 * c04 + e1 * c03 + e2 * (tt * c02) + e3 * c01 + e4 * c00;
    double t1 = fmax(fmax(fmax(fabs(ee0), fabs(ee1)), fmax(fabs(ee2), fabs(ee3))), fabs(ee4));
    if (t1 < 1.0e-40) t1 = 1.0;
    double r = rcp(t1);
    e0 = ee0 * r;
    e1 = ee1 * r;
    e2 = ee2 * r;
    e3 = ee3 * r;
    e4 = ee4 * r;
  }
  if (llw != 1) {
    float4 L = st[0];
    xka = omega * rcp((double)L.y);
    ra = sqrt((wvno + xka) * fabs(wvno - xka));
    double dpth = (double)L.x;
    rho1 = (double)L.w;
    double cosp, w, x, ep;
    var_ps(ra * dpth, ra, rcp(ra), wvno, xka, dpth, cosp, w, x, ep);
    return cosp * e0 - rho1 * w * e1;
  }
  return e0;
}

}  // namespace jdsg_pf
