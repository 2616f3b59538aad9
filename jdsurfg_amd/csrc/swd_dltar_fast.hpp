// swd_dltar_fast.hpp -- the period equations of surfdisp96.f:712-898 for the
// exact-root polish of the learned-bracket path (polish, continuation,
// exact-root chains and their scan jobs).  Same mathematics as
// swd_dltar.hpp, arranged for throughput instead of bit-faithful rounding:
// FMA contraction, reciprocals from the hardware estimate refined by two
// Newton steps (shared where one denominator serves several quotients), and
// the three exponentials of a layer reduced to two (exp(-2p) = exp(-p)^2,
// exp(-p-q) = exp(-p) exp(-q)).  The values differ from the faithful form in
// the last bits only; a root polished to 1e-10 relative does not move.
// Each layer step is the layer's coefficients times the sweep vector; the
// one-lane forms (dltar1, dltar4) and the wave forms (dltar1_wave,
// dltar4_wave: a wave evaluates one equation, its lanes forming the layers'
// coefficients in parallel) share those two halves and return the same bits.
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

// ---- one layer step, split into the layer's coefficients (which depend on
// the layer, the frequency and the phase velocity only) and their product
// with the sweep vector, so that a wave can form the coefficients of many
// layers at once (dltar1_wave, dltar4_wave) and still take the same steps.
// Both halves fix their roundings: contraction only inside one expression,
// and the vector products as explicit fma chains.

// Love: e10 = e1 cosq + e2 xmu z, e20 = e1 y / xmu + e2 cosq
struct Coef1 {
  double cosq, y, z, xmu, ixmu;
};
JDSG_DEV Coef1 layer_coef1(const float4 L, double wvno, double omega) {
#pragma clang fp contract(on)
  Coef1 K;
  const double beta1 = (double)L.z;
  const double rho1 = (double)L.w;
  K.xmu = rho1 * beta1 * beta1;
  K.ixmu = rcp(K.xmu);
  const double xkb = omega * rcp(beta1);
  const double rb = sqrt((wvno + xkb) * fabs(wvno - xkb));
  const double dm = (double)L.x;
  const double q = dm * rb;
  if (wvno < xkb) {
    double sinq, cosq;
    sincos(q, &sinq, &cosq);
    K.cosq = cosq;
    K.y = sinq * rcp(rb);
    K.z = -rb * sinq;
  } else if (wvno == xkb) {
    K.cosq = 1.0;
    K.y = dm;
    K.z = 0.0;
  } else {
    const double fac = (q < 16) ? exp(-2.0 * q) : 0.0;
    K.cosq = (1.0 + fac) * 0.5;
    const double sinq = (1.0 - fac) * 0.5;
    K.y = sinq * rcp(rb);
    K.z = rb * sinq;
  }
  return K;
}
JDSG_DEV void apply1(const Coef1 &K, double &e1, double &e2) {
  const double e10 = fma(e1, K.cosq, (e2 * K.xmu) * K.z);
  const double e20 = fma(e1 * K.y, K.ixmu, e2 * K.cosq);
  double xnor = fmax(fabs(e10), fabs(e20));
  if (xnor < 1.0e-40) xnor = 1.0;
  const double r = rcp(xnor);
  e1 = e10 * r;
  e2 = e20 * r;
}
JDSG_DEV void hs_init1(const float4 h, double wvno, double omega, double &e1, double &e2) {
#pragma clang fp contract(on)
  const double beta1 = (double)h.z;
  const double rho1 = (double)h.w;
  const double xkb = omega * rcp(beta1);
  const double rb = sqrt((wvno + xkb) * fabs(wvno - xkb));
  e1 = rho1 * rb;
  const double ib = rcp(beta1);
  e2 = ib * ib;
}

// Rayleigh: the Dunkin compound-matrix entries the step multiplies the
// 5-vector by (ca(1..5,*) with the symmetric entries of dnka folded in);
// t42 .. t02 are -2 wvno^2 times c42, c32, c12, c02
struct Coef4 {
  double c00, c01, c02, c03, c04, c10, c11, c12, c13, c22, c30, c31, c32, c40, c42;
  double t42, t32, t12, t02;
};
constexpr int kCoef4 = 19;
JDSG_DEV Coef4 layer_coef4(const float4 L, double omega, double iom, double wvno,
                           double wvno2) {
#pragma clang fp contract(on)
  Coef4 K;
  const double al = (double)L.y, be = (double)L.z;
  const double xka = omega * rcp(al);
  const double xkb = omega * rcp(be);
  const double t = be * iom;
  const double gammk = 2.0 * t * t;
  const double gam = gammk * wvno2;
  const double ra = sqrt((wvno + xka) * fabs(wvno - xka));
  const double rb = sqrt((wvno + xkb) * fabs(wvno - xkb));
  const double dpth = (double)L.x;
  const double rho1 = (double)L.w;
  const double irho = rcp(rho1);
  double cosp, w, x, ep, cosq, y, z, eq;
  var_ps(ra * dpth, ra, rcp(ra), wvno, xka, dpth, cosp, w, x, ep);
  var_ps(rb * dpth, rb, rcp(rb), wvno, xkb, dpth, cosq, y, z, eq);
  const double a0 = ep * eq;
  const double cpcq = cosp * cosq, cpy = cosp * y, cpz = cosp * z, cqw = cosq * w,
               cqx = cosq * x, xy = x * y, xz = x * z, wy = w * y, wz = w * z;
  const double gamm1_ = gam - 1.0;
  const double twgm1 = gam + gamm1_;
  const double gmgmk = gam * gammk;
  const double gmgm1 = gam * gamm1_;
  const double gm1sq = gamm1_ * gamm1_;
  const double rho2 = rho1 * rho1;
  const double a0pq = a0 - cpcq;
  K.c00 = cpcq - 2.0 * gmgm1 * a0pq - gmgmk * xz - wvno2 * gm1sq * wy;
  K.c01 = (wvno2 * cpy - cqx) * irho;
  K.c02 = -(twgm1 * a0pq + gammk * xz + wvno2 * gamm1_ * wy) * irho;
  K.c03 = (cpz - wvno2 * cqw) * irho;
  K.c04 = -(2.0 * wvno2 * a0pq + xz + wvno2 * wvno2 * wy) * (irho * irho);
  K.c10 = (gmgmk * cpz - gm1sq * cqw) * rho1;
  K.c11 = cpcq;
  K.c12 = gammk * cpz - gamm1_ * cqw;
  K.c13 = -wz;
  K.c30 = (gm1sq * cpy - gmgmk * cqx) * rho1;
  K.c31 = -xy;
  K.c32 = gamm1_ * cpy - gammk * cqx;
  K.c40 = -(2.0 * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * xz + gm1sq * gm1sq * wy) * rho2;
  K.c42 = -(gammk * gamm1_ * twgm1 * a0pq + gam * gammk * gammk * xz +
            gamm1_ * gm1sq * wy) * rho1;
  K.c22 = a0 + 2.0 * (cpcq - K.c00);
  const double tt = -2.0 * wvno2;
  K.t42 = tt * K.c42;
  K.t32 = tt * K.c32;
  K.t12 = tt * K.c12;
  K.t02 = tt * K.c02;
  return K;
}
JDSG_DEV void apply4(const Coef4 &K, double (&e)[5]) {
  const double ee0 =
      fma(e[4], K.c40, fma(e[3], K.c30, fma(e[2], K.t42, fma(e[1], K.c10, e[0] * K.c00))));
  const double ee1 =
      fma(e[4], K.c30, fma(e[3], K.c31, fma(e[2], K.t32, fma(e[1], K.c11, e[0] * K.c01))));
  const double ee2 =
      fma(e[4], K.c42, fma(e[3], K.c32, fma(e[2], K.c22, fma(e[1], K.c12, e[0] * K.c02))));
  const double ee3 =
      fma(e[4], K.c10, fma(e[3], K.c11, fma(e[2], K.t12, fma(e[1], K.c13, e[0] * K.c03))));
  const double ee4 =
      fma(e[4], K.c00, fma(e[3], K.c01, fma(e[2], K.t02, fma(e[1], K.c03, e[0] * K.c04))));
  double t1 = fmax(fmax(fmax(fabs(ee0), fabs(ee1)), fmax(fabs(ee2), fabs(ee3))), fabs(ee4));
  if (t1 < 1.0e-40) t1 = 1.0;
  const double r = rcp(t1);
  e[0] = ee0 * r;
  e[1] = ee1 * r;
  e[2] = ee2 * r;
  e[3] = ee3 * r;
  e[4] = ee4 * r;
}
JDSG_DEV void hs_init4(const float4 h, double omega, double iom, double wvno, double wvno2,
                       double (&e)[5]) {
#pragma clang fp contract(on)
  const double xka = omega * rcp((double)h.y);
  const double xkb = omega * rcp((double)h.z);
  const double ra = sqrt((wvno + xka) * fabs(wvno - xka));
  const double rb = sqrt((wvno + xkb) * fabs(wvno - xkb));
  const double t = (double)h.z * iom;
  const double gammk = 2.0 * t * t;
  const double gam = gammk * wvno2;
  const double gamm1 = gam - 1.0;
  const double rho1 = (double)h.w;
  e[0] = rho1 * rho1 * (gamm1 * gamm1 - gam * gammk * ra * rb);
  e[1] = -rho1 * ra;
  e[2] = rho1 * (gamm1 - gammk * ra * rb);
  e[3] = rho1 * rb;
  e[4] = wvno2 - ra * rb;
}
// the top layer of a model with water above it (llw != 1)
JDSG_DEV double top_fluid4(const float4 L, double omega, double wvno, const double (&e)[5]) {
#pragma clang fp contract(on)
  const double xka = omega * rcp((double)L.y);
  const double ra = sqrt((wvno + xka) * fabs(wvno - xka));
  const double dpth = (double)L.x;
  const double rho1 = (double)L.w;
  double cosp, w, x, ep;
  var_ps(ra * dpth, ra, rcp(ra), wvno, xka, dpth, cosp, w, x, ep);
  return cosp * e[0] - rho1 * w * e[1];
}

// Love (Haskell SH), half-space to surface
JDSG_DEV double dltar1(const float4 *__restrict__ st, int stride, int mmax, int llw, double wvno,
                       double omega) {
  double e1, e2;
  hs_init1(st[(size_t)(mmax - 1) * stride], wvno, omega, e1, e2);
  float4 Lnext = (mmax - 2 >= 0) ? st[(size_t)(mmax - 2) * stride] : float4{};
  for (int m = mmax - 2; m >= llw - 1; --m) {
    const float4 L = Lnext;
    if (m > 0) Lnext = st[(size_t)(m - 1) * stride];
    apply1(layer_coef1(L, wvno, omega), e1, e2);
  }
  return e1;
}

// Rayleigh (Dunkin compound matrices), half-space upward
JDSG_DEV double dltar4(const float4 *__restrict__ st, int stride, int mmax, int llw, double wvno,
                       double omga) {
  const double omega = omga < 1.0e-4 ? 1.0e-4 : omga;
  const double iom = rcp(omega);
  const double wvno2 = wvno * wvno;
  double e[5];
  hs_init4(st[(size_t)(mmax - 1) * stride], omega, iom, wvno, wvno2, e);
  float4 Lnext = (mmax - 2 >= 0) ? st[(size_t)(mmax - 2) * stride] : float4{};
  for (int m = mmax - 2; m >= llw - 1; --m) {
    const float4 L = Lnext;
    if (m > 0) Lnext = st[(size_t)(m - 1) * stride];
    apply4(layer_coef4(L, omega, iom, wvno, wvno2), e);
  }
  if (llw != 1) return top_fluid4(st[0], omega, wvno, e);
  return e[0];
}

// ---- the same values by the 64 lanes of a wave together, for one chain:
// lane l forms the coefficients of step b + l (layer mmax - 2 - b - l) into
// cf (LDS, 64 doubles per coefficient, lane-fastest), and then every lane
// takes the steps in order from there, so each gets the value the one-lane
// form returns.  Every lane of the wave calls with the same arguments (a
// workgroup of one wave); cf holds kCoef4 x 64 doubles.
JDSG_DEV double dltar4_wave(const float4 *__restrict__ st, int stride, int mmax, int llw,
                            double wvno, double omga, double *cf) {
  const int lane = threadIdx.x & 63;
  const double omega = omga < 1.0e-4 ? 1.0e-4 : omga;
  const double iom = rcp(omega);
  const double wvno2 = wvno * wvno;
  double e[5];
  hs_init4(st[(size_t)(mmax - 1) * stride], omega, iom, wvno, wvno2, e);
  const int nstep = mmax - llw;
  for (int b = 0; b < nstep; b += 64) {
    if (b + lane < nstep) {
      const Coef4 K = layer_coef4(st[(size_t)(mmax - 2 - b - lane) * stride], omega, iom, wvno,
                                  wvno2);
      const double v[kCoef4] = {K.c00, K.c01, K.c02, K.c03, K.c04, K.c10, K.c11,
                                K.c12, K.c13, K.c22, K.c30, K.c31, K.c32, K.c40,
                                K.c42, K.t42, K.t32, K.t12, K.t02};
#pragma unroll
      for (int q = 0; q < kCoef4; ++q) cf[q * 64 + lane] = v[q];
    }
    __syncthreads();
    const int n = nstep - b < 64 ? nstep - b : 64;
#pragma unroll 2
    for (int j = 0; j < n; ++j) {
      Coef4 K;
      K.c00 = cf[0 * 64 + j]; K.c01 = cf[1 * 64 + j]; K.c02 = cf[2 * 64 + j];
      K.c03 = cf[3 * 64 + j]; K.c04 = cf[4 * 64 + j]; K.c10 = cf[5 * 64 + j];
      K.c11 = cf[6 * 64 + j]; K.c12 = cf[7 * 64 + j]; K.c13 = cf[8 * 64 + j];
      K.c22 = cf[9 * 64 + j]; K.c30 = cf[10 * 64 + j]; K.c31 = cf[11 * 64 + j];
      K.c32 = cf[12 * 64 + j]; K.c40 = cf[13 * 64 + j]; K.c42 = cf[14 * 64 + j];
      K.t42 = cf[15 * 64 + j]; K.t32 = cf[16 * 64 + j]; K.t12 = cf[17 * 64 + j];
      K.t02 = cf[18 * 64 + j];
      apply4(K, e);
    }
    __syncthreads();
  }
  if (llw != 1) return top_fluid4(st[0], omega, wvno, e);
  return e[0];
}
constexpr int kCoef1 = 5;
JDSG_DEV double dltar1_wave(const float4 *__restrict__ st, int stride, int mmax, int llw,
                            double wvno, double omega, double *cf) {
  const int lane = threadIdx.x & 63;
  double e1, e2;
  hs_init1(st[(size_t)(mmax - 1) * stride], wvno, omega, e1, e2);
  const int nstep = mmax - llw;
  for (int b = 0; b < nstep; b += 64) {
    if (b + lane < nstep) {
      const Coef1 K = layer_coef1(st[(size_t)(mmax - 2 - b - lane) * stride], wvno, omega);
      cf[0 * 64 + lane] = K.cosq;
      cf[1 * 64 + lane] = K.y;
      cf[2 * 64 + lane] = K.z;
      cf[3 * 64 + lane] = K.xmu;
      cf[4 * 64 + lane] = K.ixmu;
    }
    __syncthreads();
    const int n = nstep - b < 64 ? nstep - b : 64;
#pragma unroll 2
    for (int j = 0; j < n; ++j) {
      Coef1 K;
      K.cosq = cf[0 * 64 + j];
      K.y = cf[1 * 64 + j];
      K.z = cf[2 * 64 + j];
      K.xmu = cf[3 * 64 + j];
      K.ixmu = cf[4 * 64 + j];
      apply1(K, e1, e2);
    }
    __syncthreads();
  }
  return e1;
}

}  // namespace jdsg_pf
