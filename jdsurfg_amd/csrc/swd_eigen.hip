// swd_eigen.hip -- eigenfunctions, energy integrals, group velocity and
// layer sensitivity kernels on CDNA4 (the GPU form of src/SWD/sregn96.f90
// and slegn96.f90, including the sregnpu/slegnpu group-kernel assembly).
//
// One lane owns one entry (column, wave, mode, period).  Rayleigh: the
// Dunkin up-sweep (sregn96.f90:401-489) is stored per layer in a lane-strided
// scratch area; the Haskell down-sweep (989-1059) then runs top-down fused
// with the eigenfunction assembly (svfunc :193), the energy integrals
// (energy :1061, intijr :1198 -- E/E^-1 and the layer functions are formed
// once per layer instead of once per integral) and the dc/dh terms (getdcdh
// :1430).  With group kernels a pass kernel runs each of the periods T,
// T(1+.005), T(1-.005) and a combine kernel assembles dU/dm with the
// reference's formula, which reads the dc/dm left by the T(1-.005) pass
// (sregn96.f90:1831-1834).
//
// Outputs per entry: spherical c and U, and per layer dc/d{a,b,rho,h} and
// dU/d{a,b,rho,h} (h: layer thickness, half-space 0), at
//   base(entry) + layer * lstride.
#include "swd_kernels.hpp"

#pragma clang fp contract(fast)

namespace {



constexpr double kPiF = (double)3.1415926535898f;  // REAL literal

// 1/x from the hardware reciprocal estimate refined by two Newton steps:
// within about an ulp of the IEEE quotient at a little over half its
// instructions (an IEEE FP64 division is about ten).  For the per-layer
// quotients of the sweeps, whose denominators are layer constants or
// vertical wavenumbers: finite and nonzero where the result is used.
JDSG_DEV double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

JDSG_DEV double powm5d(double x) {
  double x2 = x * x, x4 = x2 * x2, x5 = x4 * x;
  return 1.0 / x5;
}

// ------------------------------------------------------------ Rayleigh parts
struct Layer {
  double za, zb, zd, zrho;
  int iwat;
};

// Per-pass constants: frequency, wavenumber and the reciprocals the layer
// loops multiply by (a division costs about ten FP64 instructions).
struct Wk {
  double om, om2, k, k2;     // omega, omega^2, wavenumber, wavenumber^2
  double ik, ik2, ic, kw;    // 1/k, 1/k^2, k/omega = 1/c, k^2/omega^2
  double iom2;               // 1/omega^2
};
JDSG_DEV Wk make_wk(double omega, double wvno) {
  Wk K;
  K.om = omega;
  K.om2 = omega * omega;
  K.k = wvno;
  K.k2 = wvno * wvno;
  K.ik = 1.0 / wvno;
  K.ik2 = K.ik * K.ik;
  K.ic = wvno / omega;
  K.kw = K.ic * K.ic;
  K.iom2 = 1.0 / K.om2;
  return K;
}

// Vertical wavenumbers of a layer and their layer exponentials.  ra and rb
// are real (evanescent) or imaginary (propagating), so exp(-ra d) is either
// exp(-Re(ra) d) or cos(Im(ra) d) - i sin(Im(ra) d): one exp or one sincos
// per wave type.  varsv, the Dunkin and Haskell matrices and every
// exponential of the energy integrals (ffunc, gfunc, h1func, h2func) are
// formed from Ea = exp(-ra d) and Eb = exp(-rb d).
struct LayerWaves {
  cplx ra, rb;        // vertical wavenumbers
  cplx Ea, Eb;        // exp(-ra d), exp(-rb d) (Eb unset in a fluid layer)
  double pr, pi;      // ra d
  double qr, qi;      // rb d
  double ma, mb;      // exp(-Re(ra) d), exp(-Re(rb) d): 1 when imaginary
  double ia, ib;      // 1 / (the nonzero part of ra, rb)
};

// |z| of a real or imaginary number (one part is zero, so this is exact), and
// the squared modulus, for the reference's small-modulus tests without hypot.
JDSG_DEV double abs_ri(cplx z) { return fabs(z.re) + fabs(z.im); }
JDSG_DEV double norm2(cplx z) { return z.re * z.re + z.im * z.im; }

JDSG_DEV LayerWaves layer_waves(const Layer &L, double omega, double wvno2) {
  LayerWaves w;
  double xka = omega * frcp(L.za);
  double xkb = (L.zb > 0.0) ? omega * frcp(L.zb) : 0.0;
  w.ra = csqrt_real(wvno2 - xka * xka);
  w.rb = csqrt_real(wvno2 - xkb * xkb);
  w.ia = frcp((w.ra.im != 0.0) ? w.ra.im : w.ra.re);
  w.ib = frcp((w.rb.im != 0.0) ? w.rb.im : w.rb.re);
  w.pr = w.ra.re * L.zd;
  w.pi = w.ra.im * L.zd;
  w.qr = w.rb.re * L.zd;
  w.qi = w.rb.im * L.zd;
  // The branches form scalars and the fields are assigned once after them:
  // stores to different fields in the two branches are merged by the
  // compiler into stores through selected addresses, in private memory.
  double ea_re, ea_im, ma;
  if (w.pi != 0.0) {
    double s, c;
    sincos(w.pi, &s, &c);
    ea_re = c;
    ea_im = -s;
    ma = 1.0;
  } else {
    ma = (w.pr == 0.0) ? 1.0 : exp(-w.pr);
    ea_re = ma;
    ea_im = 0.0;
  }
  double eb_re = 0.0, eb_im = 0.0, mb = 1.0;
  if (L.iwat == 0) {
    if (w.qi != 0.0) {
      double s, c;
      sincos(w.qi, &s, &c);
      eb_re = c;
      eb_im = -s;
    } else {
      mb = (w.qr == 0.0) ? 1.0 : exp(-w.qr);
      eb_re = mb;
    }
  }
  w.Ea = cmk(ea_re, ea_im);
  w.ma = ma;
  w.Eb = cmk(eb_re, eb_im);
  w.mb = mb;
  return w;
}

// sregn96.f90:827-911 varsv from the layer exponentials: a propagating wave
// gives cos/sin directly, an evanescent one the exponentially scaled
// cosh/sinh (exp(-2 Re(ra) d) = Ea^2, dropped past 30 as in the reference).
JDSG_DEV void varsv_w(const LayerWaves &w, const Layer &L, double &cosp, double &cosq,
                      double &rsinp, double &rsinq, double &sinpr, double &sinqr, double &pex,
                      double &svex) {
  const double tiny = (double)1.0e-5f;
  pex = w.pr;
  svex = 0.0;
  if (w.pi != 0.0) {
    double c = w.Ea.re, s = -w.Ea.im;
    cosp = c;
    rsinp = -(w.ra.im * s);
    sinpr = (abs_ri(w.ra) < tiny) ? L.zd : s * w.ia;
  } else {
    double pfac = (w.pr == 0.0) ? 1.0 : (w.pr < 30.) ? w.ma * w.ma : 0.0;
    cosp = 0.5 + pfac * 0.5;
    double sinp = 0.5 - pfac * 0.5;
    rsinp = w.ra.re * sinp;
    sinpr = (fabs(w.pr) < tiny && abs_ri(w.ra) < tiny) ? L.zd : sinp * w.ia;
  }
  if (L.iwat == 1) {
    cosq = 1.0;
    rsinq = 0.0;
    sinqr = 0.0;
    return;
  }
  svex = w.qr;
  if (w.qi != 0.0) {
    double c = w.Eb.re, s = -w.Eb.im;
    cosq = c;
    rsinq = -(w.rb.im * s);
    sinqr = (abs_ri(w.rb) < tiny) ? L.zd : s * w.ib;
  } else {
    double svfac = (w.qr == 0.0) ? 1.0 : (w.qr < 30.) ? w.mb * w.mb : 0.0;
    cosq = 0.5 + svfac * 0.5;
    double sinq = 0.5 - svfac * 0.5;
    rsinq = w.rb.re * sinq;
    sinqr = (fabs(w.qr) < tiny && abs_ri(w.rb) < tiny) ? L.zd : sinq * w.ib;
  }
}

// Scale a sweep vector by a power of two so that its largest entry lies in
// [0.5, 1): the sweeps of sregn96.f90 (:475-485, :1040-1050) and
// slegn96.f90 (:420-440) divide by the largest entry (left alone below
// `floor`) and add its logarithm to the exponent sum; scaling by 2^-e
// instead is exact and adds e ln 2.
template <int K>
JDSG_DEV double normalize_pow2(double (&v)[K], double floor = 1.0e-40) {
  double t1 = 0.0;
#pragma unroll
  for (int i = 0; i < K; ++i)
    if (fabs(v[i]) > t1) t1 = fabs(v[i]);
  if (t1 < floor) return 0.0;
  int e;
  (void)frexp(t1, &e);
#pragma unroll
  for (int i = 0; i < K; ++i) v[i] = ldexp(v[i], -e);
  return (double)e * 0.69314718055994530942;
}

// sregn96.f90:491-647 dnka, applied: ee = cd * ca (the up-sweep's row vector
// times the layer's 5x5 Dunkin matrix, summed over j in increasing order).
// Each branch forms the matrix entries as scalars and the product itself, so
// no 25-element array is merged across the fluid/solid branches (which the
// compiler keeps in private memory, a scratch round trip per layer step).
JDSG_DEV void dnka_apply(const double (&cd)[5], double (&ee)[5], double cosp, double rsinp,
                         double sinpr, double cossv, double rsinsv, double sinsvr, float rho,
                         float b, int iwat, const LayerWaves &w, const Wk &K) {
  const double wvno = K.k, wvno2 = K.k2, om2 = K.om2;
  // exp(-ex) and exp(-exa) of the reference (ex = Re(ra) d, exa = ex +
  // Re(rb) d) are the layer magnitudes ma and ma*mb
  const double ex = w.pr, exa = w.pr + w.qr;
  if (iwat == 1) {
    // nonzero entries: 00 11 33 44 = cosp, 22 = dfac, 01 = 34, 10 = 43
    const double dfac = (ex > 35.0) ? 0.0 : w.ma;
    const double irom = frcp((double)rho * om2);
    const double c01 = -rsinp * irom;
    const double c10 = -(double)rho * sinpr * om2;
    ee[0] = cd[0] * cosp + cd[1] * c10;
    ee[1] = cd[0] * c01 + cd[1] * cosp;
    ee[2] = cd[2] * dfac;
    ee[3] = cd[3] * cosp + cd[4] * c10;
    ee[4] = cd[3] * c01 + cd[4] * cosp;
    return;
  }
  double a0 = (exa < 60.0) ? w.ma * w.mb : 0.0;
  double cpcq = cosp * cossv, cpy = cosp * sinsvr, cpz = cosp * rsinsv, cqw = cossv * sinpr,
         cqx = cossv * rsinp, xy = rsinp * sinsvr, xz = rsinp * rsinsv, wy = sinpr * sinsvr,
         wz = sinpr * rsinsv;
  float rho2 = rho * rho;
  float bb2 = 2.0f * b * b;
  double gam = (double)bb2 * K.kw;
  double gam2 = gam * gam;
  double gamm1 = gam - 1.;
  double gamm2 = gamm1 * gamm1;
  double cqww2 = cqw * wvno2;
  double cqxw2 = cqx * K.ik2;
  double gg1 = gam * gamm1;
  double a0c = 2.0 * (a0 - cpcq);
  double xz2 = xz * K.ik2;
  double gxz2 = gam * xz2;
  double g2xz2 = gam2 * xz2;
  double a0cgg1 = a0c * (gam + gamm1);
  double wy2 = wy * wvno2;
  double g2wy2 = gamm2 * wy2;
  double g1wy2 = gamm1 * wy2;
  double rom = (double)rho * om2;
  const double irom = frcp(rom);
  double temp = a0c * gg1 + g2xz2 + g2wy2;
  const double c22 = a0 + temp + temp;
  const double c00 = cpcq - temp;
  const double c01 = (-cqx + wvno2 * cpy) * irom;
  temp = 0.5 * a0cgg1 + gxz2 + g1wy2;
  const double c02 = wvno * temp * irom;
  const double c03 = (-cqww2 + cpz) * irom;
  temp = wvno2 * (a0c + wy2) + xz;
  const double c04 = -temp * frcp((double)rho2 * om2 * om2);
  const double c10 = (-gamm2 * cqw + gam2 * cpz * K.ik2) * (double)rho * om2;
  const double c11 = cpcq;
  const double c12 = (gamm1 * cqww2 - gam * cpz) * K.ik;
  const double c13 = -wz;
  const double c14 = c03;
  temp = 0.5 * a0cgg1 * gg1 + gam2 * gxz2 + gamm2 * g1wy2;
  const double c20 = -2.0 * temp * (double)rho * om2 * K.ik;
  const double c21 = -wvno * (gam * cqxw2 - gamm1 * cpy) * 2.0;
  const double c23 = -2.0 * c12;
  const double c24 = -2.0 * c02;
  const double c30 = (-gam2 * cqxw2 + gamm2 * cpy) * (double)rho * om2;
  const double c31 = -xy;
  const double c32 = -c21 / 2.0;
  const double c33 = c11;
  const double c34 = c01;
  temp = gamm2 * (a0c * gam2 + g2wy2) + gam2 * g2xz2;
  const double c40 = -(double)rho2 * om2 * om2 * temp * K.ik2;
  const double c41 = c30;
  const double c42 = -c20 / 2.0;
  const double c43 = c10;
  const double c44 = c00;
  ee[0] = cd[0] * c00 + cd[1] * c10 + cd[2] * c20 + cd[3] * c30 + cd[4] * c40;
  ee[1] = cd[0] * c01 + cd[1] * c11 + cd[2] * c21 + cd[3] * c31 + cd[4] * c41;
  ee[2] = cd[0] * c02 + cd[1] * c12 + cd[2] * c22 + cd[3] * c32 + cd[4] * c42;
  ee[3] = cd[0] * c03 + cd[1] * c13 + cd[2] * c23 + cd[3] * c33 + cd[4] * c43;
  ee[4] = cd[0] * c04 + cd[1] * c14 + cd[2] * c24 + cd[3] * c34 + cd[4] * c44;
}

// sregn96.f90:913-987 hska, applied: out = AA * vv (the down-sweep's column
// vector times the layer's 4x4 Haskell matrix, summed over j in increasing
// order), the entries formed as scalars in each branch (no array merged
// across the fluid/solid branches, which would live in private memory).
JDSG_DEV void hska_apply(const double (&vv)[4], double (&out)[4], double cosp, double rsinp,
                         double sinpr, double tcossv, double trsinsv, double tsinsvr, float rho,
                         float b, int iwat, const LayerWaves &w, const Wk &K) {
  const double wvno = K.k, wvno2 = K.k2, om2 = K.om2;
  const double pex = w.pr, svex = (iwat == 1) ? 0.0 : w.qr;
  if (iwat == 1) {
    // nonzero entries: 00 33 = dfac, 11 22 = cosp, 12, 21
    const double dfac = (pex > 35.0) ? 0.0 : w.ma;
    const double a12 = -rsinp * frcp((double)rho * om2);
    const double a21 = -(double)rho * om2 * sinpr;
    out[0] = dfac * vv[0];
    out[1] = cosp * vv[1] + a12 * vv[2];
    out[2] = a21 * vv[1] + cosp * vv[2];
    out[3] = dfac * vv[3];
    return;
  }
  // exp(svex - pex): with a propagating S wave (svex = 0) that is ma
  double dfac = ((pex - svex) > 70.0) ? 0.0 : (svex == 0.0) ? w.ma : exp(svex - pex);
  double cossv = dfac * tcossv;
  double rsinsv = dfac * trsinsv;
  double sinsvr = dfac * tsinsvr;
  float bb2 = 2.0f * b * b;
  double gam = (double)bb2 * K.kw;
  double gamm1 = gam - 1.0;
  double rom = (double)rho * om2;
  const double irom = frcp(rom);
  const double a00 = cossv + gam * (cosp - cossv);
  const double a01 = -wvno * gamm1 * sinpr + gam * rsinsv * K.ik;
  const double a02 = -wvno * (cosp - cossv) * irom;
  const double a03 = (wvno2 * sinpr - rsinsv) * irom;
  const double a10 = gam * rsinp * K.ik - wvno * gamm1 * sinsvr;
  const double a11 = cosp - gam * (cosp - cossv);
  const double a12 = (-rsinp + wvno2 * sinsvr) * irom;
  const double a13 = -a02;
  const double a20 = rom * gam * gamm1 * (cosp - cossv) * K.ik;
  const double a21 = rom * (-gamm1 * gamm1 * sinpr + gam * gam * rsinsv * K.ik2);
  const double a22 = a11;
  const double a23 = -a01;
  const double a30 = rom * (gam * gam * rsinp * K.ik2 - gamm1 * gamm1 * sinsvr);
  const double a31 = -a20;
  const double a32 = -a10;
  const double a33 = a00;
  out[0] = a00 * vv[0] + a01 * vv[1] + a02 * vv[2] + a03 * vv[3];
  out[1] = a10 * vv[0] + a11 * vv[1] + a12 * vv[2] + a13 * vv[3];
  out[2] = a20 * vv[0] + a21 * vv[1] + a22 * vv[2] + a23 * vv[3];
  out[3] = a30 * vv[0] + a31 * vv[1] + a32 * vv[2] + a33 * vv[3];
}

// sregn96.f90:649-825 evalg (jbdry = 0)
struct EMat {
  cplx e[4][4], ei[4][4], ra, rb;
};
JDSG_DEV void evalg0(const Layer &L, bool allfluid, double wvno, double om, double om2,
                     double wvno2, EMat &M, double *gbr_re) {
  double xka = om / L.za;
  double xkb = (L.zb > 0.0) ? om / L.zb : 0.0;
  cplx ra = csqrt_real(wvno2 - xka * xka);
  cplx rb = csqrt_real(wvno2 - xkb * xkb);
  M.ra = ra;
  M.rb = rb;
  double gam = L.zb * wvno / om;
  gam = 2.0 * (gam * gam);
  double gamm1 = gam - 1.0;
  double zr = L.zrho;
  if (L.iwat == 0) {
    M.e[0][0] = cmk(wvno, 0); M.e[0][1] = rb; M.e[0][2] = cmk(wvno, 0); M.e[0][3] = -rb;
    M.e[1][0] = ra; M.e[1][1] = cmk(wvno, 0); M.e[1][2] = -ra; M.e[1][3] = cmk(wvno, 0);
    M.e[2][0] = cmk(zr * om2 * gamm1, 0);
    M.e[2][1] = zr * om2 * gam * rb / wvno;
    M.e[2][2] = cmk(zr * om2 * gamm1, 0);
    M.e[2][3] = -zr * om2 * gam * rb / wvno;
    M.e[3][0] = zr * om2 * gam * ra / wvno;
    M.e[3][1] = cmk(zr * om2 * gamm1, 0);
    M.e[3][2] = -zr * om2 * gam * ra / wvno;
    M.e[3][3] = cmk(zr * om2 * gamm1, 0);
    M.ei[0][0] = cmk(0.5 * gam / wvno, 0);
    M.ei[0][1] = -0.5 * gamm1 / ra;
    M.ei[0][2] = cmk(-0.5 / (zr * om2), 0);
    M.ei[0][3] = 0.5 * wvno / (zr * om2 * ra);
    M.ei[1][0] = -0.5 * gamm1 / rb;
    M.ei[1][1] = cmk(0.5 * gam / wvno, 0);
    M.ei[1][2] = 0.5 * wvno / (zr * om2 * rb);
    M.ei[1][3] = cmk(-0.5 / (zr * om2), 0);
    M.ei[2][0] = cmk(0.5 * gam / wvno, 0);
    M.ei[2][1] = 0.5 * gamm1 / ra;
    M.ei[2][2] = cmk(-0.5 / (zr * om2), 0);
    M.ei[2][3] = -0.5 * wvno / (zr * om2 * ra);
    M.ei[3][0] = 0.5 * gamm1 / rb;
    M.ei[3][1] = cmk(0.5 * gam / wvno, 0);
    M.ei[3][2] = -0.5 * wvno / (zr * om2 * rb);
    M.ei[3][3] = cmk(-0.5 / (zr * om2), 0);
    if (gbr_re) {
      cplx g[5];
      g[0] = (zr * zr) * om2 * om2 * (-gam * gam * ra * rb + wvno2 * gamm1 * gamm1);
      g[1] = -zr * (wvno2 * ra) * om2;
      g[2] = -zr * (-gam * ra * rb + wvno2 * gamm1) * om2 * wvno;
      g[3] = zr * (wvno2 * rb) * om2;
      g[4] = wvno2 * (wvno2 - ra * rb);
      cplx den = -zr * zr * om2 * om2 * wvno2 * ra * rb;
#pragma unroll
      for (int i = 0; i < 5; ++i) gbr_re[i] = ((0.25 * g[i]) / den).re;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) M.e[i][j] = M.ei[i][j] = cmk(0, 0);
    M.e[0][0] = ra;
    M.e[0][1] = -ra;
    M.e[1][0] = cmk(-zr * om2, 0);
    M.e[1][1] = cmk(-zr * om2, 0);
    M.ei[0][0] = 0.5 / ra;
    M.ei[0][1] = cmk(-0.5 / (zr * om2), 0);
    M.ei[1][0] = -0.5 / ra;
    M.ei[1][1] = cmk(-0.5 / (zr * om2), 0);
    if (gbr_re) {
      if (allfluid) {
        gbr_re[0] = (0.5 / ra).re;
        gbr_re[1] = (cmk(0.5, 0.0) / cmk(-zr * om2, 0)).re;
        gbr_re[2] = gbr_re[3] = gbr_re[4] = 0.0;
      } else {
        gbr_re[0] = gbr_re[1] = gbr_re[2] = 0.0;
        gbr_re[3] = ((0.5 * zr * om2) / ra).re;
        gbr_re[4] = -0.5;
      }
    }
  }
}

// sregn96.f90:1319-1397
JDSG_DEV cplx ffunc(cplx nub, double dm) {
  if (cabs_(nub) < 1.0e-08) return cmk(dm, 0);
  cplx argcd = nub * dm;
  cplx exqq = (argcd.re < 40.0) ? cexp_(-2.0 * argcd) : cmk(0, 0);
  return (1.0 - exqq) / (2.0 * nub);
}
JDSG_DEV cplx gfunc(cplx nub, double dm) {
  cplx argcd = nub * dm;
  if (argcd.re < 75) return cexp_(-argcd) * dm;
  return cmk(0, 0);
}
struct U4 {
  double ur, uz, tz, tr;
};

// The layer matrices E and E^-1 of evalg (sregn96.f90:649-825, jbdry = 0)
// have entries that are real multiples of 1, ra, rb, 1/ra, 1/rb, so the
// energy integrals take them from six scalars as needed instead of holding
// two complex 4x4 matrices live (which would cost 128 registers per lane).
struct EScal {
  cplx ra, rb, ira, irb;  // ra, rb and their reciprocals
  double k, z, g, g1;     // wavenumber, rho*om^2, gamma, gamma-1
  double gk, izh;         // z*gamma/k and 0.5/z
};

JDSG_DEV EScal escal(const Layer &L, double wvno, double om, double om2, double wvno2) {
  EScal s;
  double xka = om / L.za;
  double xkb = (L.zb > 0.0) ? om / L.zb : 0.0;
  s.ra = csqrt_real(wvno2 - xka * xka);
  s.rb = csqrt_real(wvno2 - xkb * xkb);
  s.ira = 1.0 / s.ra;
  s.irb = (L.iwat == 0) ? 1.0 / s.rb : cmk(0, 0);
  double gam = L.zb * wvno / om;
  s.g = 2.0 * (gam * gam);
  s.g1 = s.g - 1.0;
  s.k = wvno;
  s.z = L.zrho * om2;
  s.gk = s.z * s.g / wvno;
  s.izh = 0.5 / s.z;
  return s;
}

// solid layer entries E(I,J), E^-1(I,J)
template <int I, int J>
JDSG_DEV cplx Es(const EScal &s) {
  if (I == 0) return (J == 0 || J == 2) ? cmk(s.k, 0) : (J == 1 ? s.rb : -s.rb);
  if (I == 1) return (J == 1 || J == 3) ? cmk(s.k, 0) : (J == 0 ? s.ra : -s.ra);
  if (I == 2)
    return (J == 0 || J == 2) ? cmk(s.z * s.g1, 0) : (J == 1 ? s.gk * s.rb : -(s.gk * s.rb));
  return (J == 1 || J == 3) ? cmk(s.z * s.g1, 0) : (J == 0 ? s.gk * s.ra : -(s.gk * s.ra));
}
template <int I, int J>
JDSG_DEV cplx Eis(const EScal &s) {
  const double hg = 0.5 * s.g / s.k, hk = s.k * s.izh;
  if (I == 0)
    return J == 0 ? cmk(hg, 0) : J == 1 ? (-0.5 * s.g1) * s.ira
           : J == 2 ? cmk(-s.izh, 0) : hk * s.ira;
  if (I == 1)
    return J == 0 ? (-0.5 * s.g1) * s.irb : J == 1 ? cmk(hg, 0)
           : J == 2 ? hk * s.irb : cmk(-s.izh, 0);
  if (I == 2)
    return J == 0 ? cmk(hg, 0) : J == 1 ? (0.5 * s.g1) * s.ira
           : J == 2 ? cmk(-s.izh, 0) : -(hk * s.ira);
  return J == 0 ? (0.5 * s.g1) * s.irb : J == 1 ? cmk(hg, 0)
         : J == 2 ? -(hk * s.irb) : cmk(-s.izh, 0);
}
// fluid layer entries (rows and columns 0-1 only)
template <int I, int J>
JDSG_DEV cplx Ef(const EScal &s) {
  if (I == 0) return J == 0 ? s.ra : -s.ra;
  return cmk(-s.z, 0);
}
template <int I, int J>
JDSG_DEV cplx Eif(const EScal &s) {
  if (J == 1) return cmk(-s.izh, 0);
  return I == 0 ? 0.5 * s.ira : -0.5 * s.ira;
}

// lower half-space (typelyr 1)
template <int I, int J>
JDSG_DEV double int_half(const EScal &s, cplx km1pd, cplx km1sd) {
  cplx i2 = Es<I, 2>(s), i3 = Es<I, 3>(s), j2 = Es<J, 2>(s), j3 = Es<J, 3>(s);
  cplx c = i2 * j2 * km1pd * km1pd / (2.0 * s.ra) +
           (i2 * j3 + i3 * j2) * km1pd * km1sd / (s.ra + s.rb) +
           i3 * j3 * km1sd * km1sd / (2.0 * s.rb);
  return c.re;
}
template <int I, int J>
JDSG_DEV double int_fluid(const EScal &s, int typelyr, cplx kmpu, cplx km1pd, cplx FA,
                          cplx GA) {
  cplx i0 = Ef<I, 0>(s), i1 = Ef<I, 1>(s), j0 = Ef<J, 0>(s), j1 = Ef<J, 1>(s);
  cplx c;
  if (typelyr == 0)
    c = i0 * j0 * kmpu * kmpu * FA + (i0 * j1 + i1 * j0) * kmpu * km1pd * GA +
        i1 * j1 * km1pd * km1pd * FA;
  else
    c = i1 * j1 * km1pd * km1pd / (2.0 * s.ra);
  return c.re;
}

// The six integrals of an internal solid layer (sregn96.f90:1198-1317,
// intijr typelyr 0) from the shared layer exponentials.  Rows 1 and 3 of E
// are k p + rb q and z(g-1) p + (z g/k) rb q, rows 2 and 4 are ra r + k s
// and (z g/k) ra r + z(g-1) s, with p = (1,0,1,0), q = (0,1,0,-1),
// r = (1,0,-1,0), s = (0,1,0,1).  So each INT_IJ = Re(E_I M E_J) -- M the
// symmetric matrix of the ten layer-function-weighted amplitude products --
// is a real combination of the six forms pMp, qMq, pMq, rMr, sMs, rMs, and
// those are sums of four amplitude products.  ra^2 and rb^2 are real.
JDSG_DEV void solid_ints(const Layer &L, const LayerWaves &w, const U4 &u0, const U4 &u1,
                         const Wk &K, double irz, double INT[6]) {
  const double k = K.k, z = L.zrho * K.om2, d = L.zd;
  const double gm = L.zb * K.ic;
  const double g = 2.0 * (gm * gm), g1 = g - 1.0;
  const double gk = z * g * K.ik, izh = 0.5 * irz;
  const double hg = 0.5 * g * K.ik, hk = k * izh;
  // 1/ra, 1/rb of a real or imaginary number
  const cplx ira = (w.ra.im == 0.0) ? cmk(w.ia, 0.0) : cmk(0.0, -w.ia);
  const cplx irb = (w.rb.im == 0.0) ? cmk(w.ib, 0.0) : cmk(0.0, -w.ib);
  // wave amplitudes E^-1 u at the layer bottom (up-going) and top (down-going)
  const cplx kmpu = (hg * u1.ur - izh * u1.tz) + (hk * u1.tr - 0.5 * g1 * u1.uz) * ira;
  const cplx kmsu = (hg * u1.uz - izh * u1.tr) + (hk * u1.tz - 0.5 * g1 * u1.ur) * irb;
  const cplx km1pd = (hg * u0.ur - izh * u0.tz) + (0.5 * g1 * u0.uz - hk * u0.tr) * ira;
  const cplx km1sd = (hg * u0.uz - izh * u0.tr) + (0.5 * g1 * u0.ur - hk * u0.tz) * irb;
  // layer functions (sregn96.f90:1319-1397) with their reference cut-offs
  const cplx zero = cmk(0.0, 0.0);
  const cplx FA = (abs_ri(w.ra) < 1.0e-08) ? cmk(d, 0.0)
                  : (0.5 * (1.0 - ((w.pr < 40.0) ? w.Ea * w.Ea : zero))) * ira;
  const cplx GA = (w.pr < 75.0) ? w.Ea * d : zero;
  const cplx FB = (abs_ri(w.rb) < 1.0e-08) ? cmk(d, 0.0)
                  : (0.5 * (1.0 - ((w.qr < 40.0) ? w.Eb * w.Eb : zero))) * irb;
  const cplx GB = (w.qr < 75.0) ? w.Eb * d : zero;
  const cplx sab = w.rb + w.ra, dab = w.ra - w.rb;
  // complex quotients as a conj(b) / |b|^2 (|ra|, |rb| are of order
  // omega / velocity, far from overflow): one division each; the reference's
  // |b| < 1e-8 tests read |b|^2 < 1e-16
  const double n1 = norm2(sab), n2 = norm2(dab);
  const cplx H1 = (n1 < 1.0e-16) ? cmk(d, 0.0)
                  : ((1.0 - ((sab.re * d < 40.0) ? w.Ea * w.Eb : zero)) * cconj(sab)) *
                        frcp(n1);
  const cplx H2 = (n2 < 1.0e-16) ? cmk(d, 0.0)
                  : ((((w.qr < 40.0) ? w.Eb : zero) - ((w.pr < 40.0) ? w.Ea : zero)) *
                     cconj(dab)) * frcp(n2);
  // the six quadratic forms of M
  const cplx A2 = kmpu * kmpu + km1pd * km1pd, X = kmpu * km1pd;
  const cplx B2 = kmsu * kmsu + km1sd * km1sd, Y = kmsu * km1sd;
  const double fa = (FA * A2).re, ga = 2.0 * (GA * X).re;
  const double fb = (FB * B2).re, gb = 2.0 * (GB * Y).re;
  const double P = fa + ga, R = fa - ga;      // pMp, rMr
  const double Q = fb - gb, S = fb + gb;      // qMq, sMs
  const cplx h1t = H1 * (kmpu * kmsu - km1pd * km1sd);
  const cplx h2t = H2 * (km1pd * kmsu - kmpu * km1sd);
  const cplx PQ = h1t + h2t, RS = h1t - h2t;  // pMq, rMs
  const double rbPQ = w.rb.re * PQ.re - w.rb.im * PQ.im;  // Re(rb pMq)
  const double raRS = w.ra.re * RS.re - w.ra.im * RS.im;  // Re(ra rMs)
  const double rb2 = w.rb.re * w.rb.re - w.rb.im * w.rb.im;
  const double ra2 = w.ra.re * w.ra.re - w.ra.im * w.ra.im;
  const double zg1 = z * g1;
  INT[0] = k * k * P + 2.0 * k * rbPQ + rb2 * Q;                       // INT11
  INT[1] = k * zg1 * P + (k * gk + zg1) * rbPQ + gk * rb2 * Q;         // INT13
  INT[2] = ra2 * R + 2.0 * k * raRS + k * k * S;                       // INT22
  INT[3] = gk * ra2 * R + (zg1 + k * gk) * raRS + k * zg1 * S;         // INT24
  INT[4] = zg1 * zg1 * P + 2.0 * zg1 * gk * rbPQ + gk * gk * rb2 * Q;  // INT33
  INT[5] = gk * gk * ra2 * R + 2.0 * gk * zg1 * raRS + zg1 * zg1 * S;  // INT44
}

struct EnergySums {
  double s0, s1, s2, s3;
};

// energy contribution of layer m (sregn96.f90:1061-1161 loop body).  Writes
// the unscaled dc/da, dc/db, dc/drho of layer m to ka/kb/kr.
JDSG_DEV void layer_energy(const Layer &L, const LayerWaves *lw, int typelyr, const U4 &u0,
                           const U4 &u1, const Wk &K, double c, EnergySums &S, double &ka,
                           double &kb, double &kr, double &ur_fluid) {
  const double om = K.om, om2 = K.om2, wvno = K.k, wvno2 = K.k2;
  double ah = L.za, av = L.za, bv = L.zb, rho = L.zrho, eta = 1.0;
  double TA = L.zrho * L.za * L.za, TC = TA;
  if (L.iwat == 1) {
    const EScal s = escal(L, wvno, om, om2, wvno2);
    double TN = 0.0;
    double TF = TA - 2. * TN;
    double a12 = -(wvno * wvno - om * om / (ah * ah)) / (rho * om * om);
    cplx km1pd = Eif<1, 0>(s) * u0.uz + Eif<1, 1>(s) * u0.tz;
    cplx kmpu = cmk(0, 0), FA = cmk(0, 0), GA = cmk(0, 0);
    if (typelyr == 0) {
      kmpu = Eif<0, 0>(s) * u1.uz + Eif<0, 1>(s) * u1.tz;
      FA = ffunc(s.ra, L.zd);
      GA = gfunc(s.ra, L.zd);
    }
    double INT11 = int_fluid<0, 0>(s, typelyr, kmpu, km1pd, FA, GA);
    double INT22 = int_fluid<1, 1>(s, typelyr, kmpu, km1pd, FA, GA);
    double wr = wvno / (rho * om2);
    double URUR = INT22 * (wr * wr);
    double UZUZ = INT11;
    double URDUZ = -(wvno / (rho * om2)) * a12 * INT22;
    double DUZDUZ = a12 * a12 * INT22;
    S.s0 = S.s0 + rho * (URUR + UZUZ);
    S.s1 = S.s1 + TA * URUR;
    S.s2 = S.s2 - TF * URDUZ;
    S.s3 = S.s3 + TC * DUZDUZ;
    double facah = rho * ah * (URUR - 2. * eta * URDUZ / wvno);
    double facav = rho * av * DUZDUZ / wvno2;
    ka = facah + facav;
    kb = 0.0;
    double facr = -0.5 * c * c * (URUR + UZUZ);
    kr = 0.5 * (av * facav + ah * facah) / rho + facr;
    ur_fluid = -wvno * u0.tz / (rho * om2);
    return;
  }
  double TL = L.zrho * L.zb * L.zb, TN = TL;
  double TF = TA - 2. * TN;
  double a12 = -wvno, a14 = frcp(TL), a23 = frcp(TC), a21 = wvno * TF * a23;
  const double irho = a23 * (L.za * L.za);
  double INT11, INT13, INT22, INT24, INT33, INT44;
  if (typelyr == 0) {
    double I6[6];
    solid_ints(L, *lw, u0, u1, K, irho * K.iom2, I6);
    INT11 = I6[0];
    INT13 = I6[1];
    INT22 = I6[2];
    INT24 = I6[3];
    INT33 = I6[4];
    INT44 = I6[5];
  } else {
    const EScal s = escal(L, wvno, om, om2, wvno2);
    cplx km1pd = Eis<2, 0>(s) * u0.ur + Eis<2, 1>(s) * u0.uz + Eis<2, 2>(s) * u0.tz +
                 Eis<2, 3>(s) * u0.tr;
    cplx km1sd = Eis<3, 0>(s) * u0.ur + Eis<3, 1>(s) * u0.uz + Eis<3, 2>(s) * u0.tz +
                 Eis<3, 3>(s) * u0.tr;
    INT11 = int_half<0, 0>(s, km1pd, km1sd);
    INT13 = int_half<0, 2>(s, km1pd, km1sd);
    INT22 = int_half<1, 1>(s, km1pd, km1sd);
    INT24 = int_half<1, 3>(s, km1pd, km1sd);
    INT33 = int_half<2, 2>(s, km1pd, km1sd);
    INT44 = int_half<3, 3>(s, km1pd, km1sd);
  }
  double URUR = INT11;
  double UZUZ = INT22;
  double DURDUR = a12 * a12 * INT22 + 2. * a12 * a14 * INT24 + a14 * a14 * INT44;
  double DUZDUZ = a21 * a21 * INT11 + 2. * a21 * a23 * INT13 + a23 * a23 * INT33;
  double URDUZ = a21 * INT11 + a23 * INT13;
  double UZDUR = a12 * INT22 + a14 * INT24;
  S.s0 = S.s0 + rho * (URUR + UZUZ);
  S.s1 = S.s1 + TL * UZUZ + TA * URUR;
  S.s2 = S.s2 + TL * UZDUR - TF * URDUZ;
  S.s3 = S.s3 + TL * DURDUR + TC * DUZDUZ;
  double facah = rho * ah * (URUR - 2. * eta * URDUZ * K.ik);
  double facav = rho * av * DUZDUZ * K.ik2;
  double facbh = 0.0;
  double facbv = rho * bv * (UZUZ + 2. * UZDUR * K.ik + DURDUR * K.ik2 + 4. * eta * URDUZ * K.ik);
  ka = facah + facav;
  kb = facbv + facbh;
  double facr = -0.5 * c * c * (URUR + UZUZ);
  kr = 0.5 * (av * facav + ah * facah + bv * facbv) * irho + facr;
  ur_fluid = u0.ur;
}

// sregn96.f90:1430-1529 getdcdh, sum of the six terms (before fac)
// ixl2, ixmu (and ixl2p, ixmup of the layer above): 1/(lambda + 2 mu) and
// 1/mu, formed once per layer and reused for the next interface.
JDSG_DEV double dcdh_terms(int m, const Layer &Lm, const Layer &Lp, double xmu, double xlam,
                           double xmup, double xlamp, double ixl2, double ixmu, double ixl2p,
                           double ixmup, const U4 &u, double om2, double wvno, double wvno2) {
  double tuz = u.uz, ttz = u.tz, ttr = u.tr, tur;
  if (Lm.iwat == 1)
    tur = -wvno * ttz / (Lm.zrho * om2);
  else
    tur = u.ur;
  double gfac1, gfac2, gfac3, gfac4, gfac5, gfac6;
  if (m == 0) {
    double drho = Lm.zrho - 0.0;
    double dmu = xmu - 0.0;
    double dlm = xlam - 0.0;
    double dl2mu = dlm + dmu + dmu;
    double xl2mp = xlam + xmu + xmu;
    double duzdzp = (ttz + wvno * xlam * tur) * ixl2;
    double durdzp = (Lm.iwat == 1) ? wvno * tuz : (ttr * ixmu) - wvno * tuz;
    double drur2 = tur * tur * drho;
    double dlur2 = tur * tur * dl2mu;
    gfac1 = om2 * drho * (tuz * tuz);
    gfac2 = om2 * drur2;
    gfac3 = -wvno2 * dmu * (tuz * tuz);
    gfac4 = -wvno2 * dlur2;
    gfac5 = (xl2mp * (duzdzp * duzdzp));
    gfac6 = (xmu * (durdzp * durdzp));
  } else {
    double drho = Lm.zrho - Lp.zrho;
    double dmu = xmu - xmup;
    double dlm = xlam - xlamp;
    double dl2mu = dlm + dmu + dmu;
    double xl2mp = xlam + xmu + xmu;
    double xl2mm = xlamp + xmup + xmup;
    double duzdzp = (ttz + wvno * xlam * tur) * ixl2;
    double durdzp = (xmu == 0.0) ? wvno * tuz : (ttr * ixmu) - wvno * tuz;
    double durdzm = (xmup == 0.0) ? wvno * tuz : (ttr * ixmup) - wvno * tuz;
    double drur2, dlur2, duzdzm;
    if (Lp.iwat == 1 && Lm.iwat == 0) {
      double URB = -wvno * u.tz / (Lp.zrho * om2);
      drur2 = tur * tur * Lm.zrho - URB * URB * Lp.zrho;
      dlur2 = tur * tur * xl2mp - URB * URB * xl2mm;
      duzdzm = (ttz + wvno * xlamp * URB) / xlamp;
    } else if (Lp.iwat == 1 && Lm.iwat == 1) {
      double URB = -wvno * u.tz / (Lp.zrho * om2);
      drur2 = tur * tur * Lm.zrho - URB * URB * Lp.zrho;
      dlur2 = tur * tur * xl2mp - URB * URB * xl2mm;
      duzdzm = (ttz + wvno * xlamp * URB) / xl2mm;
    } else {
      drur2 = tur * tur * drho;
      dlur2 = tur * tur * dl2mu;
      duzdzm = (ttz + wvno * xlamp * tur) * ixl2p;
    }
    gfac1 = om2 * drho * (tuz * tuz);
    gfac2 = om2 * drur2;
    gfac3 = -wvno2 * dmu * (tuz * tuz);
    gfac4 = -wvno2 * dlur2;
    gfac5 = (xl2mp * (duzdzp * duzdzp) - xl2mm * (duzdzm * duzdzm));
    gfac6 = (xmu * (durdzp * durdzp) - xmup * (durdzm * durdzm));
  }
  return gfac1 + gfac2 + gfac3 + gfac4 + gfac5 + gfac6;
}

// Item decoding: item -> (col, mode, k) of this wave
JDSG_DEV void decode_item(const EigenArgs &A, long long it, int &col, int &mode, int &k,
                          long long &e) {
  int ncol = A.cm.ncol;
  if (A.item_ent == nullptr) {
    col = (int)(it % ncol);
    long long r = it / ncol;
    k = (int)(r % A.grid_K);
    mode = (int)(r / A.grid_K);
    int woff = (A.wave == 0) ? 0 : A.grid_nR;
    e = A.ent_off[col] + (long long)(woff + mode) * A.grid_K + k;
    return;
  }
  e = A.item_ent[it];
  // column by binary search in ent_off
  int lo = 0, hi = ncol - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (A.ent_off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  col = lo;
  int K = A.per_off[col + 1] - A.per_off[col];
  long long el = e - A.ent_off[col];
  int wm = (int)(el / K);
  k = (int)(el % K);
  mode = (A.wave == 0) ? wm : wm - A.nmode[2 * col];
}

// Rayleigh layer i of a column: the flattened velocities and thickness and
// the density with its Rayleigh flattening factor (1 for a flat model, so
// the product is the density itself).  Five loads with no branch between
// them, so they are in flight together.
JDSG_DEV Layer load_layer(const EigenArgs &A, int col, int i) {
  size_t q = (size_t)i * A.cm.ncol_pad + col;
  Layer L;
  L.za = A.cm.za[q];
  L.zb = A.cm.zb[q];
  L.zd = A.cm.zd[q];
  L.zrho = A.cm.rho0[q] * A.cm.rtpR[q];
  L.iwat = L.zb > 0.0 ? 0 : 1;
  return L;
}

// One Rayleigh pass (svfunc + energy).  Raw per-layer kernels go to
// dst[a,b,r,h] at dst_base + layer*dst_stride (h holds the getdcdh sum before
// fac).  Returns ugr, sumi0 and fac through the references.
struct Dst {
  double *p[4];
  long long base, stride;
};

// The Dunkin up-sweep's vector (5 entries and the exponent sum) at the top
// of every layer goes to the lane's global scratch (lane-strided, so each
// store of a wave is one 512-byte line); the down-sweep reads interface i + 1
// while it works on interface i.  Measured against keeping every sixth
// vector and recomputing the segments into LDS: 120 ms against 136 ms for
// the Rayleigh stage of 256x256 columns x 60 periods (the recomputation
// cost a second Dunkin sweep; the stored vectors add about 6 GB of HBM
// traffic per launch, well inside the bandwidth left beside the sweeps'
// arithmetic).

// One step of the up-sweep (sregn96.f90:455-487) through layer L: cd, the
// vector at the top of the layer below, becomes the vector at the top of L.
JDSG_DEV void dunkin_step(const Layer &L, double omega, double wvno2, const Wk &K,
                          double (&cd)[5], double &exsum) {
  const LayerWaves lw = layer_waves(L, omega, wvno2);
  double cosp, cossv, rsinp, rsinsv, sinpr, sinsvr, pex, svex;
  varsv_w(lw, L, cosp, cossv, rsinp, rsinsv, sinpr, sinsvr, pex, svex);
  double ee[5];
  dnka_apply(cd, ee, cosp, rsinp, sinpr, cossv, rsinsv, sinsvr, (float)L.zrho, (float)L.zb,
             L.iwat, lw, K);
  double exn = normalize_pow2(ee);
#pragma unroll
  for (int i = 0; i < 5; ++i) cd[i] = ee[i];
  exsum = exsum + pex + svex + exn;
}

JDSG_DEV void rayleigh_pass(const EigenArgs &A, int col, int N, double omega, double wvno,
                            double *scr, long long sstride, const Dst &D, double &ugr,
                            double &sumi0, double &fac, double *eig, long long ebase) {
  double om2 = omega * omega, wvno2 = wvno * wvno;
  const Wk K = make_wk(omega, wvno);
  bool allfluid = true;
  int jwat = 0;
  bool topwater = true;
  for (int i = 0; i < N; ++i) {
    double zb = A.cm.zb[(size_t)i * A.cm.ncol_pad + col];
    if (zb > 0.0) allfluid = false;
    if (topwater) {
      if (zb > 0.0) topwater = false; else jwat = i + 1;
    }
  }
  if (allfluid) jwat = 0;
  auto ck = [&](int m, int q) -> double & {
    return scr[((long long)m * 6 + q) * sstride];
  };
  // ---- up: Dunkin vectors from the half-space (sregn96.f90:401-489)
  {
    Layer H = load_layer(A, col, N - 1);
    EMat M;
    double g[5];
    evalg0(H, allfluid, wvno, omega, om2, wvno2, M, g);
    double cd[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      cd[j] = g[j];
      ck(N - 1, j) = cd[j];
    }
    ck(N - 1, 5) = 0.0;
    double exsum = 0.0;
    // layer m - 1 is loaded while layer m is swept (its loads are in flight
    // during the step instead of stalling the next one)
    Layer Lc = load_layer(A, col, max(N - 2, 0));
#pragma unroll 4
    for (int m = N - 2; m >= 0; --m) {
      const Layer Ln = load_layer(A, col, max(m - 1, 0));
      dunkin_step(Lc, omega, wvno2, K, cd, exsum);
      Lc = Ln;
#pragma unroll
      for (int j = 0; j < 5; ++j) ck(m, j) = cd[j];
      ck(m, 5) = exsum;
    }
  }
  // ---- down + svfunc + energy + getdcdh, top to bottom
  double cd0_1 = ck(0, 1);
  double cd0_2 = ck(0, 2);
  double exe0 = ck(0, 5);
  double f1213 = -cd0_1;
  const double r1213 = 1.0 / f1213;
  U4 uprev;
  uprev.ur = cd0_2 / cd0_1;
  uprev.uz = 1.0;
  uprev.tz = 0.0;
  uprev.tr = 0.0;
  if (0 < jwat) { uprev.ur = 0.0; uprev.tr = 0.0; }
  double vv[4] = {1.0, 0.0, 0.0, 0.0};
  double exsum = 0.0;
  double c = omega / wvno;
  EnergySums S = {0.0, 0.0, 0.0, 0.0};
  Layer Lprev = load_layer(A, col, 0);
  double xmu_prev = Lprev.zrho * (Lprev.zb * Lprev.zb);
  double xlam_prev = Lprev.zrho * (Lprev.za * Lprev.za) - 2 * xmu_prev;
  double ixmu_prev = 1.0 / xmu_prev, ixl2_prev = 1.0 / (xlam_prev + xmu_prev + xmu_prev);
  // getdcdh for layer 0
  {
    if (D.p[3]) {
      double g = dcdh_terms(0, Lprev, Lprev, xmu_prev, xlam_prev, 0.0, 0.0, ixl2_prev, ixmu_prev,
                            0.0, 0.0, uprev, om2, wvno, wvno2);
      D.p[3][D.base] = g;
    }
  }
  Layer Lm1 = Lprev;  // layer m-1 props for getdcdh
  double nx[6];       // the stored up-sweep vector of the next interface
#pragma unroll
  for (int j = 0; j < 6; ++j) nx[j] = ck(min(1, N - 1), j);
  double xmu_m1 = xmu_prev, xlam_m1 = xlam_prev;
  for (int m = 0; m < N - 1; ++m) {
    Layer L = Lprev;  // layer m
    const LayerWaves lw = layer_waves(L, omega, wvno2);
    double cosp, cossv, rsinp, rsinsv, sinpr, sinsvr, pex, svex;
    varsv_w(lw, L, cosp, cossv, rsinp, rsinsv, sinpr, sinsvr, pex, svex);
    double aa0[4];
    hska_apply(vv, aa0, cosp, rsinp, sinpr, cossv, rsinsv, sinsvr, (float)L.zrho, (float)L.zb,
               L.iwat, lw, K);
    double ex2 = normalize_pow2(aa0);
#pragma unroll
    for (int i = 0; i < 4; ++i) vv[i] = aa0[i];
    exsum = exsum + pex + ex2;
    // eigenfunction at the top of layer m+1 (svfunc)
    int i1 = m + 1;
    // the up-sweep vector there (loaded one interface ahead)
    const double cd1 = nx[0], cd2 = nx[1], cd3 = nx[2], cd4 = -cd3, cd5 = nx[3], cd6 = nx[4],
                 exe_i = nx[5];
    {
      const int i2 = min(i1 + 1, N - 1);
#pragma unroll
      for (int j = 0; j < 6; ++j) nx[j] = ck(i2, j);
    }
    double tz1 = -vv[3], tz2 = -vv[2], tz3 = vv[1], tz4 = vv[0];
    double uu1 = tz2 * cd6 - tz3 * cd5 + tz4 * cd4;
    double uu2 = -tz1 * cd6 + tz3 * cd3 - tz4 * cd2;
    double uu3 = tz1 * cd5 - tz2 * cd3 + tz4 * cd1;
    double uu4 = -tz1 * cd4 + tz2 * cd2 - tz3 * cd1;
    double ext = exsum + exe_i - exe0;
    U4 un;
    if (ext > -80.0 && ext < 80.0) {
      double fact = exp(ext) * r1213;
      un.ur = uu1 * fact;
      un.uz = uu2 * fact;
      un.tz = uu3 * fact;
      un.tr = uu4 * fact;
    } else {
      un.ur = un.uz = un.tz = un.tr = 0.0;
    }
    if (i1 < jwat) { un.ur = 0.0; un.tr = 0.0; }
    // energy of layer m (internal)
    double ka, kb, kr, urf;
    layer_energy(L, &lw, 0, uprev, un, K, c, S, ka, kb, kr, urf);
    long long o = D.base + (long long)m * D.stride;
    if (D.p[0]) D.p[0][o] = ka;
    if (D.p[1]) D.p[1][o] = kb;
    if (D.p[2]) D.p[2][o] = kr;
    if (eig) {
      A.eig[0][ebase + m] = urf;
      A.eig[1][ebase + m] = uprev.uz;
      A.eig[2][ebase + m] = uprev.tr;
      A.eig[3][ebase + m] = uprev.tz;
    }
    // getdcdh of layer m+1
    Layer Ln = load_layer(A, col, i1);
    double xmu_n = Ln.zrho * (Ln.zb * Ln.zb);
    double xlam_n = Ln.zrho * (Ln.za * Ln.za) - 2 * xmu_n;
    // reciprocals by the refined hardware estimate (lambda + 2 mu = rho a^2 > 0;
    // mu is zero in a fluid layer, where dcdh_terms does not read 1/mu)
    double ixmu_n = (xmu_n != 0.0) ? frcp(xmu_n) : 0.0, ixl2_n = frcp(xlam_n + xmu_n + xmu_n);
    if (D.p[3]) {
      double g = dcdh_terms(i1, Ln, L, xmu_n, xlam_n, xmu_prev, xlam_prev, ixl2_n, ixmu_n,
                            ixl2_prev, ixmu_prev, un, om2, wvno, wvno2);
      D.p[3][D.base + (long long)i1 * D.stride] = g;
    }
    uprev = un;
    Lprev = Ln;
    xmu_prev = xmu_n;
    xlam_prev = xlam_n;
    ixmu_prev = ixmu_n;
    ixl2_prev = ixl2_n;
    (void)Lm1; (void)xmu_m1; (void)xlam_m1;
  }
  // half-space energy
  {
    double ka, kb, kr, urf;
    layer_energy(Lprev, nullptr, 1, uprev, uprev, K, c, S, ka, kb, kr, urf);
    long long o = D.base + (long long)(N - 1) * D.stride;
    if (D.p[0]) D.p[0][o] = ka;
    if (D.p[1]) D.p[1][o] = kb;
    if (D.p[2]) D.p[2][o] = kr;
    if (eig) {
      A.eig[0][ebase + N - 1] = urf;
      A.eig[1][ebase + N - 1] = uprev.uz;
      A.eig[2][ebase + N - 1] = uprev.tr;
      A.eig[3][ebase + N - 1] = uprev.tz;
    }
  }
  ugr = (wvno * S.s1 + S.s2) / (omega * S.s0);
  double are = wvno / (2.0 * omega * ugr * S.s0);
  fac = are * c / wvno2;
  sumi0 = S.s0;
}

// ------------------------------------------------------------ Love parts
// slegn96.f90:248-328 varl
struct VarL {
  double cosq, sinq, yl, zl, mu, rb, irb, xkb, eexl;
  cplx E1;  // exp(-nub d): (cos q, -sin q) propagating, exp(-q) evanescent
};
JDSG_DEV VarL varl(const Layer &L, double omega, double wvno, double dpth) {
  VarL v;
  v.xkb = omega * frcp(L.zb);
  double wvnop = wvno + v.xkb;
  double wvnom = fabs(wvno - v.xkb);
  v.rb = sqrt(wvnop * wvnom);
  double q = v.rb * dpth;
  v.mu = L.zrho * L.zb * L.zb;
  v.irb = (v.rb > 0.0) ? frcp(v.rb) : 0.0;
  // scalars per branch, the fields assigned once after them (stores to
  // different fields in the branches would be merged into stores through
  // selected addresses, in private memory)
  double eexl = 0.0, sinq = 0.0, cosq, yl, zl, e1re, e1im = 0.0;
  if (wvno < v.xkb) {
    sincos(q, &sinq, &cosq);
    yl = sinq * v.irb;
    zl = -v.rb * sinq;
    e1re = cosq;
    e1im = -sinq;
  } else if (wvno == v.xkb) {
    cosq = 1.0;
    yl = dpth;
    zl = 0.0;
    e1re = 1.0;
  } else {
    eexl = q;
    const double e1 = exp(-q);
    e1re = e1;
    double fac = (q < 18.0) ? e1 * e1 : 0.0;
    cosq = (1.0 + fac) * 0.5;
    sinq = (1.0 - fac) * 0.5;
    yl = sinq * v.irb;
    zl = v.rb * sinq;
  }
  v.eexl = eexl;
  v.sinq = sinq;
  v.cosq = cosq;
  v.yl = yl;
  v.zl = zl;
  v.E1 = cmk(e1re, e1im);
  return v;
}

// Love layers need no Vp: four loads per layer instead of five
JDSG_DEV Layer load_layer_love(const EigenArgs &A, int col, int i, bool pu) {
  size_t q = (size_t)i * A.cm.ncol_pad + col;
  Layer L;
  L.za = 0.0;
  L.zb = A.cm.zb[q];
  L.zd = A.cm.zd[q];
  L.zrho = (A.iflsph > 0) ? A.cm.rho0[q] * A.cm.rtpL[q] : A.cm.rho0[q];
  L.iwat = pu ? ((L.zb > 0.0) ? 0 : 1) : ((L.zb == 0.0) ? 1 : 0);
  return L;
}
JDSG_DEV bool love_solid(const EigenArgs &A, int col, int i, bool pu) {
  double zb = A.cm.zb[(size_t)i * A.cm.ncol_pad + col];
  return pu ? (zb > 0.0) : (zb != 0.0);
}

// One Love pass (shfunc + energy); scratch holds uu, tt, exl per layer.
JDSG_DEV void love_pass(const EigenArgs &A, int col, int N, double omega, double wvno, bool pu,
                        double *scr, long long sstride, const Dst &D, double &ugr,
                        double &sumi1, double &fac, bool want_eig, long long ebase) {
#define UU(i) scr[((long long)(i) * 3 + 0) * sstride]
#define TT(i) scr[((long long)(i) * 3 + 1) * sstride]
#define EXL(i) scr[((long long)(i) * 3 + 2) * sstride]
  // up (slegn96.f90:372-445)
  {
    Layer H = load_layer_love(A, col, N - 1, pu);
    double xmuH = H.zrho * (H.zb * H.zb);
    if (H.zb > 0.01) {
      VarL v = varl(H, omega, wvno, 0.0);
      UU(N - 1) = 1.0;
      TT(N - 1) = -xmuH * v.rb;
    } else {
      UU(N - 1) = 1.0;
      TT(N - 1) = 0.0;
    }
    EXL(N - 1) = 0.0;
    double u1 = UU(N - 1), t1v = TT(N - 1);
    // layer k - 1 is loaded while layer k is swept
    Layer Lc = load_layer_love(A, col, max(N - 2, 0), pu);
#pragma unroll 4
    for (int k = N - 2; k >= 0; --k) {
      const Layer L = Lc;
      Lc = load_layer_love(A, col, max(k - 1, 0), pu);
      if (L.iwat != 0) {
        UU(k) = 0.0;
        TT(k) = 0.0;
        EXL(k) = 0.0;
        continue;
      }
      VarL v = varl(L, omega, wvno, L.zd);
      double a11 = v.cosq, a22 = v.cosq, a12 = -(v.yl * frcp(v.mu)), a21 = -(v.zl * v.mu);
      double at[2] = {a11 * u1 + a12 * (t1v), a21 * u1 + a22 * (t1v)};
      EXL(k) = normalize_pow2(at, 1.0e-30) + v.eexl;
      u1 = at[0];
      t1v = at[1];
      UU(k) = u1;
      TT(k) = t1v;
    }
  }
  // shfunc normalisation (slegn96.f90:179-246): the running exponent and
  // the largest normalised amplitude; the normalised amplitudes are formed
  // where they are used, so the sweep storage is read, not rewritten
  double umax = UU(0);
  if (umax == 0.0) {  // the search below ends with umax = UU(0) otherwise
    double ext = 0.0;
    for (int k = 1; k < N; ++k) {
      double uk = 0.0;
      if (love_solid(A, col, k, pu)) {
        ext = ext + EXL(k - 1);
        double fact = (ext < 80.0) ? 1. / exp(ext) : 0.0;
        uk = UU(k) * fact;
      }
      if (fabs(uk) > fabs(umax)) umax = uk;
    }
  }
  const bool norm = fabs(umax) > 0.0;
  const double iumax = norm ? 1.0 / umax : 0.0;
  // energy (slegn96.f90:447-629) fused with the dc/dh terms (:585-610,
  // stored before the factor fac, which needs the finished sums) and the
  // eigenfunctions
  double c = omega / wvno;
  double omega2 = omega * omega, wvno2 = wvno * wvno;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  bool solid0 = love_solid(A, col, 0, pu);
  double u_k = (norm && solid0) ? UU(0) * iumax : UU(0);
  double t_k = 0.0;
  const double ik2 = 1.0 / wvno2, hik = 0.5 / wvno;
  double ext = 0.0;
  int llflag = 0;
  double zr_prev = 0.0, xmu_prev = 0.0, ixmu_prev = 0.0;
  // Layer k + 1 and the sweep values that layer k reads (uu, tt at k + 1,
  // exl at k) come from loads issued one layer earlier, so they are in
  // flight during the previous layer's work (indices clamped to the stack).
  Layer Lk = load_layer_love(A, col, 0, pu);
  Layer Ln = load_layer_love(A, col, min(1, N - 1), pu);
  double uu_n = UU(min(1, N - 1)), tt_n = TT(min(1, N - 1)), exl_k = EXL(0);
  for (int k = 0; k < N; ++k) {
    const Layer L = Lk, L1 = Ln;
    const double uu1 = uu_n, tt1 = tt_n, exl = exl_k;
    {
      const int k2 = min(k + 2, N - 1);
      Lk = L1;
      Ln = load_layer_love(A, col, k2, pu);
      uu_n = UU(k2);
      tt_n = TT(k2);
      exl_k = EXL(min(k + 1, N - 1));
    }
    long long o = D.base + (long long)k * D.stride;
    // normalised amplitudes at the top of layer k+1
    double u_n = 0.0, t_n = 0.0;
    if (k + 1 < N) {
      if (L1.iwat == 0) {  // layer k + 1 solid
        ext = ext + exl;
        double fact = (ext < 80.0) ? exp(-ext) : 0.0;
        if (norm) fact = fact * iumax;
        u_n = uu1 * fact;
        t_n = tt1 * fact;
      }
    }
    double xmu = L.zrho * (L.zb * L.zb);
    // 1/mu by the refined hardware reciprocal, carried to the next layer
    const double ixmu = (xmu != 0.0) ? frcp(xmu) : 0.0;
    if (L.iwat != 0) {
      if (D.p[1]) D.p[1][o] = 0.0;
      if (D.p[2]) D.p[2][o] = 0.0;
      if (D.p[3]) D.p[3][o] = 0.0;
    } else {
      double TN = L.zrho * L.zb * L.zb, TL = TN;
      double VSHH = L.zb, VSHV = L.zb;
      double drho = L.zrho, dpth = L.zd;
      VarL v = varl(L, omega, wvno, dpth);
      double dmu = L.zrho * (L.zb * L.zb);
      double rb = v.rb;
      if (rb < 1.0e-10) rb = 1.0e-10;
      double upup, dupdup;
      double uk = u_k;
      if (k == N - 1) {
        upup = (0.5 / rb) * uk * uk;
        dupdup = (0.5 * rb) * uk * uk;
      } else {
        // nub is i rb or rb, so 1/nub and exp(-nub d) (from varl, unless rb
        // was raised to the floor above) need no complex division
        const bool prop = wvno < v.xkb;
        const double irb = (v.rb >= 1.0e-10) ? v.irb : 1.0 / 1.0e-10;  // rb at its floor
        const cplx inub = prop ? cmk(0.0, -irb) : cmk(irb, 0.0);
        const cplx ixn = (hik * ixmu) * inub;  // 0.5 / (wvno xnub), dmu = mu
        const double hk = hik;
        const cplx km1dn = hk * uk - ixn * t_k;
        const cplx kmup = hk * u_n + ixn * t_n;
        const double f3re = prop ? 0.0 : rb * dpth;
        const cplx E1 = (v.rb < 1.0e-10) ? cexp_(-(prop ? cmk(0.0, rb) : cmk(rb, 0.0)) * dpth)
                                         : v.E1;
        cplx exqq = (f3re < 40.0) ? E1 * E1 : cmk(0, 0);
        cplx f = (0.5 * (1.0 - exqq)) * inub;
        exqq = (f3re < 75.0) ? E1 : cmk(0, 0);
        cplx g = dpth * exqq;
        const double w2 = wvno * wvno;
        cplx f1 = f * (w2 * (kmup * kmup + km1dn * km1dn));
        cplx f2 = g * ((2.0 * w2) * (kmup * km1dn));
        upup = (f1 + f2).re;
        const double nub2 = prop ? -(rb * rb) : rb * rb;
        dupdup = nub2 * (f1 - f2).re;
      }
      s0 = s0 + drho * upup;
      s1 = s1 + TN * upup;
      s2 = s2 + TL * dupdup;
      double DCDBH = c * drho * VSHH * upup;
      double DCDBV = c * drho * VSHV * dupdup * ik2;
      if (D.p[1]) D.p[1][o] = DCDBH + DCDBV;
      if (D.p[2])
        D.p[2][o] = 0.5 * c * (-c * c * upup + VSHH * VSHH * upup + VSHV * VSHV * dupdup * ik2);
      // dc/dh terms before fac
      if (D.p[3]) {
        double drh, dm, dvdz;
        if (llflag == 0) {
          drh = L.zrho;
          dm = xmu;
          dvdz = 0.0;
        } else {
          drh = L.zrho - zr_prev;
          dm = xmu - xmu_prev;
          dvdz = t_k * t_k * (ixmu - ixmu_prev);
        }
        D.p[3][o] = uk * uk * (omega2 * drh - wvno2 * dm) + dvdz;
      }
      llflag = llflag + 1;
    }
    zr_prev = L.zrho;
    xmu_prev = xmu;
    ixmu_prev = ixmu;
    if (want_eig) {
      A.eig[0][ebase + k] = u_k;
      A.eig[1][ebase + k] = t_k;
    }
    u_k = u_n;
    t_k = t_n;
  }
  ugr = s1 / (c * s0);
  double ale = 0.5 / s1;
  fac = ale * c / wvno2;
  sumi1 = s1;
#undef UU
#undef TT
#undef EXL
}

JDSG_DEV void zero_entry(const EigenArgs &A, long long e, long long kb, long long ls, int N,
                         int st) {
  A.c_out[e] = 0.0;
  A.u_out[e] = 0.0;
  A.status[e] = st;
  for (int a = 0; a < 4; ++a) {
    if (A.dc[a])
      for (int i = 0; i < N; ++i) A.dc[a][kb + i * ls] = 0.0;
    if (A.du[a])
      for (int i = 0; i < N; ++i) A.du[a][kb + i * ls] = 0.0;
  }
}

}  // namespace

// ----------------------------------------------------------------------------
// The stage runs in chunks of one entry per lane.  For each chunk a pass
// kernel per period (T, and with group kernels T(1+.005), T(1-.005)) runs
// the sweep + energy pass of one lane's entry and leaves the raw kernels and
// (U, energy, fac) in the lane's scratch; a combine kernel then forms the
// outputs.  Splitting the passes from the assembly keeps each kernel's
// register use to one pass (the fused form needed ~450 registers and ran one
// wave per SIMD).
//
// Lane scratch (slots of `lanes` doubles): [0, 6 nmax) sweep storage;
// [(6 + 4v + a) nmax, +nmax) raw kernel a of pass v; 18 nmax + 3v + {0,1,2}
// the pass's U, energy sum and dc/dh factor.
namespace {

struct EntryRef {
  int col, mode, k, N;
  long long e, el, kb, ls;
};

JDSG_DEV EntryRef entry_ref(const EigenArgs &A, long long it) {
  EntryRef r;
  decode_item(A, it, r.col, r.mode, r.k, r.e);
  r.N = A.cm.nlay[r.col];
  r.el = r.e - A.ent_off[r.col];
  if (A.layout == 0) {
    r.kb = r.col + r.el * (long long)r.N * A.cm.ncol;
    r.ls = A.cm.ncol;
  } else {
    r.kb = A.kern_off[r.col] + r.el * r.N;
    r.ls = 1;
  }
  return r;
}

// 0: solve; JDSG_NO_ROOT / JDSG_EXIT_PATH: entry without outputs; *pm_ok
// false when the T(1+-0.005) roots are missing (group kernels)
JDSG_DEV int entry_status(const EigenArgs &A, long long e, bool *pm_ok) {
  int st = A.cstat[e];
  *pm_ok = true;
  if (st & JDSG_EXIT_PATH) return JDSG_EXIT_PATH;
  if (A.cflat[e] == 0.0) return JDSG_NO_ROOT;
  if (A.group) {
    int stpm = A.cstat[A.nent + e] | A.cstat[2 * A.nent + e];
    if ((stpm & JDSG_EXIT_PATH) || A.cflat[A.nent + e] == 0.0 || A.cflat[2 * A.nent + e] == 0.0)
      *pm_ok = false;
  }
  return 0;
}

// The chunk lane a wave works on.  In a grid batch the items of a chunk are
// C whole mode/period combinations of every column (item = combination * ncol
// + column), so the C waves over one block of 64 columns read the same column
// model.  Workgroup w runs on XCD w % 8; the mapping gives the waves that run
// on one XCD, close in dispatch order, the C combinations of one column block,
// so those model lines come from that XCD's L2.  A bijection on the chunk's
// waves; partial chunks and list-ordered batches keep the identity.  The
// pass and combine kernels use the same mapping, so a lane's scratch slot is
// the same in both.
JDSG_DEV long long chunk_lane(const EigenArgs &A, long long it0, long long lanes) {
  const long long ncol = A.cm.ncol;
  const long long w = blockIdx.x;
  if (!A.item_ent && ncol % 512 == 0 && lanes % ncol == 0 && it0 % ncol == 0 &&
      (long long)gridDim.x * 64 == lanes) {
    const long long B = ncol / 64, C = lanes / ncol;  // waves per combination, combinations
    const long long x = w & 7, j = w >> 3, h = j % C, g = j / C;
    return (h * B + 8 * g + x) * 64 + threadIdx.x;
  }
  return w * 64 + threadIdx.x;
}

JDSG_DEV double entry_period(const EigenArgs &A, const EntryRef &r, int v) {
  if (A.tv) return A.tv[3 * r.e + v];
  double T = A.periods[A.per_off[r.col] + r.k];
  return v == 0 ? T : (v == 1 ? T * (1.0 + 0.5 * 0.01) : T * (1.0 - 0.5 * 0.01));
}

}  // namespace

template <int W>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W == 0 ? 2 : 3)))
jdsg_eigen_pass(EigenArgs A, long long it0, long long lanes, int v) {
  const long long lane = chunk_lane(A, it0, lanes);
  const long long it = it0 + lane;
  if (lane >= lanes || it >= A.nitem) return;
  const EntryRef r = entry_ref(A, it);
  bool pm_ok;
  if (entry_status(A, r.e, &pm_ok)) return;
  if (v > 0 && !pm_ok) return;
  double *scr = A.scr + lane;
  const long long ss = lanes;
  const double twopi = 2.0 * kPiF;
  double omega = twopi / entry_period(A, r, v);
  double wvno = omega / A.cflat[(long long)v * A.nent + r.e];
  double *draw = scr + (long long)6 * A.nmax * ss;
  Dst D;
  // raw kernels of the requested outputs only (a null pointer skips a type)
  for (int a = 0; a < 4; ++a)
    D.p[a] = (A.dc[a] || A.du[a]) ? draw + (long long)(4 * v + a) * A.nmax * ss : nullptr;
  D.base = 0;
  D.stride = ss;
  long long ebase = (A.eig[0] && A.kern_off) ? A.kern_off[r.col] + r.el * r.N : 0;
  double ugr, s0, fac;
  if constexpr (W == 0) {
    rayleigh_pass(A, r.col, r.N, omega, wvno, scr, ss, D, ugr, s0, fac,
                  (v == 0 && A.eig[0]) ? A.eig[0] : nullptr, ebase);
  } else {
    love_pass(A, r.col, r.N, omega, wvno, A.group != 0, scr, ss, D, ugr, s0, fac,
              v == 0 && A.eig[0] != nullptr, ebase);
  }
  double *sc = scr + ((long long)18 * A.nmax + 3 * v) * ss;
  sc[0] = ugr;
  sc[ss] = s0;
  sc[2 * ss] = fac;
}

template <int W>
__global__ void __launch_bounds__(64)
jdsg_eigen_combine(EigenArgs A, long long it0, long long lanes) {
  const long long lane = chunk_lane(A, it0, lanes);
  const long long it = it0 + lane;
  if (lane >= lanes || it >= A.nitem) return;
  const bool ray = (W == 0);
  const EntryRef r = entry_ref(A, it);
  const int col = r.col, N = r.N;
  const long long e = r.e, kb = r.kb, ls = r.ls;
  bool pm_ok;
  int status = entry_status(A, e, &pm_ok);
  if (status) {
    zero_entry(A, e, kb, ls, N, status);
    return;
  }
  if (A.group && !pm_ok) {
    int stpm = A.cstat[A.nent + e] | A.cstat[2 * A.nent + e];
    status = (stpm & JDSG_EXIT_PATH) ? JDSG_EXIT_PATH_PM : JDSG_NO_ROOT_PM;
  }
  const double *scr = A.scr + lane;
  const long long ss = lanes;
  const double *draw = scr + (long long)6 * A.nmax * ss;
  const double *scal = scr + (long long)18 * A.nmax * ss;
  const double twopi = 2.0 * kPiF;
  const double ar = 6371.0;
  const double T = entry_period(A, r, 0);
  double u = 0.0, cout = 0.0;
  if (!A.group) {
    // ---------------- sregn96_/slegn96_ form
    double cc = A.cflat[e];
    double omega = twopi / T;
    double ugr = scal[0], s0 = scal[ss], fac = scal[2 * ss];
    if (ray) {
      if (fabs(cc) < 1.0e-36) cc = 0.0;
      if (fabs(ugr) < 1.0e-36) ugr = 0.0;
    }
    double tm = 1.0, tm3 = 1.0;
    if (A.iflsph > 0) {
      double rr = ray ? cc / (2. * ar * omega) : 3.0 * cc / (2. * ar * omega);
      tm = sqrt(1. + rr * rr);
      tm3 = tm * tm * tm;
      cout = cc / tm;
      u = ugr * tm;
    } else {
      cout = cc;
      u = ugr;
    }
    // finalise kernels bottom-up with the thickness suffix sums
    double hsum = 0.0;
    // flattening factors of the requested types only (the stage is bound by
    // HBM traffic)
    const bool need_v = A.dc[0] || A.dc[1], need_r = A.dc[2] != nullptr,
               need_d = A.dc[3] != nullptr;
    for (int i = N - 1; i >= 0; --i) {
      size_t q = (size_t)i * A.cm.ncol_pad + col;
      double vtp = need_v ? A.cm.vtp[q] : 0.0, dtp = need_d ? A.cm.dtp[q] : 0.0;
      double rtp = need_r ? (ray ? A.cm.rtpR[q] : A.cm.rtpL[q]) : 0.0;
      long long o = (long long)i * ss;
      const double *p0 = draw, *p1 = draw + (long long)A.nmax * ss,
                   *p2 = draw + (long long)2 * A.nmax * ss, *p3 = draw + (long long)3 * A.nmax * ss;
      // only the requested types were stored by the passes
      double ka = 0.0, kbv = 0.0, kr = 0.0, kh = 0.0;
      if (ray) {
        double us0 = ugr * s0;
        if (A.dc[0]) ka = p0[o] / us0;
        if (A.dc[1]) kbv = p1[o] / us0;
        if (A.dc[2]) kr = p2[o] / us0;
        if (A.dc[3]) kh = fac * p3[o];
        if (fabs(kh) < 1.0e-38) kh = 0.0;
      } else {
        // Love: dcdb, dcdr divided by sumi1, dcdh final
        Layer L = load_layer_love(A, col, i, false);
        if (A.dc[1]) kbv = (L.iwat == 0) ? p1[o] / s0 : 0.0;
        if (A.dc[2]) kr = (L.iwat == 0) ? p2[o] / s0 : 0.0;
        if (A.dc[3]) kh = fac * p3[o];
        if (fabs(kh) < 1.0e-38) kh = 0.0;
      }
      if (A.iflsph > 0) {
        ka = ka * vtp / tm3;
        kbv = kbv * vtp / tm3;
        kh = kh * dtp / tm3;
        kr = kr * rtp / tm3;
      }
      long long oo = kb + (long long)i * ls;
      if (A.dc[0]) A.dc[0][oo] = ka;
      if (A.dc[1]) A.dc[1][oo] = kbv;
      if (A.dc[2]) A.dc[2][oo] = kr;
      if (A.dc[3]) A.dc[3][oo] = hsum;
      hsum = hsum + kh;
    }
  } else {
    // ---------------- sregnpu_/slegnpu_ form
    double tt1 = entry_period(A, r, 1), tt2 = entry_period(A, r, 2);
    double usc[3], facv[3];
    for (int v = 0; v < 3; ++v) {
      double ugr = scal[(long long)(3 * v) * ss], s0 = scal[(long long)(3 * v + 1) * ss];
      usc[v] = ray ? ugr * s0 : s0;
      facv[v] = scal[(long long)(3 * v + 2) * ss];
    }
    double cp = A.cflat[e], cg = scal[0];
    double uc1 = cg / cp;
    double dtv = tt2 - tt1;
    double tm = 1.0, tm1 = 0.0, tm3 = 1.0;
    if (A.iflsph > 0) {
      double omega = twopi / T;
      double rr = ray ? cp / (2. * ar * omega) : 3.0 * cp / (2. * ar * omega);
      tm = sqrt(1. + rr * rr);
      double q = (ray ? 0.5 : 1.5) / (ar * omega);
      tm1 = q * q / tm;
      tm3 = tm * tm * tm;
    }
    double hsc = 0.0, hsu = 0.0;
    const bool need_v = A.dc[0] || A.du[0] || A.dc[1] || A.du[1],
               need_r = A.dc[2] || A.du[2], need_d = A.dc[3] || A.du[3];
    // the entry's divisors as reciprocals: the layer loop is bound by its
    // loads, and a division is about ten instructions between them
    const double iusc[3] = {1.0 / usc[0], 1.0 / usc[1], 1.0 / usc[2]};
    const double udu = uc1 * uc1 * (T / dtv), itm3 = 1.0 / tm3;
#pragma unroll 2
    for (int i = N - 1; i >= 0; --i) {
      size_t q = (size_t)i * A.cm.ncol_pad + col;
      double vtp = need_v ? A.cm.vtp[q] : 0.0, dtp = need_d ? A.cm.dtp[q] : 0.0;
      double rtp = need_r ? (ray ? A.cm.rtpR[q] : A.cm.rtpL[q]) : 0.0;
      long long oo = kb + (long long)i * ls;
      long long os = (long long)i * ss;
      bool solid = true;
      if (!ray) solid = (A.cm.zb[q] > 0.0);
      double fct[4] = {vtp, vtp, rtp, dtp};
      for (int a = 0; a < 4; ++a) {
        if (!A.dc[a] && !A.du[a]) continue;  // not stored by the passes
        double r0 = draw[(long long)(0 + a) * A.nmax * ss + os];
        double r1 = pm_ok ? draw[(long long)(4 + a) * A.nmax * ss + os] : 0.0;
        double r2 = pm_ok ? draw[(long long)(8 + a) * A.nmax * ss + os] : 0.0;
        double d0, d1, d2;
        if (a == 3) {  // dc/dh: the passes leave it before the factor fac
          d0 = facv[0] * r0; if (fabs(d0) < 1.0e-38) d0 = 0.0;
          d1 = facv[1] * r1; if (fabs(d1) < 1.0e-38) d1 = 0.0;
          d2 = facv[2] * r2; if (fabs(d2) < 1.0e-38) d2 = 0.0;
        } else if (ray) {
          d0 = r0 * iusc[0];
          d1 = pm_ok ? r1 * iusc[1] : 0.0;
          d2 = pm_ok ? r2 * iusc[2] : 0.0;
        } else {
          d0 = (a != 0 && solid) ? r0 * iusc[0] : 0.0;
          d1 = (a != 0 && solid && pm_ok) ? r1 * iusc[1] : 0.0;
          d2 = (a != 0 && solid && pm_ok) ? r2 * iusc[2] : 0.0;
        }
        double du = uc1 * (2.0 - uc1) * d2 - udu * (d2 - d1);
        double dc = d0;
        if (A.iflsph > 0) {
          du = (tm * du + cg * cp * dc * tm1) * fct[a];
          dc = dc * itm3 * fct[a];
        }
        if (!ray && a == 0) { du = 0.0; dc = 0.0; }
        if (!pm_ok) du = 0.0;
        if (a < 3) {
          if (A.dc[a]) A.dc[a][oo] = dc;
          if (A.du[a]) A.du[a][oo] = du;
        } else {
          if (A.dc[3]) A.dc[3][oo] = hsc;
          if (A.du[3]) A.du[3][oo] = hsu;
          hsc = hsc + dc;
          hsu = hsu + du;
        }
      }
    }
    if (A.iflsph > 0) {
      cout = cp / tm;
      u = cg * tm;
    } else {
      cout = cp;
      u = cg;
    }
  }
  A.c_out[e] = cout;
  A.u_out[e] = u;
  A.status[e] = status;
}

int jdsg_eigen_pass_launches(const EigenArgs &A, int nwaves) {
  const long long lanes = (long long)nwaves * 64;
  return (int)((A.nitem + lanes - 1) / lanes) * (A.group ? 3 : 1);
}

// Chunks of `lanes` entries: the pass(es), then the assembly.
void jdsg_launch_eigen(const EigenArgs &A, int nwaves, hipStream_t st, hipEvent_t *ev) {
  const long long lanes = (long long)nwaves * 64;
  const int nv = A.group ? 3 : 1;
  int np = 0;
  for (long long it0 = 0; it0 < A.nitem; it0 += lanes) {
    long long n = A.nitem - it0 < lanes ? A.nitem - it0 : lanes;
    dim3 g((unsigned)((n + 63) / 64));
    for (int v = 0; v < nv; ++v, ++np) {
      if (ev) (void)hipEventRecord(ev[2 * np], st);
      if (A.wave == 0)
        hipLaunchKernelGGL(jdsg_eigen_pass<0>, g, dim3(64), 0, st, A, it0, lanes, v);
      else
        hipLaunchKernelGGL(jdsg_eigen_pass<1>, g, dim3(64), 0, st, A, it0, lanes, v);
      if (ev) (void)hipEventRecord(ev[2 * np + 1], st);
    }
    if (A.wave == 0)
      hipLaunchKernelGGL(jdsg_eigen_combine<0>, g, dim3(64), 0, st, A, it0, lanes);
    else
      hipLaunchKernelGGL(jdsg_eigen_combine<1>, g, dim3(64), 0, st, A, it0, lanes);
  }
}
