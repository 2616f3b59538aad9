// swd_common.hpp -- shared device helpers for the CDNA4 dispersion/kernel
// path: a small complex-double type with the operation order the Fortran
// COMPLEX(8) code uses, the batch layout descriptors, and constants.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define JDSG_DEV __device__ __forceinline__

// ---------------------------------------------------------------- complex
struct cplx {
  double re, im;
};
JDSG_DEV cplx cmk(double r, double i) { return cplx{r, i}; }
JDSG_DEV cplx operator+(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
JDSG_DEV cplx operator-(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
JDSG_DEV cplx operator-(cplx a) { return cplx{-a.re, -a.im}; }
JDSG_DEV cplx operator*(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
JDSG_DEV cplx operator*(double s, cplx a) { return cplx{s * a.re, s * a.im}; }
JDSG_DEV cplx operator*(cplx a, double s) { return cplx{a.re * s, a.im * s}; }
JDSG_DEV cplx operator/(cplx a, double s) { return cplx{a.re / s, a.im / s}; }
JDSG_DEV cplx operator+(cplx a, double s) { return cplx{a.re + s, a.im}; }
JDSG_DEV cplx operator+(double s, cplx a) { return cplx{s + a.re, a.im}; }
JDSG_DEV cplx operator-(double s, cplx a) { return cplx{s - a.re, -a.im}; }
JDSG_DEV cplx operator-(cplx a, double s) { return cplx{a.re - s, a.im}; }
// Smith's division, the libgcc/libm form for operands of ordinary range.
JDSG_DEV cplx operator/(cplx a, cplx b) {
  double c = b.re, d = b.im;
  if (fabs(c) < fabs(d)) {
    double r = c / d, den = c * r + d;
    return cplx{(a.re * r + a.im) / den, (a.im * r - a.re) / den};
  }
  double r = d / c, den = d * r + c;
  return cplx{(a.im * r + a.re) / den, (a.im - a.re * r) / den};
}
JDSG_DEV cplx operator/(double s, cplx b) { return cplx{s, 0.0} / b; }
JDSG_DEV cplx cconj(cplx a) { return cplx{a.re, -a.im}; }
// principal square root of a real number seen as complex (x + 0i)
// (one square root and two scalar selects: a select between two aggregates
// becomes a select between their addresses, which puts them in private memory)
JDSG_DEV cplx csqrt_real(double x) {
  const bool pos = x >= 0.0;
  const double s = sqrt(pos ? x : -x);
  return cplx{pos ? s : 0.0, pos ? 0.0 : s};
}
// exp of a complex number; a real or imaginary argument takes only the one
// function it needs (exp(0) = 1, sincos(0) = (0, 1) exactly, so the value is
// the same)
JDSG_DEV cplx cexp_(cplx z) {
  double e = (z.re == 0.0) ? 1.0 : exp(z.re), s = 0.0, c = 1.0;
  if (z.im != 0.0) sincos(z.im, &s, &c);
  return cplx{e * c, e * s};
}
JDSG_DEV double cabs_(cplx z) { return hypot(z.re, z.im); }
JDSG_DEV double sign1(double x) { return copysign(1.0, x); }

// ---------------------------------------------------------------- layout
// Per-column flattened model for the root chains (surfdisp96 `sphere`,
// R = 6370 km): float4 {d, a, b, rho_wave} per layer, stored
// [layer * ncol_pad + col] so a wavefront over columns loads contiguously.
// Per-column flattened model for the eigen stage (bldsph, R = 6371 km), in
// double, same [layer * ncol_pad + col] order:
//   zd, za, zb, rtpR, rtpL, vtp, dtp and the unflattened rho (rho0).
// With iflsph = 0 the flattening factors are 1 and zd/za/zb the inputs.
struct ColumnModel {
  int ncol, ncol_pad, nmax;
  const int *nlay;        // [ncol] layers incl. half-space
  const float4 *chainR;   // [nmax * ncol_pad]
  const float4 *chainL;
  const float *cc;        // [ncol] starting phase velocity (float, cc1 after backing off)
  const float *betmx;     // [ncol]
  const int *llw;         // [ncol] 1, or 2 for a water top layer
  const double *zd, *za, *zb, *rho0, *rtpR, *rtpL, *vtp, *dtp;  // [nmax * ncol_pad]
};

// status bits per entry
enum : int {
  JDSG_OK = 0,
  JDSG_NO_ROOT = 1,        // no root at T (mode beyond cutoff, or chain failure)
  JDSG_NO_ROOT_PM = 2,     // root at T but none at T*(1 +- 0.005): dU/dm zero
  JDSG_EXIT_PATH = 4,      // the reference wrapper would exit(1) in the T chain
  JDSG_EXIT_PATH_PM = 8,   // ... in a T*(1 +- 0.005) chain: dU/dm zero
};

#define JDSG_MAXMODE 8
