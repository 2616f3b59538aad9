// flopd.hpp -- an op-counting scalar type for the oracle restatement
// (TEST INFRASTRUCTURE ONLY).  The restatement's C sources are compiled as
// C++ with `double` replaced by FlopD and `double complex` by FlopC (the
// recipe in oracle/Makefile, target `flopcount`), so running them counts
// their floating-point operations.  Counting rule (the frozen flop model's):
// +, -, *, /, sqrt, exp, log, sin, cos count 1 each; pow counts 3 (log, mul,
// exp); comparisons, fabs and copies count 0.  Complex operations count the
// real operations of their textbook forms: add 2, multiply 6, divide 9
// (Smith's form), exp 5 (exp, sin, cos, two multiplies), sqrt of a real
// argument 1, abs 4 (two multiplies, an add, a sqrt).
#pragma once
#include <cmath>
#include <cstdint>

extern unsigned long long g_flops;

struct FlopD {
  double v;
  FlopD() = default;
  constexpr FlopD(double x) : v(x) {}
  explicit operator double() const { return v; }
  explicit operator float() const { return (float)v; }
  explicit operator int() const { return (int)v; }
  explicit operator long long() const { return (long long)v; }
  explicit operator bool() const { return v != 0.0; }
  FlopD &operator+=(FlopD o) { ++g_flops; v += o.v; return *this; }
  FlopD &operator-=(FlopD o) { ++g_flops; v -= o.v; return *this; }
  FlopD &operator*=(FlopD o) { ++g_flops; v *= o.v; return *this; }
  FlopD &operator/=(FlopD o) { ++g_flops; v /= o.v; return *this; }
  FlopD operator-() const { return FlopD(-v); }
  FlopD operator+() const { return *this; }
};
inline FlopD operator+(FlopD a, FlopD b) { ++g_flops; return FlopD(a.v + b.v); }
inline FlopD operator-(FlopD a, FlopD b) { ++g_flops; return FlopD(a.v - b.v); }
inline FlopD operator*(FlopD a, FlopD b) { ++g_flops; return FlopD(a.v * b.v); }
inline FlopD operator/(FlopD a, FlopD b) { ++g_flops; return FlopD(a.v / b.v); }
#define FLOPD_CMP(op)                                                   \
  inline bool operator op(FlopD a, FlopD b) { return a.v op b.v; }        \
  inline bool operator op(FlopD a, double b) { return a.v op b; }         \
  inline bool operator op(double a, FlopD b) { return a op b.v; }         \
  inline bool operator op(FlopD a, float b) { return a.v op (double)b; }  \
  inline bool operator op(float a, FlopD b) { return (double)a op b.v; }  \
  inline bool operator op(FlopD a, int b) { return a.v op (double)b; }    \
  inline bool operator op(int a, FlopD b) { return (double)a op b.v; }
FLOPD_CMP(<)
FLOPD_CMP(>)
FLOPD_CMP(<=)
FLOPD_CMP(>=)
FLOPD_CMP(==)
FLOPD_CMP(!=)
#undef FLOPD_CMP
#define FLOPD_MIX(op)                                                        \
  inline FlopD operator op(FlopD a, double b) { return a op FlopD(b); }     \
  inline FlopD operator op(double a, FlopD b) { return FlopD(a) op b; }     \
  inline FlopD operator op(FlopD a, float b) { return a op FlopD(b); }      \
  inline FlopD operator op(float a, FlopD b) { return FlopD(a) op b; }      \
  inline FlopD operator op(FlopD a, int b) { return a op FlopD(b); }        \
  inline FlopD operator op(int a, FlopD b) { return FlopD(a) op b; }
FLOPD_MIX(+)
FLOPD_MIX(-)
FLOPD_MIX(*)
FLOPD_MIX(/)
#undef FLOPD_MIX

inline FlopD sqrt(FlopD a) { ++g_flops; return FlopD(std::sqrt(a.v)); }
inline FlopD exp(FlopD a) { ++g_flops; return FlopD(std::exp(a.v)); }
inline FlopD log(FlopD a) { ++g_flops; return FlopD(std::log(a.v)); }
inline FlopD sin(FlopD a) { ++g_flops; return FlopD(std::sin(a.v)); }
inline FlopD cos(FlopD a) { ++g_flops; return FlopD(std::cos(a.v)); }
inline FlopD pow(FlopD a, FlopD b) { g_flops += 3; return FlopD(std::pow(a.v, b.v)); }
inline FlopD pow(FlopD a, double b) { g_flops += 3; return FlopD(std::pow(a.v, b)); }
inline FlopD pow(FlopD a, int b) { g_flops += 3; return FlopD(std::pow(a.v, b)); }
inline FlopD fabs(FlopD a) { return FlopD(std::fabs(a.v)); }
inline FlopD copysign(FlopD a, FlopD b) { return FlopD(std::copysign(a.v, b.v)); }
inline FlopD copysign(double a, FlopD b) { return FlopD(std::copysign(a, b.v)); }
inline FlopD fmin(FlopD a, FlopD b) { return FlopD(std::fmin(a.v, b.v)); }
inline FlopD fmax(FlopD a, FlopD b) { return FlopD(std::fmax(a.v, b.v)); }
inline bool isnan(FlopD a) { return std::isnan(a.v); }
using std::sqrt;
using std::exp;
using std::log;
using std::pow;
using std::fabs;

struct FlopC {
  FlopD re, im;
  FlopC() = default;
  constexpr FlopC(double r) : re(r), im(0.0) {}
  constexpr FlopC(FlopD r) : re(r), im(0.0) {}
  FlopC(FlopD r, FlopD i) : re(r), im(i) {}
};
inline FlopC CMPLX_(FlopD r, FlopD i) { return FlopC(r, i); }
inline FlopC operator+(FlopC a, FlopC b) { return FlopC(a.re + b.re, a.im + b.im); }
inline FlopC operator-(FlopC a, FlopC b) { return FlopC(a.re - b.re, a.im - b.im); }
inline FlopC operator-(FlopC a) { return FlopC(-a.re, -a.im); }
inline FlopC operator*(FlopC a, FlopC b) {
  return FlopC(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
inline FlopC operator/(FlopC a, FlopC b) {
  // Smith's form: 9 operations whichever branch
  if (std::fabs(b.re.v) >= std::fabs(b.im.v)) {
    FlopD r = b.im / b.re, den = b.re + b.im * r;
    return FlopC((a.re + a.im * r) / den, (a.im - a.re * r) / den);
  }
  FlopD r = b.re / b.im, den = b.re * r + b.im;
  return FlopC((a.re * r + a.im) / den, (a.im * r - a.re) / den);
}
#define FLOPC_MIX(op)                                                           \
  inline FlopC operator op(FlopC a, FlopD b) { return a op FlopC(b); }         \
  inline FlopC operator op(FlopD a, FlopC b) { return FlopC(a) op b; }         \
  inline FlopC operator op(FlopC a, double b) { return a op FlopC(FlopD(b)); } \
  inline FlopC operator op(double a, FlopC b) { return FlopC(FlopD(a)) op b; } \
  inline FlopC operator op(FlopC a, int b) { return a op FlopC(FlopD(b)); }    \
  inline FlopC operator op(int a, FlopC b) { return FlopC(FlopD(a)) op b; }
FLOPC_MIX(+)
FLOPC_MIX(-)
FLOPC_MIX(*)
FLOPC_MIX(/)
#undef FLOPC_MIX
inline FlopC &operator+=(FlopC &a, FlopC b) { return a = a + b; }
inline FlopC &operator-=(FlopC &a, FlopC b) { return a = a - b; }
inline FlopC &operator*=(FlopC &a, FlopC b) { return a = a * b; }
inline bool operator==(FlopC a, FlopC b) { return a.re == b.re && a.im == b.im; }
inline bool operator!=(FlopC a, FlopC b) { return !(a == b); }
inline FlopD creal(FlopC a) { return a.re; }
inline FlopD cimag(FlopC a) { return a.im; }
inline FlopC conj(FlopC a) { return FlopC(a.re, -a.im); }
inline FlopD cabs(FlopC a) { return sqrt(a.re * a.re + a.im * a.im); }
inline FlopC cexp(FlopC a) {
  FlopD e = exp(a.re);
  return FlopC(e * cos(a.im), e * sin(a.im));
}
// square root of the real-axis arguments the restatement passes
inline FlopC csqrt(FlopC a) {
  if (a.im.v != 0.0) {  // general principal root (not reached by the restatement)
    double m = std::sqrt(a.re.v * a.re.v + a.im.v * a.im.v);
    double r = std::sqrt(0.5 * (m + a.re.v)), i = std::copysign(std::sqrt(0.5 * (m - a.re.v)), a.im.v);
    g_flops += 8;
    return FlopC(FlopD(r), FlopD(i));
  }
  if (a.re.v >= 0.0) return FlopC(sqrt(a.re), FlopD(0.0));
  return FlopC(FlopD(0.0), sqrt(-a.re));
}
