// swd_surrogate.hip -- learned-bracket root finding.
//
// The MLP predicts the flat phase velocity of every (column, wave, mode,
// period).  Each root is then bracketed at +-half_width around the prediction
// (one period-equation evaluation per end), accepted only if the period
// equation changes sign, and polished in FP64 by Illinois false position to
// |c1-c2| <= tol*c.  A polished root is kept only if it is the root the
// reference's chain would reach: the mode-order pass checks, with the signs
// already evaluated, that no odd number of roots lies between the previous
// mode's bracket (or, for the lowest mode, the chain's floor cm) and this
// bracket.  Roots that fail any check are left to the plain chain, which walks
// each chain in the reference order and solves only those (the `known` input
// of the chain kernel).  Every lane works on an independent root: no
// sequential chain.
#include "swd_kernels.hpp"
#include "swd_dltar.hpp"

#pragma clang fp contract(off)

namespace {
constexpr double kTwoPi = 2.0 * 3.141592653589793;

__device__ __forceinline__ void bracket_ends(double cen, double hw, double cm, double *a,
                                             double *b) {
  *a = cen * (1.0 - hw);
  *b = cen * (1.0 + hw);
  if (*a < cm) *a = cm;
}
}  // namespace

// One lane per entry e (all waves): decode column by binary search.
extern "C" __global__ void __launch_bounds__(64)
jdsg_polish(PolishArgs A) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long nev = 0;
  unsigned nf = 0;
  if (e < A.nitem) {
  int ncol = A.cm.ncol;
  int lo = 0, hi = ncol - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (A.ent_off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  int col = lo;
  int K = A.per_off[col + 1] - A.per_off[col];
  long long el = e - A.ent_off[col];
  int wm = (int)(el / K), k = (int)(el % K);
  int nR = A.nmode[2 * col];
  bool ray = wm < nR;
  double T = A.periods[A.per_off[col] + k];
  if (A.variant == 1) T = T * (1.0 + 0.5 * 0.01);
  if (A.variant == 2) T = T * (1.0 - 0.5 * 0.01);
  double omega = kTwoPi / T;
  const float4 *st = (ray ? A.cm.chainR : A.cm.chainL) + col;
  int stride = A.cm.ncol_pad, mmax = A.cm.nlay[col], llw = A.cm.llw[col];
  auto F = [&](double c) {
    double wv = omega / c;
    return ray ? jdsg_pe::dltar4(st, stride, mmax, llw, wv, omega)
               : jdsg_pe::dltar1(st, stride, mmax, llw, wv, omega);
  };
  double cen = A.center[e];
  double root = 0.0;
  double betmx = (double)A.cm.betmx[col], cm = (double)A.cm.cc[col];
  signed char *sg = A.sg + ((long long)A.variant * A.nent + e) * 3;
  signed char s_lo = 0, s_hi = 0, s_floor = 0;
  int mode = ray ? wm : wm - nR;
  if (mode == 0) {
    s_floor = (signed char)sign1(F(cm));
    ++nev;
  }
  if (cen > 0.0 && cen < betmx + 0.5) {
    double a, b;
    bracket_ends(cen, A.half_width, cm, &a, &b);
    double fa = F(a), fb = F(b);
    nev += 2;
    s_lo = (signed char)sign1(fa);
    s_hi = (signed char)sign1(fb);
    if (sign1(fa) != sign1(fb)) {
      // Illinois false position
      for (int it = 0; it < 80; ++it) {
        double c = (a * fb - b * fa) / (fb - fa);
        if (!(c > fmin(a, b) && c < fmax(a, b))) c = 0.5 * (a + b);
        double fc = F(c);
        ++nev;
        if (fc == 0.0) { a = b = c; break; }
        if (sign1(fc) != sign1(fb)) {
          a = b;
          fa = fb;
        } else {
          fa = 0.5 * fa;
        }
        b = c;
        fb = fc;
        if (fabs(b - a) <= A.tol * fabs(b)) break;
      }
      root = b;
      if (!(root > cm) || root > betmx) root = 0.0;
    }
  }
  A.known[(long long)A.variant * A.nent + e] = root;
  sg[0] = s_lo;
  sg[1] = s_hi;
  sg[2] = s_floor;
  if (root == 0.0) nf = 1;
  }
  // one atomic per wavefront for the counters
  for (int o = 32; o > 0; o >>= 1) {
    nev += __shfl_xor(nev, o);
    nf += __shfl_xor(nf, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (nf) atomicAdd(A.nfail, nf);
    if (A.nevals && nev) atomicAdd(A.nevals, nev);
  }
}

// Mode order and root identity, one thread per (column, wave, period).  The
// chain reaches mode m as the first sign change above mode m-1 (the lowest
// mode: above its floor cm), with c_m > c_{m-1} + 0.01 dc
// (surfdisp96.f:270-276).  A polished root is kept only if
//   - the mode below it is kept (otherwise its chain start is not known),
//   - the period equation has the same sign at cm and at the bottom of the
//     lowest mode's bracket, and at the top of bracket m-1 and the bottom of
//     bracket m, so no odd number of roots is skipped in between,
//   - brackets m-1 and m do not overlap and c_m clears the floor gap.
// A failed pairwise check between m-1 and m leaves it open which of the two
// is off, so both go back to the chain; the first failure sends every mode
// above it to the chain as well.  What these checks cannot see is an even
// number of skipped roots with consistent brackets above it.
extern "C" __global__ void jdsg_mode_order(PolishArgs A, int kmax) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int ncol = A.cm.ncol;
  if (t >= 2LL * ncol * kmax) return;
  int col = (int)(t % ncol);
  long long r = t / ncol;
  int w = (int)(r & 1);
  int k = (int)(r >> 1);
  int K = A.per_off[col + 1] - A.per_off[col];
  if (k >= K) return;
  int nR = A.nmode[2 * col], nL = A.nmode[2 * col + 1];
  int M = w == 0 ? nR : nL, off = w == 0 ? 0 : nR;
  const double floor_gap = 1.0e-2 * (double)0.005f;
  const double cm = (double)A.cm.cc[col];
  double prev = 0.0, prev_hi = cm;
  double *prev_p = nullptr;
  int prev_sign = 0;
  bool bad = false;
  unsigned nf = 0;
  for (int m = 0; m < M; ++m) {
    long long e = A.ent_off[col] + (long long)(off + m) * K + k;
    long long q = (long long)A.variant * A.nent + e;
    double *p = A.known + q;
    double v = *p;
    const signed char *sg = A.sg + q * 3;
    if (m == 0) prev_sign = sg[2];
    if (!bad) {
      if (!(v > 0.0)) {  // not bracketed: this mode and the ones above go to the chain
        bad = true;
        continue;
      }
      double a, b;
      bracket_ends(A.center[e], A.half_width, cm, &a, &b);
      bool ok = prev_sign != 0 && sg[0] == prev_sign &&
                (m == 0 || (v > prev + floor_gap && a > prev_hi));
      if (ok) {
        prev = v;
        prev_hi = b;
        prev_sign = sg[1];
        prev_p = p;
        continue;
      }
      bad = true;
      if (prev_p) {
        *prev_p = 0.0;
        ++nf;
      }
    }
    if (v > 0.0) {
      *p = 0.0;
      ++nf;
    }
  }
  if (nf) atomicAdd(A.nfail, nf);
}

// Bracket centres for T(1 +- .005) from the polished T roots and the slope
// d ln c / d ln T of the MLP predictions along the period axis.
extern "C" __global__ void jdsg_pm_centers(const long long *ent_off, const int *per_off,
                                           const double *periods, int ncol, long long nent,
                                           const double *known0, const double *pred,
                                           double *cen1, double *cen2) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nent) return;
  int lo = 0, hi = ncol - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (ent_off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  int col = lo;
  int K = per_off[col + 1] - per_off[col];
  long long el = e - ent_off[col];
  int k = (int)(el % K);
  const double *T = periods + per_off[col];
  double c0 = known0[e] > 0.0 ? known0[e] : pred[e];
  double slope = 0.0;
  int ka = k > 0 ? k - 1 : k, kb = k < K - 1 ? k + 1 : k;
  if (kb > ka) {
    double pa = pred[e - (k - ka)], pb = pred[e + (kb - k)];
    if (pa > 0.0 && pb > 0.0) slope = log(pb / pa) / log(T[kb] / T[ka]);
  }
  double l1 = log(1.0 + 0.5 * 0.01), l2 = log(1.0 - 0.5 * 0.01);
  cen1[e] = c0 * exp(slope * l1);
  cen2[e] = c0 * exp(slope * l2);
}

// Per-column MLP features: Vs, Vp, rho sampled at 48 depths (2 km apart,
// layer mid-points z = 1, 3, ..., 95 km; below the stack the half-space),
// normalised.  Entry features append ln T, the wave and the mode one-hot.
// X row layout (ld = 192 bf16): [0,144) profile, 144 ln T, 145-146 wave,
// 147-150 mode, rest zero.
extern "C" __global__ void jdsg_column_features(int ncol, const int *layer_off, const float *thk,
                                                const float *vp, const float *vs,
                                                const float *rho, float *colf) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  int l0 = layer_off[col], N = layer_off[col + 1] - l0;
  float ztop = 0.f;
  int i = 0;
  for (int j = 0; j < 48; ++j) {
    float z = 1.0f + 2.0f * j;
    while (i < N - 1 && ztop + thk[l0 + i] <= z) {
      ztop += thk[l0 + i];
      ++i;
    }
    float s = vs[l0 + i], p = vp[l0 + i], r = rho[l0 + i];
    colf[(size_t)col * 144 + j] = (s - 3.5f) / 0.8f;
    colf[(size_t)col * 144 + 48 + j] = (p - 6.2f) / 1.2f;
    colf[(size_t)col * 144 + 96 + j] = (r - 2.8f) / 0.4f;
  }
}

__device__ __forceinline__ uint16_t f2bf_(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// One thread per (entry, feature): X[e][f] (bf16, ld 192) for the entries
// of variant-0 periods.
extern "C" __global__ void jdsg_entry_features(long long e0, long long nrow, int ncol,
                                               const long long *ent_off, const int *per_off,
                                               const double *periods, const int *nmode,
                                               const float *colf, const long long *rows,
                                               uint16_t *X) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrow * 192) return;
  long long e = rows ? rows[t / 192] : e0 + t / 192;
  int f = (int)(t % 192);
  int lo = 0, hi = ncol - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (ent_off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  int col = lo;
  int K = per_off[col + 1] - per_off[col];
  long long el = e - ent_off[col];
  int wm = (int)(el / K), k = (int)(el % K);
  int nR = nmode[2 * col];
  int wave = wm < nR ? 0 : 1;
  int mode = wave == 0 ? wm : wm - nR;
  float v = 0.f;
  if (f < 144) v = colf[(size_t)col * 144 + f];
  else if (f == 144) v = (float)((log(periods[per_off[col] + k]) - 2.0) / 1.5);
  else if (f == 145) v = wave == 0 ? 1.f : 0.f;
  else if (f == 146) v = wave == 1 ? 1.f : 0.f;
  else if (f >= 147 && f < 151) v = (mode == f - 147) ? 1.f : 0.f;
  X[t] = f2bf_(v);
}

void jdsg_launch_polish(const PolishArgs &A, hipStream_t st) {
  hipLaunchKernelGGL(jdsg_polish, dim3((unsigned)((A.nitem + 63) / 64)), dim3(64), 0, st, A);
}
void jdsg_launch_mode_order(const PolishArgs &A, int kmax, hipStream_t st) {
  long long n = 2LL * A.cm.ncol * kmax;
  hipLaunchKernelGGL(jdsg_mode_order, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, A,
                     kmax);
}
void jdsg_launch_pm_centers(const long long *ent_off, const int *per_off, const double *periods,
                            int ncol, long long nent, const double *known0, const double *pred,
                            double *cen1, double *cen2, hipStream_t st) {
  hipLaunchKernelGGL(jdsg_pm_centers, dim3((unsigned)((nent + 255) / 256)), dim3(256), 0, st,
                     ent_off, per_off, periods, ncol, nent, known0, pred, cen1, cen2);
}
void jdsg_launch_column_features(int ncol, const int *layer_off, const float *thk,
                                 const float *vp, const float *vs, const float *rho, float *colf,
                                 hipStream_t st) {
  hipLaunchKernelGGL(jdsg_column_features, dim3((ncol + 255) / 256), dim3(256), 0, st, ncol,
                     layer_off, thk, vp, vs, rho, colf);
}
void jdsg_launch_entry_features(long long e0, long long nrow, int ncol, const long long *ent_off,
                                const int *per_off, const double *periods, const int *nmode,
                                const float *colf, const long long *rows, uint16_t *X,
                                hipStream_t st) {
  long long n = nrow * 192;
  if (n <= 0) return;
  hipLaunchKernelGGL(jdsg_entry_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, e0,
                     nrow, ncol, ent_off, per_off, periods, nmode, colf, rows, X);
}

// pred[r] = out[r*ldo] * scale + shift  (MLP output column 0 to flat c)
__global__ void jdsg_denorm(long long n, const float *out, int ldo, double scale, double shift,
                            double *pred, long long e0) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pred[e0 + i] = (double)out[i * ldo] * scale + shift;
}
void jdsg_launch_denorm(long long n, const float *out, int ldo, double scale, double shift,
                        double *pred, long long e0, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(jdsg_denorm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, out,
                     ldo, scale, shift, pred, e0);
}
