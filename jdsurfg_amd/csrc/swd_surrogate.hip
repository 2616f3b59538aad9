// swd_surrogate.hip -- learned-bracket root finding.
//
// The MLP predicts the flat phase velocity of every (column, wave, mode,
// period).  Each root is then bracketed at +-half_width around the prediction
// (one period-equation evaluation per end; without a sign change the bracket
// is widened to three half widths, below and then above), accepted only if
// the period equation changes sign, and polished in FP64 by false position
// with the Anderson-Bjorck weight until the bracket is within tol*c, or the
// secant correction falls below tol*c right after one below 1e-4*c, or the
// error estimate of the corrected value (correction^2 / previous correction)
// falls below tol*c/10.  A polished root is kept only if it is the root the
// reference's chain would reach: the mode-order pass checks, with the signs
// already evaluated, that no odd number of roots lies between the previous
// mode's bracket (or, for the lowest mode, the chain's floor cm) and this
// bracket.  Roots that fail any check are left to the plain chain, which walks
// each chain in the reference order and solves only those (the `known` input
// of the chain kernel).  Every lane works on an independent root: no
// sequential chain.
#include <algorithm>

#include "swd_kernels.hpp"
#include "swd_dltar_fast.hpp"

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

// Persistent polish, one launch per wave type (W = 0 Rayleigh, 1 Love).  Each
// lane walks its items (item = variant-major, then the wave's entries in
// grid or list order) with a small state machine -- floor sign (lowest mode
// only), bracket low, bracket high, false-position steps -- so every trip of the
// main loop evaluates exactly one period equation per live lane, and a lane
// that finishes a root takes its next item at once instead of idling until
// the slowest lane of its wave is done.
namespace {
enum { PP_FLOOR = 0, PP_LO, PP_HI, PP_LO2, PP_HI2, PP_ILL };

struct PLane {
  long long e;          // entry
  int col, variant, phase, nit;
  double omega, cm, betmx, a, b, fa, fb, c;
  double fat;           // F(a) itself (fa carries the Anderson-Bjorck weight)
  double cp;            // the previous iterate
  double cen;           // the bracket's centre
  double a2, b2;        // the widened bracket ends (three half widths)
  double lo, hi;        // the bracket the root was taken from
  signed char s_floor, s_lo;
};

JDSG_DEV void polish_finish(const PolishArgs &A, const PLane &S, double root, signed char s_hi,
                            unsigned &nf) {
  long long q = (long long)S.variant * A.nent + S.e;
  A.known[q] = root;
  A.sg[q * 3 + 0] = S.s_lo;
  A.sg[q * 3 + 1] = s_hi;
  A.sg[q * 3 + 2] = S.s_floor;
  A.brk[2 * q] = S.lo;
  A.brk[2 * q + 1] = S.hi;
  if (root == 0.0) ++nf;
}

// Decode item `it` into the lane; false when `it` is past the last item.
// Items whose centre is unusable are finished here without an evaluation.
template <int W>
JDSG_DEV bool polish_begin(const PolishArgs &A, PLane &S, long long &it, long long stride_items,
                           unsigned &nf) {
  const long long total = A.nitem * A.nv;
  const int ncol = A.cm.ncol;
  for (; it < total; it += stride_items) {
    int v = A.v0 + (int)(it / A.nitem);
    long long i = it % A.nitem;
    int col, mode, k, K;
    long long e;
    if (!A.item_ent) {
      K = A.grid_K;
      col = (int)(i % ncol);
      long long r = i / ncol;
      k = (int)(r % K);
      mode = (int)(r / K);
      e = A.ent_off[col] + (long long)((W == 0 ? 0 : A.grid_nR) + mode) * K + k;
    } else {
      e = A.item_ent[i];
      int lo = 0, hi = ncol - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (A.ent_off[mid] <= e) lo = mid; else hi = mid - 1;
      }
      col = lo;
      K = A.per_off[col + 1] - A.per_off[col];
      long long el = e - A.ent_off[col];
      int wm = (int)(el / K);
      k = (int)(el % K);
      mode = (W == 0) ? wm : wm - A.nmode[2 * col];
    }
    double T = A.periods[A.per_off[col] + k];
    if (v == 1) T = T * (1.0 + 0.5 * 0.01);
    if (v == 2) T = T * (1.0 - 0.5 * 0.01);
    S.e = e;
    S.col = col;
    S.variant = v;
    S.omega = kTwoPi / T;
    S.cm = (double)A.cm.cc[col];
    S.betmx = (double)A.cm.betmx[col];
    S.s_floor = 0;
    S.s_lo = 0;
    S.nit = 0;
    S.cp = 0.0;
    double cen = A.centers[v][e];
    S.cen = cen;
    // a root above betmx is rejected, so a centre whose widest bracket starts
    // above betmx cannot give one: no evaluation is spent on it
    bool usable = cen > 0.0 && cen * (1.0 - 3.0 * A.hw[v]) < S.betmx;
    if (usable) {
      bracket_ends(cen, A.hw[v], S.cm, &S.a, &S.b);
      bracket_ends(cen, 3.0 * A.hw[v], S.cm, &S.a2, &S.b2);
      S.lo = S.a;
      S.hi = S.b;
    } else {
      S.lo = S.hi = 0.0;
    }
    // The floor sign of a T(1 +- 0.005) entry is its T entry's: cm lies
    // below every root at every period (the chain starts its scans there), so
    // F(cm, omega) has no zero between the three frequencies and keeps its
    // sign.  Only where the T pass left no sign is it evaluated again.
    const signed char s0 = (mode == 0 && v > 0) ? A.sg[e * 3 + 2] : (signed char)0;
    if (mode == 0 && s0 == 0) {
      S.phase = PP_FLOOR;
      S.c = S.cm;
      if (!usable) S.a = S.b = 0.0;  // floor sign still wanted by mode order
      return true;
    }
    if (mode == 0) {
      S.s_floor = s0;
      if (usable) {
        S.phase = PP_LO;
        S.c = S.a;
        return true;
      }
      polish_finish(A, S, 0.0, 0, nf);
      continue;
    }
    if (usable) {
      S.phase = PP_LO;
      S.c = S.a;
      return true;
    }
    long long q = (long long)v * A.nent + e;
    A.known[q] = 0.0;
    A.sg[q * 3 + 0] = 0;
    A.sg[q * 3 + 1] = 0;
    A.sg[q * 3 + 2] = 0;
    A.brk[2 * q] = 0.0;
    A.brk[2 * q + 1] = 0.0;
    ++nf;
  }
  return false;
}

// Feed F(S.c); true when the lane's root is finished (results written).
JDSG_DEV bool polish_step(const PolishArgs &A, PLane &S, double F, unsigned &nf) {
  switch (S.phase) {
    case PP_FLOOR:
      S.s_floor = (signed char)sign1(F);
      if (S.a > 0.0) {
        S.phase = PP_LO;
        S.c = S.a;
        return false;
      }
      polish_finish(A, S, 0.0, 0, nf);
      return true;
    case PP_LO:
      S.fa = F;
      S.fat = F;
      S.s_lo = (signed char)sign1(F);
      S.phase = PP_HI;
      S.c = S.b;
      return false;
    case PP_HI: {
      S.fb = F;
      if (sign1(S.fa) == sign1(S.fb)) {
        // no sign change: widen to three half widths below, then above
        // (not past betmx, where a root would be rejected anyway)
        if (S.a2 < S.a) {
          S.phase = PP_LO2;
          S.c = S.a2;
          return false;
        }
        if (S.b < S.betmx) {
          S.phase = PP_HI2;
          S.c = S.b2;
          return false;
        }
        polish_finish(A, S, 0.0, (signed char)sign1(F), nf);
        return true;
      }
      S.phase = PP_ILL;
      break;
    }
    case PP_LO2:
      if (sign1(F) != sign1(S.fa)) {  // root in [a2, a]
        S.b = S.a;
        S.fb = S.fat;
        S.a = S.a2;
        S.fa = F;
        S.fat = F;
        S.lo = S.a2;
        S.hi = S.b;
        S.s_lo = (signed char)sign1(F);
        S.phase = PP_ILL;
        break;
      }
      if (S.b < S.betmx) {
        S.phase = PP_HI2;
        S.c = S.b2;
        return false;
      }
      S.lo = S.a2;
      polish_finish(A, S, 0.0, (signed char)sign1(S.fb), nf);
      return true;
    case PP_HI2:
      if (sign1(F) != sign1(S.fb)) {  // root in [b, b2]
        S.a = S.b;
        S.fa = S.fb;
        S.fat = S.fb;
        S.b = S.b2;
        S.fb = F;
        S.lo = S.a;
        S.hi = S.b2;
        S.s_lo = (signed char)sign1(S.fa);
        S.phase = PP_ILL;
        break;
      }
      polish_finish(A, S, 0.0, (signed char)sign1(F), nf);
      return true;
    default: {  // false position step with F = F(S.c)
      double fc = F;
      double xo = 0.0, fo = 0.0;  // the point that stops being an end
      if (fc == 0.0) {
        S.a = S.b = S.c;
      } else {
        if (sign1(fc) != sign1(S.fb)) {
          xo = S.a;
          fo = S.fat;
          S.a = S.b;
          S.fa = S.fb;
          S.fat = S.fb;
        } else {
          xo = S.b;
          fo = S.fb;
          // the retained end keeps its place: Anderson-Bjorck scales its
          // value by 1 - fc/fb (0.5, the Illinois factor, when that is <= 0)
          double m = 1.0 - fc / S.fb;
          S.fa = ((m > 0.0) ? m : 0.5) * S.fa;
        }
        S.b = S.c;
        S.fb = fc;
      }
      ++S.nit;
      bool done = fc == 0.0 || fabs(S.b - S.a) <= A.tol * fabs(S.b) || S.nit >= 80;
      double root = S.b;
      if (!done) {
        double cn = (S.a * S.fb - S.b * S.fa) / (S.fb - S.fa);
        if (!(cn > fmin(S.a, S.b) && cn < fmax(S.a, S.b))) cn = 0.5 * (S.a + S.b);
        // inverse quadratic interpolation through F at the two ends and at
        // the point just dropped (as in Brent's method), taken when it falls
        // inside the bracket: from a close centre or a first interpolation
        // it reaches the tolerance one evaluation earlier than false position
        if (fo != S.fat && fo != S.fb && S.fat != S.fb) {
          const double x = xo * S.fat * S.fb / ((fo - S.fat) * (fo - S.fb)) +
                           S.a * fo * S.fb / ((S.fat - fo) * (S.fat - S.fb)) +
                           S.b * fo * S.fat / ((S.fb - fo) * (S.fb - S.fat));
          if (x > fmin(S.a, S.b) && x < fmax(S.a, S.b)) cn = x;
        }
        // the corrected value's error is about dc^2 / dc_prev for a linearly
        // converging step and less once convergence is superlinear: accept
        // it when that estimate is a tenth of the tolerance
        const double d1 = fabs(cn - S.c), d0 = fabs(S.c - S.cp), cc = fabs(S.c);
        if ((d1 <= A.tol * cc && d0 <= 1.0e-4 * cc) ||
            (S.cp != 0.0 && d0 <= 1.0e-3 * cc && d1 * d1 <= 0.1 * A.tol * cc * d0)) {
          done = true;
          root = cn;
        } else {
          S.cp = S.c;
          S.c = cn;
          return false;
        }
      }
      if (!(root > S.cm) || root > S.betmx) root = 0.0;
      // the bracket changed sign, so F at its top had the sign opposite to
      // its bottom: that is the sign the mode-order pass compares
      polish_finish(A, S, root, (signed char)(-S.s_lo), nf);
      return true;
    }
  }
  double c = (S.a * S.fb - S.b * S.fa) / (S.fb - S.fa);
  if (!(c > fmin(S.a, S.b) && c < fmax(S.a, S.b))) c = 0.5 * (S.a + S.b);
  // A T(1 +- 0.005) centre, the polished T root moved along the slope of
  // the neighbouring periods' roots, lies about 1e-6 from its root, much
  // closer than the interpolation between the ends of its bracket (which is
  // kept as wide as the chain's scan step, so that a close pair of roots
  // shows as no sign change, as the chain sees it): false position starts
  // at the centre when it lies inside the bracket (not after a widening)
  if (S.variant > 0 && S.cen > fmin(S.a, S.b) && S.cen < fmax(S.a, S.b)) c = S.cen;
  S.c = c;
  return false;
}
}  // namespace

// Love lanes fit three waves per SIMD (168 VGPRs, a 32-byte spill); the
// Rayleigh ones need 214 and slowed down when capped (192 bytes of spill)
template <int W>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W == 1 ? 3 : 1)))
jdsg_polish_w(PolishArgs A) {
  const long long nl = (long long)gridDim.x * blockDim.x;
  long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long nev = 0;
  unsigned nf = 0;
  PLane S;
  bool live = polish_begin<W>(A, S, it, nl, nf);
  const int stride = A.cm.ncol_pad;
  while (__any(live)) {
    if (live) {
      const float4 *st = (W == 0 ? A.cm.chainR : A.cm.chainL) + S.col;
      int mmax = A.cm.nlay[S.col], llw = A.cm.llw[S.col];
      double wv = S.omega / S.c;
      double F = (W == 0) ? jdsg_pf::dltar4(st, stride, mmax, llw, wv, S.omega)
                          : jdsg_pf::dltar1(st, stride, mmax, llw, wv, S.omega);
      ++nev;
      if (polish_step(A, S, F, nf)) {
        it += nl;
        live = polish_begin<W>(A, S, it, nl, nf);
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    nev += __shfl_xor(nev, o);
    nf += __shfl_xor(nf, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (nf) atomicAdd(A.nfail, nf);
    if (A.nevals && nev) atomicAdd(A.nevals, nev);
  }
}

// Continuation of open roots from their polished neighbours, one lane per
// (column, mode) of one wave type and variant.  An entry left open (its
// prediction was off, or a check failed) whose neighbouring period's root of
// the same mode is known is bracketed around that root moved by the
// network's ratio between the two periods -- the chain's own start, the
// previous period's root -- and polished like any other entry; the lane
// walks each run of open entries up the periods, then down, so a run is
// continued entry by entry.  Only entries whose own prediction lies with
// its whole widened bracket below betmx are continued (where the network
// sees no root, or one near the cut-off, the chain may miss one that exists,
// and a continued root would then differ from the chain's), and a continued
// root is kept only within 5 % of that prediction.
// The mode-order pass checks these roots as it checks the others.  Mode 0
// reuses the floor sign of the first polish.
template <int W>
JDSG_DEV bool continue_begin(const PolishArgs &A, PLane &S, long long e, int col, int mode, int k,
                             int v, double cen) {
  double T = A.periods[A.per_off[col] + k];
  if (v == 1) T = T * (1.0 + 0.5 * 0.01);
  if (v == 2) T = T * (1.0 - 0.5 * 0.01);
  S.e = e;
  S.col = col;
  S.variant = v;
  S.omega = kTwoPi / T;
  S.cm = (double)A.cm.cc[col];
  S.betmx = (double)A.cm.betmx[col];
  S.nit = 0;
  S.cp = 0.0;
  S.s_lo = 0;
  S.cen = cen;
  if (!(cen > 0.0 && cen * (1.0 - 3.0 * A.hw[v]) < S.betmx)) return false;
  bracket_ends(cen, A.hw[v], S.cm, &S.a, &S.b);
  bracket_ends(cen, 3.0 * A.hw[v], S.cm, &S.a2, &S.b2);
  S.lo = S.a;
  S.hi = S.b;
  const long long q = (long long)v * A.nent + e;
  S.s_floor = (mode == 0) ? A.sg[q * 3 + 2] : 0;
  if (mode == 0 && S.s_floor == 0) return false;  // no floor sign from the first pass
  S.phase = PP_LO;
  S.c = S.a;
  return true;
}

// COOP: one (column, mode) walk per wave: every lane holds the same lane
// state and the wave forms each period equation together (the one-lane
// value in a fraction of its latency); the walks are few and long.
template <int W, bool COOP>
__global__ void __launch_bounds__(64) jdsg_continue_w(PolishArgs A, int v, int mmax_modes) {
  __shared__ double cf[COOP ? 64 * (W == 0 ? jdsg_pf::kCoef4 : jdsg_pf::kCoef1) : 1];
  const long long t =
      COOP ? (long long)blockIdx.x : (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ncol = A.cm.ncol;
  if (t >= (long long)ncol * mmax_modes) return;
  const int col = (int)(t / mmax_modes), m = (int)(t % mmax_modes);
  const int nR = A.nmode[2 * col], nL = A.nmode[2 * col + 1];
  if (m >= (W == 0 ? nR : nL)) return;
  const int K = A.per_off[col + 1] - A.per_off[col];
  const long long e0 = A.ent_off[col] + (long long)((W == 0 ? 0 : nR) + m) * K;
  const double *known = A.known + (long long)v * A.nent;
  const double *pred = A.centers[0];
  const float4 *st = (W == 0 ? A.cm.chainR : A.cm.chainL) + col;
  const int mmax = A.cm.nlay[col], llw = A.cm.llw[col], stride = A.cm.ncol_pad;
  unsigned long long nev = 0;
  unsigned nf = 0;
  if (COOP) {
    // the walk continues only open entries whose prediction lies, with its
    // widened bracket, below betmx: the lanes look for one at once, and a
    // walk without any ends here instead of visiting every period
    bool cand = false;
    for (int k = threadIdx.x; k < K; k += 64) {
      const double p1 = pred[e0 + k];
      cand = cand || (known[e0 + k] == 0.0 && p1 > 0.0 &&
                      p1 * (1.0 + 3.0 * A.hw[0]) < (double)A.cm.betmx[col]);
    }
    if (!__any(cand)) return;
  }
  for (int dir = 0; dir < 2; ++dir) {
    for (int i = 1; i < K; ++i) {
      const int k = dir == 0 ? i : K - 1 - i;
      const int kn = dir == 0 ? k - 1 : k + 1;
      const double cn = known[e0 + kn];
      if (known[e0 + k] != 0.0 || !(cn > 0.0)) continue;
      const double p1 = pred[e0 + k], p0 = pred[e0 + kn];
      if (!(p1 > 0.0 && p0 > 0.0) || p1 * (1.0 + 3.0 * A.hw[0]) >= (double)A.cm.betmx[col])
        continue;
      const double cen = cn * (p1 / p0);
      PLane S;
      if (!continue_begin<W>(A, S, e0 + k, col, m, k, v, cen)) continue;
      for (;;) {
        const double wv = S.omega / S.c;
        double F;
        if (COOP)
          F = (W == 0) ? jdsg_pf::dltar4_wave(st, stride, mmax, llw, wv, S.omega, cf)
                       : jdsg_pf::dltar1_wave(st, stride, mmax, llw, wv, S.omega, cf);
        else
          F = (W == 0) ? jdsg_pf::dltar4(st, stride, mmax, llw, wv, S.omega)
                       : jdsg_pf::dltar1(st, stride, mmax, llw, wv, S.omega);
        ++nev;
        if (polish_step(A, S, F, nf)) break;
      }
      double *kp = A.known + (long long)v * A.nent + e0 + k;
      if (*kp > 0.0 && fabs(*kp - p1) > 0.05 * p1) *kp = 0.0;
    }
  }
  if ((!COOP || threadIdx.x == 0) && A.nevals && nev) atomicAdd(A.nevals, nev);
}

// Entries still open after the surrogate stages (known = 0): the roots the
// chain has to walk to.
__global__ void __launch_bounds__(256)
jdsg_count_open(const double *known, long long n, unsigned *out) {
  __shared__ unsigned part[4];
  unsigned c = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    c += !(known[i] > 0.0);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned s = part[0] + part[1] + part[2] + part[3];
    if (s) atomicAdd(out, s);
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
  // period fastest: a wave reads consecutive entries (e = ... + mode*K + k)
  int k = (int)(t % kmax);
  long long r = t / kmax;
  int w = (int)(r & 1);
  int col = (int)(r >> 1);
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
      double a = A.brk[2 * q], b = A.brk[2 * q + 1];
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
  // d ln c / d ln T from the neighbouring periods' polished roots (exact, so
  // only the chord error is left), else from their predictions
  double slope = 0.0;
  int ka = k > 0 ? k - 1 : k, kb = k < K - 1 ? k + 1 : k;
  if (kb > ka) {
    double ka0 = known0[e - (k - ka)], kb0 = known0[e + (kb - k)];
    double pa = pred[e - (k - ka)], pb = pred[e + (kb - k)];
    if (ka0 > 0.0 && kb0 > 0.0)
      slope = log(kb0 / ka0) / log(T[kb] / T[ka]);
    else if (pa > 0.0 && pb > 0.0)
      slope = log(pb / pa) / log(T[kb] / T[ka]);
  }
  double l1 = log(1.0 + 0.5 * 0.01), l2 = log(1.0 - 0.5 * 0.01);
  cen1[e] = c0 * exp(slope * l1);
  cen2[e] = c0 * exp(slope * l2);
}

// Per-column MLP features: Vs, Vp and rho averaged (thickness-weighted) over
// 48 depth bins with graded edges -- 8 of 0.5 km to 4 km, 16 of 1 km to 20 km,
// 24 of 3 km to 92 km; below the stack the half-space -- so thin soft layers
// near the surface still show, then normalised; plus the half-space Vs, Vp,
// rho and the depth of its top.  colf row: [0,144) bins, [144,148) extras.
// Entry features append ln T, the wave and the mode one-hot.
// X row layout (ld = 192 bf16): [0,144) profile, 144 ln T, 145-146 wave,
// 147-150 mode, 151-154 half-space, rest zero.
constexpr int kColF = 148;

__device__ __forceinline__ float bin_top(int j) {
  return j < 8 ? 0.5f * j : (j < 24 ? 4.0f + (j - 8) : 20.0f + 3.0f * (j - 24));
}

extern "C" __global__ void jdsg_column_features(int ncol, const int *layer_off, const float *thk,
                                                const float *vp, const float *vs,
                                                const float *rho, float *colf) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  int l0 = layer_off[col], N = layer_off[col + 1] - l0;
  float *out = colf + (size_t)col * kColF;
  int i = 0;
  float ltop = 0.f;  // top of layer i
  for (int j = 0; j < 48; ++j) {
    float z0 = bin_top(j), z1 = bin_top(j + 1);
    float as = 0.f, ap = 0.f, ar = 0.f;
    float z = z0;
    while (z < z1) {
      float lbot = (i < N - 1) ? ltop + thk[l0 + i] : 3.0e38f;
      float seg = fminf(lbot, z1) - z;
      if (seg > 0.f) {
        as += seg * vs[l0 + i];
        ap += seg * vp[l0 + i];
        ar += seg * rho[l0 + i];
        z += seg;
      }
      if (lbot <= z1 && i < N - 1) {
        ltop = lbot;
        ++i;
      } else {
        break;
      }
    }
    float inv = 1.0f / (z1 - z0);
    out[j] = (as * inv - 3.5f) / 0.8f;
    out[48 + j] = (ap * inv - 6.2f) / 1.2f;
    out[96 + j] = (ar * inv - 2.8f) / 0.4f;
  }
  float zh = 0.f;
  for (int m = 0; m < N - 1; ++m) zh += thk[l0 + m];
  int h = l0 + N - 1;
  out[144] = (vs[h] - 4.5f) / 0.5f;
  out[145] = (vp[h] - 8.0f) / 1.0f;
  out[146] = (rho[h] - 3.3f) / 0.3f;
  out[147] = zh / 100.0f - 1.0f;
}

__device__ __forceinline__ uint16_t f2bf_(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// One thread per entry row: decodes (column, wave, mode, period) once and
// writes the 192 bf16 features of the row as 24 16-byte stores.
extern "C" __global__ void jdsg_entry_features(long long e0, long long nrow, int ncol,
                                               const long long *ent_off, const int *per_off,
                                               const double *periods, const int *nmode,
                                               const float *colf, const long long *rows,
                                               uint16_t *X) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrow) return;
  long long e = rows ? rows[t] : e0 + t;
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
  const float *cf = colf + (size_t)col * kColF;
  float lnT = (float)((log(periods[per_off[col] + k]) - 2.0) / 1.5);
  uint4 *dst = reinterpret_cast<uint4 *>(X + (size_t)t * 192);
  for (int g = 0; g < 24; ++g) {
    uint32_t w[4];
    for (int h = 0; h < 4; ++h) {
      uint32_t pr[2];
      for (int q = 0; q < 2; ++q) {
        int f = g * 8 + h * 2 + q;
        float v = 0.f;
        if (f < 144) v = cf[f];
        else if (f == 144) v = lnT;
        else if (f == 145) v = wave == 0 ? 1.f : 0.f;
        else if (f == 146) v = wave == 1 ? 1.f : 0.f;
        else if (f < 151) v = (mode == f - 147) ? 1.f : 0.f;
        else if (f < 155) v = cf[144 + (f - 151)];
        pr[q] = f2bf_(v);
      }
      w[h] = pr[0] | (pr[1] << 16);
    }
    dst[g] = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

void jdsg_launch_polish(const PolishArgs &A, int wave, int nwaves, hipStream_t st) {
  long long items = A.nitem * A.nv;
  if (items <= 0) return;
  long long need = (items + 63) / 64;
  unsigned g = (unsigned)(need < nwaves ? need : nwaves);
  if (wave == 0)
    hipLaunchKernelGGL(jdsg_polish_w<0>, dim3(g), dim3(64), 0, st, A);
  else
    hipLaunchKernelGGL(jdsg_polish_w<1>, dim3(g), dim3(64), 0, st, A);
}
void jdsg_launch_continue(const PolishArgs &A, int wave, int v, int mmax_modes, bool coop,
                          hipStream_t st) {
  const long long n = (long long)A.cm.ncol * mmax_modes;
  if (n <= 0) return;
  if (coop) {
    const dim3 g((unsigned)n);
    if (wave == 0)
      hipLaunchKernelGGL((jdsg_continue_w<0, true>), g, dim3(64), 0, st, A, v, mmax_modes);
    else
      hipLaunchKernelGGL((jdsg_continue_w<1, true>), g, dim3(64), 0, st, A, v, mmax_modes);
    return;
  }
  const dim3 g((unsigned)((n + 63) / 64));
  if (wave == 0)
    hipLaunchKernelGGL((jdsg_continue_w<0, false>), g, dim3(64), 0, st, A, v, mmax_modes);
  else
    hipLaunchKernelGGL((jdsg_continue_w<1, false>), g, dim3(64), 0, st, A, v, mmax_modes);
}
void jdsg_launch_count_open(const double *known, long long n, unsigned *out, hipStream_t st) {
  if (n <= 0) return;
  const long long blocks = std::min<long long>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(jdsg_count_open, dim3((unsigned)blocks), dim3(256), 0, st, known, n, out);
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
  long long n = nrow;
  if (n <= 0) return;
  hipLaunchKernelGGL(jdsg_entry_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, e0,
                     nrow, ncol, ent_off, per_off, periods, nmode, colf, rows, X);
}

// pred[r] = out[r*ldo] * scale + shift  (MLP output column 0 to flat c)
__global__ void jdsg_denorm(long long n, const float *out, int ldo, double scale, double shift,
                            double *pred, long long e0) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pred[e0 + i] = exp((double)out[i * ldo] * scale + shift);
}
void jdsg_launch_denorm(long long n, const float *out, int ldo, double scale, double shift,
                        double *pred, long long e0, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(jdsg_denorm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, out,
                     ldo, scale, shift, pred, e0);
}
