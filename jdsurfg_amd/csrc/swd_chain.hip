// swd_chain.hip -- per-column model preparation and the period-equation root
// chains (the CDNA4 form of src/SWD/surfdisp96.f).
//
// One lane owns one chain: a (column, wave type, period scale) triple whose
// modes are solved in the reference's sequential order (period k starts from
// the root at k-1, mode iq is floored by mode iq-1, surfdisp96.f:262-281).
// The chain runs as an explicit state machine that stops exactly where the
// reference calls dltar (getsol scan, nevill halving / Neville steps), so
// every loop trip of the kernel is one period-equation evaluation for every
// live lane: lanes in different phases never serialise the layer recursion.
// Lanes whose chain is finished fetch the next chain from a global counter.
//
// Arithmetic follows the Fortran: float model arrays and start values,
// double recursion, sngl() of the returned root.  FP contraction is off so
// the recursion rounds like the reference build.
#include <algorithm>
#include <type_traits>

#include "swd_kernels.hpp"
#include "swd_dltar.hpp"
#include "swd_dltar_fast.hpp"

using jdsg_pe::dltar1;
using jdsg_pe::dltar4;

#pragma clang fp contract(off)

namespace {

constexpr double kTwoPi = 2.0 * 3.141592653589793;

JDSG_DEV float powm5f(float x) {
  float x2 = x * x, x4 = x2 * x2, x5 = x4 * x;
  return 1.0f / x5;
}
JDSG_DEV double powm5d(double x) {
  double x2 = x * x, x4 = x2 * x2, x5 = x4 * x;
  return 1.0 / x5;
}

// surfdisp96.f:377-398 gtsolh, single precision
JDSG_DEV float gtsolh(float a, float b) {
  float c = 0.95f * b;
  for (int i = 0; i < 5; ++i) {
    float gamma = b / a;
    float kappa = c / b;
    float k2 = kappa * kappa;
    float gk = gamma * kappa;
    float gk2 = gk * gk;
    float fac1 = sqrtf(1.0f - gk2);
    float fac2 = sqrtf(1.0f - k2);
    float tk = 2.0f - k2;
    float fr = tk * tk - 4.0f * fac1 * fac2;
    float frp = -(4.0f * (2.0f - k2) * kappa) + 4.0f * fac2 * gamma * gamma * kappa / fac1 +
                4.0f * fac1 * kappa / fac2;
    frp = frp / b;
    c = c - fr / frp;
  }
  return c;
}

}  // namespace

// ----------------------------------------------------------------------------
// Column preparation: one thread per column.
//   * surfdisp96 `sphere` (surfdisp96.f:499-571, R = 6370 km, float storage)
//     and the starting value / betmx scan (surfdisp96.f:148-221);
//   * sregn96/slegn96 `bldsph` factors (R = 6371 km, double).


extern "C" __global__ void __launch_bounds__(256)
jdsg_prep_columns(int ncol, int ncol_pad, const int *__restrict__ layer_off,
                  const float *__restrict__ thk, const float *__restrict__ vp,
                  const float *__restrict__ vs, const float *__restrict__ rho,
                  int iflsph, PrepOut o) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  int l0 = layer_off[col];
  int N = layer_off[col + 1] - l0;
  o.nlay[col] = N;
  o.llw[col] = (vs[l0] <= 0.0f) ? 2 : 1;
  // ---- root-chain model
  {
    const double ar = 6370.0;
    double dr = 0.0, r0 = ar;
    float betmn = 1.e20f, betmx = -1.e20f;
    int jmn = 0, jsol = 1;
    float amn = 0.f, bmn = 0.f;
    for (int i = 0; i < N; ++i) {
      float d = thk[l0 + i], a = vp[l0 + i], b = vs[l0 + i], rh = rho[l0 + i];
      float rR = rh, rL = rh;
      if (iflsph == 1) {
        float di = (i == N - 1) ? 1.0f : d;
        dr = dr + (double)di;
        double r1 = ar - dr;
        double z0 = ar * log(ar / r0);
        double z1 = ar * log(ar / r1);
        d = (float)(z1 - z0);
        double tmp = (ar + ar) / (r0 + r1);
        a = (float)(a * tmp);
        b = (float)(b * tmp);
        float btp = (float)tmp;
        r0 = r1;
        rL = rh * powm5f(btp);
        // REAL**REAL: the float power rounded from a double evaluation, as
        // the host libm's powf delivers it (correctly rounded in practice)
        rR = rh * (float)pow((double)btp, (double)(-2.275f));
      }
      if (i == N - 1) d = 0.0f;
      size_t q = (size_t)i * ncol_pad + col;
      o.chainR[q] = make_float4(d, a, b, rR);
      o.chainL[q] = make_float4(d, a, b, rL);
      if (b > 0.01f && b < betmn) {
        betmn = b; jmn = i; jsol = 1; amn = a; bmn = b;
      } else if (b <= 0.01f && a < betmn) {
        betmn = a; jmn = i; jsol = 0;
      }
      if (b > betmx) betmx = b;
    }
    (void)jmn;
    float cc1 = (jsol == 0) ? betmn : gtsolh(amn, bmn);
    cc1 = 0.95f * cc1;
    cc1 = 0.90f * cc1;
    o.cc[col] = cc1;
    o.betmx[col] = betmx;
  }
  // ---- eigen-stage model (sregn96.f90:130-181, slegn96.f90:107-151)
  {
    const double ar = 6371.0;
    double dr = 0.0, r0 = ar;
    for (int i = 0; i < N; ++i) {
      size_t q = (size_t)i * ncol_pad + col;
      double zd = (double)thk[l0 + i];
      double za = (double)vp[l0 + i], zb = (double)vs[l0 + i];
      double vtp = 1.0, dtp = 1.0, rR = 1.0, rL = 1.0;
      if (iflsph > 0) {
        dr = dr + ((i == N - 1) ? 1.0 : zd);
        double r1 = ar - dr;
        double z0 = ar * log(ar / r0);
        double z1 = ar * log(ar / r1);
        double tmp = (2.0 * ar) / (r0 + r1);
        vtp = tmp;
        rR = pow(tmp, (double)(-2.275f));
        rL = powm5d(tmp);
        dtp = ar / r0;
        za = za * tmp;
        zb = zb * tmp;
        zd = z1 - z0;
        r0 = r1;
      }
      if (i == N - 1) zd = 0.0;
      o.zd[q] = zd;
      o.za[q] = za;
      o.zb[q] = zb;
      o.rho0[q] = (double)rho[l0 + i];
      o.rtpR[q] = rR;
      o.rtpL[q] = rL;
      o.vtp[q] = vtp;
      o.dtp[q] = dtp;
    }
  }
}

namespace {

// ----------------------------------------------------------------------------
// Chain state machine.
enum Phase : int {
  PH_START_K = 0,
  PH_GS_FIRST,
  PH_SCAN_NEXT,
  PH_GS_SCAN,
  PH_NV_INIT,
  PH_NV_LOOP,
  PH_NV_TOPHALF,
  PH_NV_BODY,
  PH_NV_END,
  PH_RETRY_NEXT,
  PH_DEL1ST,
  PH_DONE,
  PH_SUSPEND,   // a long upward scan handed to a scan job; the chain is saved
  PH_SCAN_HIT,  // resumed: the scan job found the sign change
  PH_SCAN_END,  // resumed: the scan job ran to its end without one
  PH_FP         // exact-root mode: false-position polish of the bracket
};

struct ChainState {
  // job
  int col, wave, variant, K, M;
  long long e0;        // entry index of (mode 0, period 0) of this wave
  double tscale;
  // constants
  double cc, cm, dc, onea, one;
  float betmx;
  // driver
  int iq, k, ift, ierr, iverb, sub, Mrun, Krun, kbase, retry;
  unsigned need, exitmask;
  double t1, omega;
  float t1a, t1b;
  // getsol
  double c1, c2, clow, del1, del2, del1st;
  int idir, ifirst;
  // nevill
  double c3, del3;
  double del1t;  // exact-root polish: F(c1) itself (del1 carries the weight)
  int nev, nctrl, m;
  int phase;
  double ec;  // requested evaluation point
  // known-root walk: start point of the mode's first period (for del1st)
  double tfirst, cfirst;
  int del1st_ok;
};
// state of a chain in the suspension form (exact-root fallback)
struct ChainStateSusp : ChainState {
  int nscan;        // evaluations of the current upward scan
  unsigned evals;   // evaluations of this chain
};



// Exact-root mode polishes a bracket [c1, c2] (del1, del2 of opposite sign)
// by false position with the Anderson-Bjorck weight (and the inverse quadratic
// step where it falls inside) instead of nevill's
// Neville/bisection hybrid, which bisects most of the way down to a 1e-10
// bracket (about 25 evaluations a root against 5-8).  The iterate is c3, the
// previous one del3 (0 before the second), the step count nctrl.
JDSG_DEV void fp_begin(ChainState &S) {
  S.nctrl = 0;
  S.del3 = 0.0;
  S.del1t = S.del1;
  double c = (S.c1 * S.del2 - S.c2 * S.del1) / (S.del2 - S.del1);
  if (!(c > fmin(S.c1, S.c2) && c < fmax(S.c1, S.c2))) c = 0.5 * (S.c1 + S.c2);
  S.c3 = c;
  S.ec = c;
  S.phase = PH_FP;
}

JDSG_DEV double period_at(const ChainArgs &A, const ChainState &S, int kk) {
  double T = A.periods[A.per_off[S.col] + kk];
  return (S.variant == 0) ? T : T * S.tscale;
}

JDSG_DEV void write_out(const ChainArgs &A, const ChainState &S, int mode, int kk, double v) {
  long long e = S.e0 + (long long)mode * S.K + kk;
  A.cflat[(long long)S.variant * A.nent + e] = v;
}
JDSG_DEV double read_out(const ChainArgs &A, const ChainState &S, int mode, int kk) {
  long long e = S.e0 + (long long)mode * S.K + kk;
  return A.cflat[(long long)S.variant * A.nent + e];
}

// Start (or restart) the chain proper over Krun periods from kbase.
JDSG_DEV void begin_pass(ChainState &S, double *cs, int kpad) {
  S.iq = 1;
  S.k = 0;
  S.ift = 998;
  S.ierr = 0;
  S.iverb = 0;
  for (int i = 0; i < S.Krun; ++i) {
    cs[i] = 0.0;
    cs[kpad + i] = 0.0;
  }
  S.phase = PH_START_K;
}

// Returns true with S.ec set when a period-equation value is needed at
// phase velocity S.ec (period S.t1); false when the chain is complete.
// EXACT: exact-root mode (tolerance below 1e-6), brackets polished by
// false position.
template <bool SUSP, bool EXACT>
__device__ __attribute__((noinline)) bool advance(
    typename std::conditional<SUSP, ChainStateSusp, ChainState>::type &S, const ChainArgs &A,
    double *cs, double F) {
  const int kpad = A.kmax_pad;
  for (;;) {
    switch (S.phase) {
      case PH_START_K: {
        if (S.iq > S.Mrun) {
          if (!S.retry) {
            if (A.retry && S.ierr) {
              S.retry = 1;
              S.kbase = -1;
              S.phase = PH_RETRY_NEXT;
              continue;
            }
            return false;
          }
          // end of one single-period retry pass
          S.phase = PH_RETRY_NEXT;
          if (S.ierr) S.exitmask |= S.need;
          continue;
        }
        if (S.k >= S.Krun) {
          S.iq++;
          S.k = 0;
          continue;
        }
        int k0 = S.k;
        if (k0 >= S.ift) goto fail_k;
        {
          double t1 = period_at(A, S, S.kbase + k0);
          if (A.igr > 0) {
            const float h = 0.005f;
            S.t1a = (float)(t1 / (double)(1.0f + h));
            S.t1b = (float)(t1 / (double)(1.0f - h));
            t1 = (double)S.t1a;
          } else {
            S.t1a = (float)t1;
          }
          S.t1 = t1;
          if (k0 == 0 && S.iq == 1) {
            S.c1 = S.cc; S.clow = S.cc; S.ifirst = 1;
          } else if (k0 == 0 && S.iq > 1) {
            S.c1 = cs[0] + S.one * S.dc; S.clow = S.c1; S.ifirst = 1;
          } else if (k0 > 0 && S.iq > 1) {
            S.ifirst = 0;
            S.clow = cs[k0] + S.one * S.dc;
            S.c1 = cs[k0 - 1];
            if (S.c1 < S.clow) S.c1 = S.clow;
          } else {
            S.ifirst = 0;
            S.c1 = cs[k0 - 1] - S.onea * S.dc;
            S.clow = S.cm;
          }
          S.sub = 1;
          if (k0 == 0) {
            S.tfirst = S.t1;
            S.cfirst = S.c1;
            S.del1st_ok = 0;
          }
          if (A.known && !S.retry && A.igr == 0) {
            long long e = S.e0 + (long long)(S.iq - 1) * S.K + k0;
            double kr = A.known[(long long)S.variant * A.nent + e];
            if (kr > 0.0) {  // root already bracketed and polished
              S.c3 = kr;
              goto nv_done;
            }
            if (S.ifirst == 0 && !S.del1st_ok) {
              // the polarity reference of this mode's first period was skipped
              S.omega = kTwoPi / S.tfirst;
              S.ec = S.cfirst;
              S.phase = PH_DEL1ST;
              return true;
            }
          }
          S.omega = kTwoPi / S.t1;
          S.ec = S.c1;
          S.phase = PH_GS_FIRST;
          return true;
        }
      }
      case PH_DEL1ST: {
        S.del1st = F;
        S.del1st_ok = 1;
        S.omega = kTwoPi / S.t1;
        S.ec = S.c1;
        S.phase = PH_GS_FIRST;
        return true;
      }
      case PH_GS_FIRST: {
        S.del1 = F;
        if constexpr (SUSP) S.nscan = 0;
        if (S.ifirst == 1) {
          S.del1st = S.del1;
          S.del1st_ok = 1;
        }
        double plmn = sign1(S.del1st) * sign1(S.del1);
        S.idir = (S.ifirst == 1 || plmn >= 0.0) ? 1 : -1;
        S.phase = PH_SCAN_NEXT;
        continue;
      }
      case PH_SCAN_NEXT: {
        // exact-root fallback: an upward scan that has gone min_steps points
        // without a sign change and has more than 64 to go before betmx is
        // handed to a scan job, which evaluates the same points side by side
        if constexpr (SUSP) {
          if (A.susp.allow && S.idir > 0 && S.nscan >= A.susp.min_steps && S.c1 >= S.clow &&
              S.c1 >= S.cm && ((double)S.betmx + S.dc) - S.c1 > 64.0 * S.dc) {
            S.phase = PH_SUSPEND;
            return false;
          }
        }
        S.c2 = (S.idir > 0) ? S.c1 + S.dc : S.c1 - S.dc;
        if (S.c2 <= S.clow) {
          S.idir = 1;
          S.c1 = S.clow;
          continue;
        }
        S.ec = S.c2;
        S.phase = PH_GS_SCAN;
        return true;
      }
      case PH_GS_SCAN: {
        S.del2 = F;
        if constexpr (SUSP) S.nscan++;
        if (sign1(S.del1) != sign1(S.del2)) {
          if constexpr (EXACT) {
            fp_begin(S);
            return true;
          }
          S.c3 = 0.5 * (S.c1 + S.c2);
          S.ec = S.c3;
          S.phase = PH_NV_INIT;
          return true;
        }
        S.c1 = S.c2;
        S.del1 = S.del2;
        if (S.c1 < S.cm || S.c1 >= ((double)S.betmx + S.dc)) goto getsol_fail;
        S.phase = PH_SCAN_NEXT;
        continue;
      }
      case PH_SCAN_HIT:  // as PH_GS_SCAN on a sign change between c1 and c2
        if constexpr (SUSP) {
          if constexpr (EXACT) {
            fp_begin(S);
            return true;
          }
          S.c3 = 0.5 * (S.c1 + S.c2);
          S.ec = S.c3;
          S.phase = PH_NV_INIT;
          return true;
        } else {
          return false;
        }
      case PH_SCAN_END:  // as PH_GS_SCAN past betmx: c1 is the scan's last point
        if constexpr (SUSP) goto getsol_fail;
        else return false;
      case PH_NV_INIT: {
        S.del3 = F;
        S.nev = 1;
        S.nctrl = 1;
        S.m = 1;
        S.phase = PH_NV_LOOP;
        continue;
      }
      case PH_NV_LOOP: {
        S.nctrl++;
        if (S.nctrl >= 100) goto nv_done;
        if (S.c3 < fmin(S.c1, S.c2) || S.c3 > fmax(S.c1, S.c2)) {
          S.nev = 0;
          S.c3 = 0.5 * (S.c1 + S.c2);
          S.ec = S.c3;
          S.phase = PH_NV_TOPHALF;
          return true;
        }
        S.phase = PH_NV_BODY;
        continue;
      }
      case PH_NV_TOPHALF: {
        S.del3 = F;
        S.phase = PH_NV_BODY;
        continue;
      }
      case PH_NV_BODY: {
        double s13 = S.del1 - S.del3;
        double s32 = S.del3 - S.del2;
        if (sign1(S.del3) * sign1(S.del1) < 0.0) {
          S.c2 = S.c3;
          S.del2 = S.del3;
        } else {
          S.c1 = S.c3;
          S.del1 = S.del3;
        }
        if (fabs(S.c1 - S.c2) <= A.tol * S.c1) goto nv_done;
        if (sign1(s13) != sign1(s32)) S.nev = 0;
        double ss1 = fabs(S.del1);
        double s1 = (double)0.01f * ss1;
        double ss2 = fabs(S.del2);
        double s2 = (double)0.01f * ss2;
        if (s1 > ss2 || s2 > ss1 || S.nev == 0) {
          S.c3 = 0.5 * (S.c1 + S.c2);
          S.nev = 1;
          S.m = 1;
          S.ec = S.c3;
          S.phase = PH_NV_END;
          return true;
        }
        // Neville tableau x(y) in the lane's scratch slot
        double *xx = cs + 2 * kpad, *yy = xx + 11;
        if (S.nev == 2) {
          xx[S.m] = S.c3;
          yy[S.m] = S.del3;
        } else {
          xx[0] = S.c1; yy[0] = S.del1;
          xx[1] = S.c2; yy[1] = S.del2;
          S.m = 1;
        }
        // Neville on x(y), newest point at index m (surfdisp96.f:660-668)
        double ym = yy[S.m];
        bool bad = false;
        for (int j = S.m - 1; j >= 0; --j) {
          double denom = ym - yy[j];
          if (fabs(denom) < 1.0e-10 * fabs(ym)) {
            bad = true;
            break;
          }
          xx[j] = (-yy[j] * xx[j + 1] + ym * xx[j]) / denom;
        }
        if (bad) {
          S.c3 = 0.5 * (S.c1 + S.c2);
          S.nev = 1;
          S.m = 1;
        } else {
          S.c3 = xx[0];
          S.nev = 2;
          S.m = S.m + 1;
          if (S.m > 10) S.m = 10;
        }
        S.ec = S.c3;
        S.phase = PH_NV_END;
        return true;
      }
      case PH_NV_END: {
        S.del3 = F;
        S.phase = PH_NV_LOOP;
        continue;
      }
      case PH_FP: {
        if constexpr (!EXACT) return false;
        // c1 the retained end, c2 the newest; the retained end's value is
        // scaled by 1 - F/del2 (0.5 when that is <= 0) while it stays
        const double fc = F;
        double xo = 0.0, fo = 0.0;  // the point that stops being an end
        if (fc == 0.0) {
          S.c1 = S.c3;
          S.c2 = S.c3;
        } else {
          if (sign1(fc) != sign1(S.del2)) {
            xo = S.c1;
            fo = S.del1t;
            S.c1 = S.c2;
            S.del1 = S.del2;
            S.del1t = S.del2;
          } else {
            xo = S.c2;
            fo = S.del2;
            const double m = 1.0 - fc / S.del2;
            S.del1 = ((m > 0.0) ? m : 0.5) * S.del1;
          }
          S.c2 = S.c3;
          S.del2 = fc;
        }
        ++S.nctrl;
        bool done = fc == 0.0 || fabs(S.c2 - S.c1) <= A.tol * fabs(S.c2) || S.nctrl >= 80;
        double root = S.c2;
        if (!done) {
          double cn = (S.c1 * S.del2 - S.c2 * S.del1) / (S.del2 - S.del1);
          if (!(cn > fmin(S.c1, S.c2) && cn < fmax(S.c1, S.c2))) cn = 0.5 * (S.c1 + S.c2);
          // inverse quadratic interpolation through F at both ends and at
          // the point just dropped (Brent's step), when it falls inside
          if (fo != S.del1t && fo != S.del2 && S.del1t != S.del2) {
            const double x = xo * S.del1t * S.del2 / ((fo - S.del1t) * (fo - S.del2)) +
                             S.c1 * fo * S.del2 / ((S.del1t - fo) * (S.del1t - S.del2)) +
                             S.c2 * fo * S.del1t / ((S.del2 - fo) * (S.del2 - S.del1t));
            if (x > fmin(S.c1, S.c2) && x < fmax(S.c1, S.c2)) cn = x;
          }
          // the corrected value's error is about d1^2 / d0 for a linearly
          // converging step and less once convergence is superlinear
          const double d1 = fabs(cn - S.c3), d0 = fabs(S.c3 - S.del3), cc = fabs(S.c3);
          if ((d1 <= A.tol * cc && d0 <= 1.0e-4 * cc) ||
              (S.del3 != 0.0 && d0 <= 1.0e-3 * cc && d1 * d1 <= 0.1 * A.tol * cc * d0)) {
            done = true;
            root = cn;
          } else {
            S.del3 = S.c3;
            S.c3 = cn;
            S.ec = cn;
            return true;
          }
        }
        S.c3 = root;
        goto nv_done;
      }
      case PH_RETRY_NEXT: {
        // surfdisp.cpp:95-109: each missing period is re-solved alone; a
        // period still missing is where the reference exit(1)s.
        for (;;) {
          S.kbase++;
          if (S.kbase >= S.K) { S.phase = PH_DONE; return false; }
          unsigned need = 0;
          for (int md = 0; md < S.M; ++md)
            if (read_out(A, S, md, S.kbase) == 0.0) need |= 1u << md;
          if (need) {
            S.need = need;
            S.Mrun = 32 - __builtin_clz(need);
            S.Krun = 1;
            begin_pass(S, cs, kpad);
            break;
          }
        }
        continue;
      }
      default:
        return false;
    }

  getsol_fail:
    if (S.sub == 2) {
      S.c1 = cs[S.k];
      cs[kpad + S.k] = S.c1;
      goto record;
    }
    goto fail_k;

  nv_done:
    S.c1 = S.c3;
    if (S.c1 > (double)S.betmx) goto getsol_fail;
    // getsol success
    if (S.sub == 1) {
      cs[S.k] = S.c1;
      if (A.igr > 0) {
        S.t1 = (double)S.t1b;
        S.ifirst = 0;
        S.clow = cs[kpad + S.k] + S.one * S.dc;
        S.c1 = S.c1 - S.onea * S.dc;
        S.sub = 2;
        S.omega = kTwoPi / S.t1;
        S.ec = S.c1;
        S.phase = PH_GS_FIRST;
        return true;
      }
      S.c1 = 0.0;
      goto record;
    }
    cs[kpad + S.k] = S.c1;
    goto record;

  record : {
    float cc0 = (float)cs[S.k];
    float ccg = (float)S.c1;
    double v;
    if (A.igr == 0) {
      v = (A.tol < 1.0e-6) ? cs[S.k] : (double)cc0;
    } else {
      float gvel = (1 / S.t1a - 1 / S.t1b) / (1 / (S.t1a * cc0) - 1 / (S.t1b * ccg));
      v = (double)gvel;
    }
    int md = S.iq - 1;
    if (!S.retry) {
      write_out(A, S, md, S.k, v);
    } else if (S.need & (1u << md)) {
      write_out(A, S, md, S.kbase, v);
      if (v == 0.0) S.exitmask |= 1u << md;
    }
    S.k++;
    S.phase = PH_START_K;
    continue;
  }

  fail_k : {
    if (S.iq <= 1 && S.iverb == 0) {
      S.iverb = 1;
      S.ierr = 1;
    }
    int k0 = S.k;
    S.ift = k0;
    int md = S.iq - 1;
    if (!S.retry) {
      for (int i = k0; i < S.Krun; ++i) write_out(A, S, md, i, 0.0);
    } else if (S.need & (1u << md)) {
      S.exitmask |= 1u << md;
    }
    S.iq++;
    S.k = 0;
    S.phase = PH_START_K;
    continue;
  }
  }
}

// The chain taken i-th: period variant i / ncol, column corder[i % ncol] (the
// columns deepest first, so a wave's lanes hold chains of about one depth,
// whose layer loops end together), or chain i without an order.  Chain ids
// keep their meaning (column g % ncol, variant g / ncol) for every array
// indexed by chain.
JDSG_DEV long long chain_at(const ChainArgs &A, long long i) {
  if (!A.corder) return i;
  const long long nc = A.cm.ncol;
  return (i / nc) * nc + A.corder[i % nc];
}

// Decode chain id -> job; returns false when the chain has no work.
JDSG_DEV bool setup_chain(ChainState &S, const ChainArgs &A, long long g, double *cs, int w) {
  int ncol = A.cm.ncol;
  int col = (int)(g % ncol);
  int variant = (int)(g / ncol);
  int nR = A.nmode[2 * col], nL = A.nmode[2 * col + 1];
  int M = (w == 0) ? nR : nL;
  int K = A.per_off[col + 1] - A.per_off[col];
  S.col = col;
  S.variant = variant;
  S.wave = (w == 0) ? 2 : 1;
  S.K = K;
  S.M = M;
  if (M <= 0 || K <= 0) return false;
  S.e0 = A.ent_off[col] + (long long)((w == 0) ? 0 : nR) * K;
  S.tscale = (variant == 1) ? (1.0 + 0.5 * 0.01) : (1.0 - 0.5 * 0.01);
  S.cc = (double)A.cm.cc[col];
  S.cm = S.cc;
  S.dc = fabs((double)0.005f);
  S.onea = (double)1.5f;
  S.one = 1.0e-2;
  S.betmx = A.cm.betmx[col];
  S.del1st = 0.0;
  S.retry = 0;
  S.del1st_ok = 0;
  S.tfirst = 0.0;
  S.cfirst = 0.0;
  S.exitmask = 0;
  S.need = 0;
  S.Mrun = M;
  S.Krun = K;
  S.kbase = 0;
  begin_pass(S, cs, A.kmax_pad);
  return true;
}

JDSG_DEV void finish_chain(ChainState &S, const ChainArgs &A, long long g, int w) {
  if (S.exitmask) {
    for (int md = 0; md < S.M; ++md) {
      if (!(S.exitmask & (1u << md))) continue;
      for (int kk = 0; kk < S.K; ++kk) {
        long long e = S.e0 + (long long)md * S.K + kk;
        A.cstat[(long long)S.variant * A.nent + e] = JDSG_EXIT_PATH;
      }
    }
  }
  if (A.ierr_out) A.ierr_out[(long long)w * A.cm.ncol * A.nvar + g] = S.ierr;
}

}  // namespace

// Persistent chain kernel, one launch per wave type (W = 0 Rayleigh, 1 Love):
// lane-level work fetching.  Every trip of the main loop evaluates one period
// equation per live lane.  Chain g of the launch is column g % ncol, period
// scale g / ncol.  Three waves per SIMD (at most 168 VGPRs, a small spill)
// measured faster than two (202 VGPRs, no spill) or four.  FAST selects the
// throughput form of the period equations for exact-root mode (tolerance
// below 1e-6), where the roots are polished past the last-bit differences
// between the two forms; the reference-tolerance mode keeps the faithful one.
template <int W, bool FAST>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3)))
jdsg_root_chains(ChainArgs A) {
  const long long nlanes = (long long)gridDim.x * blockDim.x;
  const long long lane_id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double *cs = A.cscr + lane_id * (2 * A.kmax_pad + 22);
  unsigned *counter = A.counter + W;
  const long long nch = (long long)A.cm.ncol * A.nvar;
  unsigned long long nev = 0;
  unsigned cev = 0;  // evaluations of the current chain
  ChainState S;
  long long i = lane_id, g = 0;
  bool live = false;
  while (i < nch) {
    g = chain_at(A, i);
    if (setup_chain(S, A, g, cs, W)) { live = true; break; }
    if (A.ierr_out) A.ierr_out[(long long)W * nch + g] = 0;
    i = nlanes + (long long)atomicAdd(counter, 1u);
  }
  double F = 0.0;
  bool need = live ? advance<false, FAST>(S, A, cs, F) : false;
  const float4 *base = (W == 0) ? A.cm.chainR : A.cm.chainL;
  while (live) {
    if (need) {
      const float4 *st = base + S.col;
      int mmax = A.cm.nlay[S.col];
      int llw = A.cm.llw[S.col];
      double wvno = S.omega / S.ec;
      if (FAST)
        F = (W == 0) ? jdsg_pf::dltar4(st, A.cm.ncol_pad, mmax, llw, wvno, S.omega)
                     : jdsg_pf::dltar1(st, A.cm.ncol_pad, mmax, llw, wvno, S.omega);
      else
        F = (W == 0) ? dltar4(st, A.cm.ncol_pad, mmax, llw, wvno, S.omega)
                     : dltar1(st, A.cm.ncol_pad, mmax, llw, wvno, S.omega);
      ++nev;
      ++cev;
      need = advance<false, FAST>(S, A, cs, F);
      continue;
    }
    finish_chain(S, A, g, W);
    if (A.chain_evals) A.chain_evals[(long long)W * nch + g] = cev;
    cev = 0;
    live = false;
    for (;;) {
      i = nlanes + (long long)atomicAdd(counter, 1u);
      if (i >= nch) break;
      g = chain_at(A, i);
      if (setup_chain(S, A, g, cs, W)) { live = true; break; }
      if (A.ierr_out) A.ierr_out[(long long)W * nch + g] = 0;
    }
    if (live) need = advance<false, FAST>(S, A, cs, 0.0);
  }
  if (A.nevals && nev) atomicAdd(A.nevals, nev);
}

// The exact-root fallback form of jdsg_root_chains<W, true>: a chain that enters a long upward scan
// is saved and leaves a scan job (PH_SUSPEND); a later launch resumes the
// chains of a job list from their saved state once jdsg_scan_jobs has run
// the scans.  Each chain then keeps its scratch in a slot of its own.
// COOP: one chain per wave instead of one per lane.  Every lane then holds
// the same chain state and takes the same branches, the period equation is
// formed by the whole wave (the layers' coefficients in parallel, the sweep
// in every lane: the one-lane value in a fraction of its latency, for about
// ten times the instructions), and lane 0 alone takes work and job slots
// and posts the saved state.  It serves the resumed chains, a tail of long
// walks whose latency sets the time of each cycle.
template <int W, bool COOP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3)))
jdsg_root_chains_susp(ChainArgs A) {
  __shared__ double cf[COOP ? 64 * (W == 0 ? jdsg_pf::kCoef4 : jdsg_pf::kCoef1) : 1];
  const long long nlanes = COOP ? (long long)gridDim.x : (long long)gridDim.x * blockDim.x;
  const long long lane_id =
      COOP ? (long long)blockIdx.x : (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool lead = !COOP || threadIdx.x == 0;
  const long long cstride = 2 * A.kmax_pad + 22;
  double *cs = nullptr;  // the chain's own slot, set by start_item
  unsigned *counter = A.counter + W;
  const long long nch = (long long)A.cm.ncol * A.nvar;
  // work items: every chain, or the chains of a job list to resume
  const ScanJob *jin = nullptr;
  long long nitems = nch;
  if (A.susp.list_in >= 0) {
    jin = A.susp.jobs[W][A.susp.list_in];
    nitems = A.susp.njobs[W * 2 + A.susp.list_in];
  }
  ChainStateSusp *saved = reinterpret_cast<ChainStateSusp *>(A.susp.saved) + (long long)W * nch;
  unsigned long long nev = 0;
  ChainStateSusp S;
  long long g = 0;
  // set up work item i in S (and cs); false when it has no work
  auto start_item = [&](long long i) -> bool {
    if (jin) {
      const ScanJob &J = jin[i];
      g = J.g;
      cs = A.susp.cs + ((long long)W * nch + g) * cstride;
      S = saved[g];
      if (J.found) {
        S.c1 = J.c_prev;
        S.del1 = J.f_prev;
        S.c2 = J.c_hit;
        S.del2 = J.f_hit;
        S.phase = PH_SCAN_HIT;
      } else {
        S.c1 = J.c_hit;
        S.del1 = J.f_hit;
        S.phase = PH_SCAN_END;
      }
      S.evals += J.evals;
      return true;
    }
    g = chain_at(A, i);
    cs = A.susp.cs + ((long long)W * nch + g) * cstride;
    if (setup_chain(S, A, g, cs, W)) {
      S.nscan = 0;
      S.evals = 0;
      return true;
    }
    if (A.ierr_out) A.ierr_out[(long long)W * nch + g] = 0;
    return false;
  };
  // the next work item of this lane (COOP: of this wave)
  auto next_item = [&]() -> long long {
    unsigned v = 0;
    if (lead) v = atomicAdd(counter, 1u);
    if (COOP) v = __shfl(v, 0);
    return nlanes + (long long)v;
  };
  long long item = lane_id;
  bool live = false;
  while (item < nitems) {
    if (start_item(item)) { live = true; break; }
    item = next_item();
  }
  double F = 0.0;
  bool need = live ? advance<true, true>(S, A, cs, F) : false;
  const float4 *base = (W == 0) ? A.cm.chainR : A.cm.chainL;
  while (live) {
    if (need) {
      const float4 *st = base + S.col;
      int mmax = A.cm.nlay[S.col];
      int llw = A.cm.llw[S.col];
      double wvno = S.omega / S.ec;
      if (COOP)
        F = (W == 0) ? jdsg_pf::dltar4_wave(st, A.cm.ncol_pad, mmax, llw, wvno, S.omega, cf)
                     : jdsg_pf::dltar1_wave(st, A.cm.ncol_pad, mmax, llw, wvno, S.omega, cf);
      else
        F = (W == 0) ? jdsg_pf::dltar4(st, A.cm.ncol_pad, mmax, llw, wvno, S.omega)
                     : jdsg_pf::dltar1(st, A.cm.ncol_pad, mmax, llw, wvno, S.omega);
      ++nev;
      ++S.evals;
      need = advance<true, true>(S, A, cs, F);
      continue;
    }
    if (S.phase == PH_SUSPEND) {
      // at most one job per chain and list, so the slot stays below nch
      unsigned slot = 0;
      if (lead) {
        slot = atomicAdd(&A.susp.njobs[W * 2 + A.susp.list_out], 1u);
        atomicAdd(&A.susp.total[W], 1u);
      }
      if (COOP) slot = __shfl(slot, 0);
      ScanJob J;
      J.g = g;
      J.c0 = S.c1;
      J.f0 = S.del1;
      J.omega = S.omega;
      J.bound = (double)S.betmx + S.dc;
      J.dc = S.dc;
      J.col = S.col;
      J.found = 0;
      J.evals = 0;
      J.c_prev = J.f_prev = J.c_hit = J.f_hit = 0.0;
      if (lead) {
        A.susp.jobs[W][A.susp.list_out][slot] = J;
        saved[g] = S;
      }
    } else {
      finish_chain(S, A, g, W);
      if (A.chain_evals) A.chain_evals[(long long)W * nch + g] = S.evals;
    }
    live = false;
    for (;;) {
      item = next_item();
      if (item >= nitems) break;
      if (start_item(item)) { live = true; break; }
    }
    if (live) need = advance<true, true>(S, A, cs, 0.0);
  }
  if (lead && A.nevals && nev) atomicAdd(A.nevals, nev);
}

// Scan jobs of suspended chains, one wave per job: each round the 64 lanes
// take the next 64 points of the chain's scan (lane l adds dc l+1 times to
// the round's base, the additions the chain itself would make), evaluate the
// period equation where the chain would (its previous point still below the
// bound), and the first lane with a sign change against its predecessor, or
// the first point past the bound, ends the scan -- the chain's own outcome.
// It also clears, for the chain launch that follows, the work counter and
// the count of the other list (consumed by the previous chain launch).
template <int W>
__global__ void __launch_bounds__(64) jdsg_scan_jobs(ChainArgs A, int list) {
  ScanJob *jobs = A.susp.jobs[W][list];
  const unsigned n = A.susp.njobs[W * 2 + list];
  const int lane = threadIdx.x;
  if (blockIdx.x == 0 && lane == 0) {
    A.counter[W] = 0u;
    A.susp.njobs[W * 2 + (list ^ 1)] = 0u;
  }
  const float4 *base = (W == 0) ? A.cm.chainR : A.cm.chainL;
  unsigned long long nev = 0;
  for (unsigned j = blockIdx.x; j < n; j += gridDim.x) {
    ScanJob J = jobs[j];
    const float4 *st = base + J.col;
    const int mmax = A.cm.nlay[J.col], llw = A.cm.llw[J.col];
    double cb = J.c0, fb = J.f0;
    unsigned evals = 0;
    for (int round = 0;; ++round) {
      double c = cb;
      for (int s = 0; s <= lane; ++s) c = c + J.dc;
      double cprev = __shfl_up(c, 1);
      if (lane == 0) cprev = cb;
      const bool eval = cprev < J.bound;
      double F = 0.0;
      if (eval) {
        const double wvno = J.omega / c;
        F = (W == 0) ? jdsg_pf::dltar4(st, A.cm.ncol_pad, mmax, llw, wvno, J.omega)
                     : jdsg_pf::dltar1(st, A.cm.ncol_pad, mmax, llw, wvno, J.omega);
      }
      double fprev = __shfl_up(F, 1);
      if (lane == 0) fprev = fb;
      const bool change = eval && sign1(F) != sign1(fprev);
      const bool end = eval && !change && c >= J.bound;
      const unsigned long long me = __ballot(eval);
      evals += (unsigned)__popcll(me);
      const unsigned long long ev = __ballot(change) | __ballot(end);
      if (ev) {
        const int first = __ffsll((long long)ev) - 1;
        const double cp = __shfl(cprev, first), fp = __shfl(fprev, first);
        const double ch = __shfl(c, first), fh = __shfl(F, first);
        const bool found = (__ballot(change) >> first) & 1ull;
        if (lane == 0) {
          J.found = found ? 1 : 0;
          J.c_prev = cp;
          J.f_prev = fp;
          J.c_hit = ch;
          J.f_hit = fh;
          J.evals = evals;
          jobs[j] = J;
        }
        break;
      }
      // all 64 points inside and no sign change: the next round starts at
      // the last of them
      cb = __shfl(c, 63);
      fb = __shfl(F, 63);
    }
    nev += evals;
  }
  if (lane == 0 && A.nevals && nev) atomicAdd(A.nevals, nev);
}

template __global__ void jdsg_root_chains<0, false>(ChainArgs);
template __global__ void jdsg_root_chains<1, false>(ChainArgs);
template __global__ void jdsg_root_chains<0, true>(ChainArgs);
template __global__ void jdsg_root_chains<1, true>(ChainArgs);
template __global__ void jdsg_root_chains_susp<0, false>(ChainArgs);
template __global__ void jdsg_root_chains_susp<1, false>(ChainArgs);
template __global__ void jdsg_root_chains_susp<0, true>(ChainArgs);
template __global__ void jdsg_root_chains_susp<1, true>(ChainArgs);

size_t jdsg_chain_state_bytes() { return sizeof(ChainStateSusp); }

// ----------------------------------------------------------------------------
// host launchers
void jdsg_launch_prep(int ncol, int ncol_pad, const int *layer_off, const float *thk,
                      const float *vp, const float *vs, const float *rho, int iflsph,
                      const PrepOut &o, hipStream_t st) {
  int blocks = (ncol + 255) / 256;
  hipLaunchKernelGGL(jdsg_prep_columns, dim3(blocks), dim3(256), 0, st, ncol, ncol_pad,
                     layer_off, thk, vp, vs, rho, iflsph, o);
}

void jdsg_launch_chains(const ChainArgs &A, int nwaves, hipStream_t st) {
  // Rayleigh chains, then Love chains (the Love launch is skipped when no
  // column asks for Love modes, and likewise for Rayleigh)
  const bool fast = A.tol < 1.0e-6;
  if (A.any_wave[0]) {
    if (fast)
      hipLaunchKernelGGL((jdsg_root_chains<0, true>), dim3(nwaves), dim3(64), 0, st, A);
    else
      hipLaunchKernelGGL((jdsg_root_chains<0, false>), dim3(nwaves), dim3(64), 0, st, A);
  }
  if (A.any_wave[1]) {
    if (fast)
      hipLaunchKernelGGL((jdsg_root_chains<1, true>), dim3(nwaves), dim3(64), 0, st, A);
    else
      hipLaunchKernelGGL((jdsg_root_chains<1, false>), dim3(nwaves), dim3(64), 0, st, A);
  }
}

void jdsg_launch_chains_susp(ChainArgs A, int nwaves, int cycles, bool coop, hipStream_t st) {
  A.susp_on = 1;
  const long long nch = (long long)A.cm.ncol * A.nvar;
  const unsigned scan_blocks = (unsigned)std::min<long long>(nch, 2048);
  // after a few cycles only the long chains are left: a smaller grid keeps
  // the near-empty resume launches cheap (lanes still take every job)
  const int tail_waves = std::max(64, nwaves / 8);
  // coop: the first launch (every chain, most with little or no work) runs
  // a chain per lane, for throughput; the resumed chains -- a tail of long
  // walks -- a chain per wave, three waves on each SIMD for the first
  // cycles and one after them (a launch without jobs then costs what the
  // lane form's small grid does)
  auto chains = [&](int w, int gw, bool first, bool early) {
    if (coop && !first) {
      const unsigned cw = (unsigned)std::min<long long>(nch, early ? 3072 : 1024);
      if (w == 0)
        hipLaunchKernelGGL((jdsg_root_chains_susp<0, true>), dim3(cw), dim3(64), 0, st, A);
      else
        hipLaunchKernelGGL((jdsg_root_chains_susp<1, true>), dim3(cw), dim3(64), 0, st, A);
    } else {
      if (w == 0)
        hipLaunchKernelGGL((jdsg_root_chains_susp<0, false>), dim3(gw), dim3(64), 0, st, A);
      else
        hipLaunchKernelGGL((jdsg_root_chains_susp<1, false>), dim3(gw), dim3(64), 0, st, A);
    }
  };
  for (int w = 0; w < 2; ++w) {
    if (!A.any_wave[w]) continue;
    A.susp.list_in = -1;
    A.susp.list_out = 0;
    A.susp.allow = cycles > 0;
    jdsg_launch_fill_u32(A.counter + w, 1, 0u, st);
    jdsg_launch_fill_u32(A.susp.njobs + 2 * w, 2, 0u, st);
    chains(w, nwaves, true, true);
    for (int c = 0; c < cycles; ++c) {
      const int lin = c & 1, lout = lin ^ 1;
      if (w == 0)
        hipLaunchKernelGGL((jdsg_scan_jobs<0>), dim3(scan_blocks), dim3(64), 0, st, A, lin);
      else
        hipLaunchKernelGGL((jdsg_scan_jobs<1>), dim3(scan_blocks), dim3(64), 0, st, A, lin);
      A.susp.list_in = lin;
      A.susp.list_out = lout;
      A.susp.allow = c + 1 < cycles;  // the last pass finishes every chain
      chains(w, c < 4 ? nwaves : std::min(nwaves, tail_waves), false, c < 4);
    }
  }
}

__global__ void jdsg_fill_u32(unsigned *p, long long n, unsigned v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void jdsg_launch_fill_u32(unsigned *p, long long n, unsigned v, hipStream_t st) {
  if (n <= 0) return;
  long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(jdsg_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
}
