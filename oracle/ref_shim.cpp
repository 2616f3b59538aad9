// ref_shim.cpp -- binds the compiled reference (src/SWD/surfdisp96.f,
// sregn96.f90, slegn96.f90, surfdisp.cpp, built from their own sources into
// oracle/_ref) to the oracle's or_backend table, so the oracle glue can drive
// the reference numerics.  TEST INFRASTRUCTURE ONLY; built into oracle/_ref.
#include <string>
#include "SWD/swd.hpp"
#include "swd_oracle.h"

static const char *kCode[4] = {"Rc", "Rg", "Lc", "Lg"};

// surfdisp.cpp:64 -- the reference wrapper exit(1)s on a missing root.
static int ref_surfdisp(const float *thk, const float *vp, const float *vs,
                        const float *rho, int nlayer, const double *t,
                        double *cg, int kmax, int swdtp, int mode, int sphere,
                        int keep_flat) {
  surfdisp(thk, vp, vs, rho, nlayer, t, cg, kmax, kCode[swdtp], mode,
           sphere != 0, keep_flat != 0);
  return 0;
}
static void ref_sregnpu(const float *a, const float *b, const float *c,
                        const float *d, int n, const double *t, double *cp,
                        double *cg, double *e, double *f, double *g, double *h,
                        double *t1, double *cp1, double *t2, double *cp2,
                        double *i, double *j, double *k, double *l, double *m,
                        double *o, double *p, double *q, int fl) {
  sregnpu_(a, b, c, d, n, t, cp, cg, e, f, g, h, t1, cp1, t2, cp2, i, j, k, l,
           m, o, p, q, fl);
}
static void ref_slegnpu(const float *a, const float *b, const float *c, int n,
                        const double *t, double *cp, double *cg, double *d,
                        double *e, double *t1, double *cp1, double *t2,
                        double *cp2, double *f, double *g, double *h, double *i,
                        double *j, double *k, int fl) {
  slegnpu_(const_cast<float *>(a), const_cast<float *>(b),
           const_cast<float *>(c), n, t, cp, cg, d, e, t1, cp1, t2, cp2, f, g,
           h, i, j, k, fl);
}
static const or_backend REF = {ref_surfdisp, sregn96_, ref_sregnpu, slegn96_,
                               ref_slegnpu};

extern "C" const or_backend *ref_backend(void) { return &REF; }

// Direct access to the reference root chain (surfdisp96.f:57).
extern "C" void ref_surfdisp96(const float *thk, const float *vp,
                               const float *vs, const float *rho, int nlayer,
                               int iflsph, int iwave, int mode, int igr,
                               int kmax, const double *t, double *cg,
                               int *ierr) {
  surfdisp96_(thk, vp, vs, rho, nlayer, iflsph, iwave, mode, igr, kmax, t, cg,
              ierr);
}
