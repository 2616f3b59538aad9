// Driver of the op-counting build for the root-chain restatement
// (swd_oracle_disp.c compiled with FlopD; TEST INFRASTRUCTURE ONLY).
#include <vector>
#include "swd_oracle_disp.cpp"

// flops of one period-equation evaluation (dltar4 Rayleigh, dltar1 Love) on
// the flattened model surfdisp96 builds, at phase velocity c and period T
double count_disp_eval(const float *thk, const float *vp, const float *vs, const float *rho,
                       int n, int wave, double T, double c, int iflsph) {
  std::vector<float> m(7 * (size_t)n);
  float cc, bx;
  or_chain_model(thk, vp, vs, rho, n, iflsph, &m[0], &m[n], &m[2 * n], &m[3 * n], &m[4 * n],
                 &cc, &bx);
  chain_ctx X;
  X.mmax = n;
  X.llw = (m[2 * n] <= 0.0f) ? 2 : 1;
  X.d = &m[0];
  X.a = &m[n];
  X.b = &m[2 * n];
  X.rho = (wave == 0) ? &m[3 * n] : &m[4 * n];
  X.nevals = 0;
  double omega = 2.0 * 3.141592653589793 / T;
  unsigned long long f0 = g_flops;
  FlopD wvno = FlopD(omega / c);
  if (wave == 0)
    (void)dltar4(&X, wvno, FlopD(omega));
  else
    (void)dltar1(&X, wvno, FlopD(omega));
  return (double)(g_flops - f0);
}

// flat roots of mode `mode` (0-based) at T, T(1+.005), T(1-.005) as the
// chain finds them (surfdisp96, spherical model), 0 where none
void count_roots(const float *thk, const float *vp, const float *vs, const float *rho, int n,
                 int wave, int mode, const double *T, int K, double *cflat) {
  std::vector<FlopD> t(K), cg(K);
  const double sc[3] = {1.0, 1.005, 0.995};
  for (int v = 0; v < 3; ++v) {
    for (int k = 0; k < K; ++k) t[k] = FlopD(T[k] * sc[v]);
    int ierr = 0;
    long long nev = 0;
    or_surfdisp96(thk, vp, vs, rho, n, 1, wave == 0 ? 2 : 1, mode + 1, 0, K, t.data(), cg.data(),
                  &ierr, &nev);
    for (int k = 0; k < K; ++k) cflat[(size_t)v * K + k] = (double)cg[k];
  }
}
