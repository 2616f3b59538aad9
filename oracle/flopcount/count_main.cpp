// count_main.cpp -- runs the op-counting build of the oracle restatement on
// the benchmark's layer stacks and prints the per-layer flop counts of the
// frozen model (TEST INFRASTRUCTURE ONLY; built by `make -C oracle flopcount`).
//
//   f_disp(wave) = flops of one period-equation evaluation / N
//                  (dltar4 Rayleigh, dltar1 Love; surfdisp96.f:712-898)
//   f_eig(wave)  = flops of the eigen/kernel call / (passes * N)
//                  (sregn96/slegn96: 1 pass; sregnpu/slegnpu: 3 passes)
// averaged over the stacks, periods and modes given on stdin:
//   line 1: nstack; then per stack: N, then N lines "thk vp vs rho";
//   then: nper, the periods.
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <vector>

#include "flopd.hpp"

unsigned long long g_flops = 0;
unsigned long long g_shared = 0;  // flops of intijr's per-layer shared part

// from the generated translation units
double count_disp_eval(const float *thk, const float *vp, const float *vs, const float *rho,
                       int n, int wave, double T, double c, int iflsph);
void count_roots(const float *thk, const float *vp, const float *vs, const float *rho, int n,
                 int wave, int mode, const double *T, int K, double *cflat);
void count_eigen(const float *thk, const float *vp, const float *vs, const float *rho, int n,
                 int wave, int group, double T, const double c3[3]);

int main() {
  int nst = 0;
  if (scanf("%d", &nst) != 1) return 1;
  std::vector<std::vector<float>> S(nst);
  std::vector<int> Ns(nst);
  for (int s = 0; s < nst; ++s) {
    scanf("%d", &Ns[s]);
    S[s].resize(4 * (size_t)Ns[s]);
    for (int i = 0; i < Ns[s]; ++i)
      for (int q = 0; q < 4; ++q) scanf("%f", &S[s][(size_t)q * Ns[s] + i]);
  }
  int K = 0;
  scanf("%d", &K);
  std::vector<double> T(K);
  for (int k = 0; k < K; ++k) scanf("%lf", &T[k]);
  // per wave: flops and layer-evaluations of the period equation, eigen
  // phase form, eigen group form
  double fd[2] = {0, 0}, nd[2] = {0, 0}, fe[2][2] = {{0, 0}, {0, 0}}, ne[2][2] = {{0, 0}, {0, 0}};
  double sh[2][2] = {{0, 0}, {0, 0}};  // intijr shared-part flops
  const int modes[2] = {4, 2};
  for (int s = 0; s < nst; ++s) {
    int n = Ns[s];
    const float *thk = &S[s][0], *vp = thk + n, *vs = vp + n, *rho = vs + n;
    for (int w = 0; w < 2; ++w)
      for (int m = 0; m < modes[w]; ++m) {
        std::vector<double> c(3 * (size_t)K);
        count_roots(thk, vp, vs, rho, n, w, m, T.data(), K, c.data());
        for (int k = 0; k < K; ++k) {
          if (c[k] <= 0.0) continue;
          // the period equation at the root, spherical model as the chain uses
          unsigned long long f0 = g_flops;
          count_disp_eval(thk, vp, vs, rho, n, w, T[k], c[k], 1);
          fd[w] += (double)(g_flops - f0);
          nd[w] += n;
          for (int g = 0; g < 2; ++g) {
            if (g && (c[K + k] <= 0.0 || c[2 * K + k] <= 0.0)) continue;
            double c3[3] = {c[k], c[K + k], c[2 * K + k]};
            f0 = g_flops;
            unsigned long long s0 = g_shared;
            count_eigen(thk, vp, vs, rho, n, w, g, T[k], c3);
            fe[w][g] += (double)(g_flops - f0);
            sh[w][g] += (double)(g_shared - s0);
            ne[w][g] += (double)n * (g ? 3 : 1);
          }
        }
      }
  }
  // non-redundant: intijr's six calls per layer share E, E^-1, the amplitude
  // projections and the layer functions; count them once
  auto nr = [&](int w, int g) { return (fe[w][g] - sh[w][g] * 5.0 / 6.0) / ne[w][g]; };
  printf("{\"f_disp\": {\"rayleigh\": %.2f, \"love\": %.2f}, "
         "\"f_eig_phase\": {\"rayleigh\": %.2f, \"love\": %.2f}, "
         "\"f_eig\": {\"rayleigh\": %.2f, \"love\": %.2f}, "
         "\"f_eig_as_executed\": {\"rayleigh\": %.2f, \"love\": %.2f}}\n",
         fd[0] / nd[0], fd[1] / nd[1], nr(0, 0), nr(1, 0), nr(0, 1), nr(1, 1),
         fe[0][1] / ne[0][1], fe[1][1] / ne[1][1]);
  return 0;
}
