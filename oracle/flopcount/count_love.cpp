// Driver of the op-counting build for slegn96/slegnpu (swd_oracle_love.c
// compiled with FlopD; TEST INFRASTRUCTURE ONLY).
#include <vector>
#include "swd_oracle_love.cpp"

void count_eigen_l(const float *thk, const float *vs, const float *rho, int n, int group,
                   double T, const double c3[3]) {
  std::vector<FlopD> w(16 * (size_t)n);
  FlopD t = FlopD(T), cp = FlopD(c3[0]), cg = FlopD(0.0);
  FlopD *o[16];
  for (int i = 0; i < 16; ++i) o[i] = &w[(size_t)i * n];
  if (!group) {
    or_slegn96(thk, vs, rho, n, &t, &cp, &cg, o[0], o[1], o[2], o[3], o[4], 1);
  } else {
    FlopD t1 = FlopD(T * (1.0 + 0.5 * 0.01)), t2 = FlopD(T * (1.0 - 0.5 * 0.01));
    FlopD c1 = FlopD(c3[1]), c2 = FlopD(c3[2]);
    or_slegnpu(thk, vs, rho, n, &t, &cp, &cg, o[0], o[1], &t1, &c1, &t2, &c2, o[2], o[3], o[4],
               o[5], o[6], o[7], 1);
  }
}

void count_eigen_r(const float *thk, const float *vp, const float *vs, const float *rho, int n,
                   int group, double T, const double c3[3]);
void count_eigen(const float *thk, const float *vp, const float *vs, const float *rho, int n,
                 int wave, int group, double T, const double c3[3]) {
  if (wave == 0)
    count_eigen_r(thk, vp, vs, rho, n, group, T, c3);
  else
    count_eigen_l(thk, vs, rho, n, group, T, c3);
}
