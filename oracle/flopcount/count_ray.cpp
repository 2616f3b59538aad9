// Driver of the op-counting build for sregn96/sregnpu (swd_oracle_ray.c
// compiled with FlopD; TEST INFRASTRUCTURE ONLY).
#include <vector>
#include "swd_oracle_ray.cpp"

void count_eigen_r(const float *thk, const float *vp, const float *vs, const float *rho, int n,
                   int group, double T, const double c3[3]) {
  std::vector<FlopD> w(16 * (size_t)n);
  FlopD t = FlopD(T), cp = FlopD(c3[0]), cg = FlopD(0.0);
  FlopD *o[16];
  for (int i = 0; i < 16; ++i) o[i] = &w[(size_t)i * n];
  if (!group) {
    or_sregn96(thk, vp, vs, rho, n, &t, &cp, &cg, o[0], o[1], o[2], o[3], o[4], o[5], o[6],
               o[7], 1);
  } else {
    FlopD t1 = FlopD(T * (1.0 + 0.5 * 0.01)), t2 = FlopD(T * (1.0 - 0.5 * 0.01));
    FlopD c1 = FlopD(c3[1]), c2 = FlopD(c3[2]);
    or_sregnpu(thk, vp, vs, rho, n, &t, &cp, &cg, o[0], o[1], o[2], o[3], &t1, &c1, &t2, &c2,
               o[4], o[5], o[6], o[7], o[8], o[9], o[10], o[11], 1);
  }
}
