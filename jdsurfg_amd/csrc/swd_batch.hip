// swd_batch.hip -- device helpers of the iterated batched solve: the model
// update of one inversion iteration (a seeded relative Vs perturbation with
// Vp and rho following Brocher's relations, empirical.cpp:9-18) and the
// iteration counter that seeds it, both captured with the solve in one
// hipGraph so that an iteration is a single graph launch.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

}  // namespace

// vs[i] *= 1 + amp * u_i with u_i uniform in [-1, 1) from (seed, *iter, i);
// Vp and rho recomputed from the new Vs (double polynomials rounded to
// float).  Fluid layers (vs == 0) are left as they are.
extern "C" __global__ void jdsg_perturb_model(long long n, float *vs, float *vp, float *rho,
                                              const int *iter, float amp, uint32_t seed) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = vs[i];
  if (!(v > 0.0f)) return;
  uint32_t h = mix32(mix32(seed ^ (uint32_t)(*iter) * 0x9e3779b9U) ^ (uint32_t)i ^
                     (uint32_t)(i >> 32) * 0x85ebca6bU);
  float u = (float)(h >> 8) * (1.0f / 8388608.0f) - 1.0f;
  v = v * (1.0f + amp * u);
  vs[i] = v;
  double s = (double)v;
  double a = 0.9409 + 2.0947 * s - 0.8206 * s * s + 0.2683 * s * s * s - 0.0251 * s * s * s * s;
  float af = (float)a;
  vp[i] = af;
  double q = (double)af;
  rho[i] = (float)(1.6612 * q - 0.4721 * q * q + 0.0671 * q * q * q - 0.0043 * q * q * q * q +
                   0.000106 * q * q * q * q * q);
}

extern "C" __global__ void jdsg_bump(int *iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *iter = *iter + 1;
}

void jdsg_launch_perturb(long long n, float *vs, float *vp, float *rho, const int *iter,
                         float amp, uint32_t seed, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(jdsg_perturb_model, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     vs, vp, rho, iter, amp, seed);
}
void jdsg_launch_bump(int *iter, hipStream_t st) {
  hipLaunchKernelGGL(jdsg_bump, dim3(1), dim3(64), 0, st, iter);
}
