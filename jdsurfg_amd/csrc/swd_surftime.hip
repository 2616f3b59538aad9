// swd_surftime.hip -- the SurfTime column glue on the GPU
// (src/SWD/SurfaceWave.cpp:59-155, 458-502; src/SWD/empirical.cpp:9-34):
//   * grid -> layer refinement: Brocher Vp/rho per node (float), then
//     grid2LayerModel with 3 sublayers and mid-point values;
//   * epilogue: layer kernels -> node kernels (_ParamConvert) and the
//     Brocher chain rule  K = Kb + Ka dVp/dVs + Krho drho/dVp dVp/dVs,
//     written in SurfTime's column-major (nx*ny, nz, kmax) order.
#include "swd_kernels.hpp"

#pragma clang fp contract(off)

namespace {

// empirical.cpp:9-18 (double polynomial of float inputs, float results)
JDSG_DEV void brocher(float vsz, float &vpz, float &rhoz) {
  double v = (double)vsz;
  vpz = (float)(0.9409 + 2.0947 * v - 0.8206 * (v * v) + 0.2683 * pow(v, 3.0) -
                0.0251 * pow(v, 4.0));
  double p = (double)vpz;
  rhoz = (float)(1.6612 * p - 0.4721 * (p * p) + 0.0671 * pow(p, 3.0) - 0.0043 * pow(p, 4.0) +
                 0.000106 * pow(p, 5.0));
}
// empirical.cpp:26-34
JDSG_DEV void brocher_deriv(float vp, float vs, float &drda, float &dadb) {
  double a = (double)vp, b = (double)vs;
  drda = (float)(1.6612 - 0.4721 * 2 * a + 0.0671 * 3 * (a * a) - 0.0043 * 4 * pow(a, 3.0) +
                 0.000106 * 5 * pow(a, 4.0));
  dadb = (float)(2.0947 - 0.8206 * 2 * b + 0.2683 * 3 * (b * b) - 0.0251 * 4 * pow(b, 3.0));
}

}  // namespace

// mask[n] = 1 where column n has a water top (vs(i,j,0) == 0), else 0:
// grid2LayerModel and _ParamConvert decide this per column
// (SurfaceWave.cpp:120, 71, 495).
extern "C" __global__ void __launch_bounds__(256)
jdsg_water_mask(int nxy, const float *__restrict__ vs, int *__restrict__ mask) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < nxy) mask[n] = vs[n] == 0.0f ? 1 : 0;
}

// One thread per column of a chunk: chunk column nl is grid column
// n = cols[col0 + nl] (all columns of a chunk share the water class).  vs is
// (nx, ny, nz) column-major; column n = i + j*nx (SurfaceWave.cpp:433-434).
// The chunk's layer stacks are written column-contiguous (column nl at
// nl*N); with a water top N = 4nz-6, else 4nz-3.  vp2, when not null,
// receives 1.732 Vs per layer: the Vp that groupvel_l hands to its root
// chain (surfdisp.cpp:139).  node_vp/node_vs keep the node values for the
// chain rule.
extern "C" __global__ void __launch_bounds__(256)
jdsg_grid_to_layers(int col0, int nchunk, int nxy, int nz, int N, int water,
                    const int *__restrict__ cols, const float *__restrict__ vs,
                    const float *__restrict__ dep, float *thk, float *vp, float *vsl, float *rho,
                    float *vp2, float *node_vp, float *node_vs) {
  int nl = blockIdx.x * blockDim.x + threadIdx.x;
  if (nl >= nchunk) return;
  int n = cols[col0 + nl];
  const int sub = 3, ns = sub + 1;
  float vs0 = vs[n];
  for (int k = 0; k < nz; ++k) {
    float a, b;
    float s = vs[n + (size_t)k * nxy];
    brocher(s, a, b);
    node_vp[(size_t)k * nxy + n] = a;
    node_vs[(size_t)k * nxy + n] = s;
  }
  size_t base = (size_t)nl * N;
  int istart = 0;
  if (water) {
    istart = 1;
    float a0, r0;
    brocher(vs0, a0, r0);
    thk[base] = dep[1] - dep[0];
    rho[base] = r0;
    vsl[base] = 0.0f;
    vp[base] = a0;
  }
  int k = istart;
  for (int i = istart; i < nz - 1; ++i) {
    float s0 = vs[n + (size_t)i * nxy], s1 = vs[n + (size_t)(i + 1) * nxy];
    float p0, r0, p1, r1;
    brocher(s0, p0, r0);
    brocher(s1, p1, r1);
    float th = (dep[i + 1] - dep[i]) / ns;
    for (int j = 0; j < ns; ++j) {
      float f = (float)(2 * j + 1);
      thk[base + k] = th;
      vp[base + k] = (float)(p0 + f * (p1 - p0) / (2.0 * ns));
      vsl[base + k] = (float)(s0 + f * (s1 - s0) / (2.0 * ns));
      rho[base + k] = (float)(r0 + f * (r1 - r0) / (2.0 * ns));
      k++;
    }
  }
  k = N - 1;
  float sb = vs[n + (size_t)(nz - 1) * nxy], pb, rb;
  brocher(sb, pb, rb);
  thk[base + k] = 0.0f;
  vp[base + k] = pb;
  vsl[base + k] = sb;
  rho[base + k] = rb;
  if (vp2)
    for (int i = 0; i < N; ++i) vp2[base + i] = (float)(1.732 * (double)vsl[base + i]);
}

// One thread per (column of the chunk, SurfTime index idx).
extern "C" __global__ void __launch_bounds__(256)
jdsg_surftime_epilogue(int col0, int nchunk, int nxy, int nz, int N, int water, int kmax,
                       const int *__restrict__ cols, const StEntry *__restrict__ ents,
                       const StJob *__restrict__ jobs, const float *__restrict__ node_vp,
                       const float *__restrict__ node_vs, double *vc, double *vout,
                       double *kernel, int *status, int want_kernel) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nchunk * kmax) return;
  int nl = (int)(t % nchunk);
  int idx = (int)(t / nchunk);
  int n = cols[col0 + nl];
  StEntry E = ents[idx];
  const StJob &J = jobs[E.job];
  int woff = E.love ? J.nR : 0;
  long long ne = (long long)(J.nR + J.nL) * J.K;
  long long el = (long long)(woff + E.mode) * J.K + E.k;
  long long e = (long long)nl * ne + el;
  int st = J.status[e];
  size_t o = (size_t)n + (size_t)idx * nxy;
  vc[o] = J.c[e];
  vout[o] = (E.kind == 1) ? J.u[e] : J.c[e];
  status[o] = st;
  if (!want_kernel) return;
  const int sublayer = 3, ns = sublayer + 1;
  const int rmax = N;
  const double *const *src = (E.kind == 1) ? J.du : J.dc;
  for (int i = 0; i < nz; ++i) {
    if (water && i == 0) continue;  // SurfaceWave.cpp:495
    double acc[3];
    for (int a = 0; a < 3; ++a) {
      if (E.love && a == 1) {
        acc[a] = 0.0;
        continue;
      }
      const double *r = src[a];
      auto R = [&](int layer) { return r[((size_t)el * N + layer) * nchunk + nl]; };
      double tup = 0.0, tdwn = 0.0, d;
      if (!water) {
        int idup = (i - 1) * ns, iddw = i * ns;
        for (int j = 0; j < ns; ++j) {
          double tmp = (2. * j + 1.0) / (2. * ns);
          if (idup + j >= 0) tup += R(idup + j) * tmp;
          if (iddw + j < rmax) tdwn += R(iddw + j) * (1.0 - tmp);
          if (iddw == rmax - 1) tdwn += R(rmax - 1);
        }
        d = 0.0 + (tup + tdwn);
      } else {
        int idup = (i - 2) * ns + 1, iddw = (i - 1) * ns + 1;
        for (int j = 0; j < ns; ++j) {
          double tmp = (2. * j + 1.0) / (2. * ns);
          if (idup + j >= 1) tup += R(idup + j) * tmp;
          if (iddw + j < rmax) tdwn += R(iddw + j) * (1.0 - tmp);
          if (iddw == rmax - 1) tdwn += R(rmax - 1);
        }
        d = 0.0 + (tup + tdwn);
      }
      acc[a] = d;
    }
    size_t ko = (size_t)n + (size_t)i * nxy + (size_t)idx * nxy * nz;
    float v = node_vs[(size_t)i * nxy + n], p = node_vp[(size_t)i * nxy + n];
    float drda, dadb;
    brocher_deriv(p, v, drda, dadb);
    kernel[ko] = acc[0] + acc[1] * dadb + acc[2] * drda * dadb;
  }
}

// get_2d_map (SurfaceWave.cpp:420-450 over disp1D :177-256): one thread per
// (column of the chunk, SurfTime index).  vc is the root of the entry's
// chain taken to the sphere by _flat2sphere (surfdisp.cpp:17-51: R = 6371 km,
// the Rayleigh factor 0.5 for Rc/Rg and the Love factor 1.5 for Lc/Lg, phase
// form); vout is vc for Rc/Lc and U of the group job (groupvel_r/_l,
// surfdisp.cpp:126-174) for Rg/Lg.  status: the vc chain's (4 on the exit
// path, 1 without a root), else the group job's.
extern "C" __global__ void __launch_bounds__(256)
jdsg_surftime_map_epilogue(int col0, int nchunk, int nxy, int kmax, const int *__restrict__ cols,
                           const StMapEntry *__restrict__ ents, const StJob *__restrict__ jobs,
                           const double *__restrict__ tlist, double *vc, double *vout,
                           int *status) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nchunk * kmax) return;
  int nl = (int)(t % nchunk);
  int idx = (int)(t / nchunk);
  int n = cols[col0 + nl];
  StMapEntry E = ents[idx];
  const StJob &J = jobs[E.job_c];
  long long e = (long long)nl * (J.nR + J.nL) * J.K + (long long)E.mode * J.K + E.k;
  double c = J.cflat[e];
  int st = (J.cstat[e] & JDSG_EXIT_PATH) ? JDSG_EXIT_PATH : (c == 0.0 ? JDSG_NO_ROOT : 0);
  const double ar = 6371.0;
  double T = tlist[idx];
  double omega = 2.0 * M_PI / T;
  double x = (E.love ? 1.5 : 0.5) * c / (ar * omega);
  double tm = sqrt(1. + x * x);
  double cs = c / tm;
  size_t o = (size_t)n + (size_t)idx * nxy;
  vc[o] = cs;
  double v = cs;
  if (E.job_u >= 0) {
    const StJob &G = jobs[E.job_u];
    long long eg = (long long)nl * (G.nR + G.nL) * G.K + (long long)E.mode * G.K + E.k;
    v = G.u[eg];
    if (st == 0) st = G.status[eg];
  }
  vout[o] = v;
  status[o] = st;
}

void jdsg_launch_water_mask(int nxy, const float *vs, int *mask, hipStream_t st) {
  hipLaunchKernelGGL(jdsg_water_mask, dim3((nxy + 255) / 256), dim3(256), 0, st, nxy, vs, mask);
}

void jdsg_launch_grid_to_layers(int col0, int nchunk, int nxy, int nz, int N, int water,
                                const int *cols, const float *vs, const float *dep, float *thk,
                                float *vp, float *vsl, float *rho, float *vp2, float *node_vp,
                                float *node_vs, hipStream_t st) {
  hipLaunchKernelGGL(jdsg_grid_to_layers, dim3((nchunk + 255) / 256), dim3(256), 0, st, col0,
                     nchunk, nxy, nz, N, water, cols, vs, dep, thk, vp, vsl, rho, vp2, node_vp,
                     node_vs);
}

void jdsg_launch_surftime_epilogue(int col0, int nchunk, int nxy, int nz, int N, int water,
                                   int kmax, const int *cols, const void *ents, const void *jobs,
                                   const float *node_vp, const float *node_vs, double *vc,
                                   double *vout, double *kernel, int *status, int want_kernel,
                                   hipStream_t st) {
  long long n = (long long)nchunk * kmax;
  if (n <= 0) return;
  hipLaunchKernelGGL(jdsg_surftime_epilogue, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     st, col0, nchunk, nxy, nz, N, water, kmax, cols, (const StEntry *)ents,
                     (const StJob *)jobs, node_vp, node_vs, vc, vout, kernel, status,
                     want_kernel);
}

void jdsg_launch_surftime_map_epilogue(int col0, int nchunk, int nxy, int kmax, const int *cols,
                                       const void *ents, const void *jobs, const double *tlist,
                                       double *vc, double *vout, int *status, hipStream_t st) {
  long long n = (long long)nchunk * kmax;
  if (n <= 0) return;
  hipLaunchKernelGGL(jdsg_surftime_map_epilogue, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     st, col0, nchunk, nxy, kmax, cols, (const StMapEntry *)ents,
                     (const StJob *)jobs, tlist, vc, vout, status);
}
