// swd_smooth.hip -- model smoothing and regularisation rows on the GPU:
// the Gaussian smoothing of a gradient by explicit diffusion steps
// (src/shared/smoothing.cpp: smooth_cart_pde :177-249, smooth_sph_pde
// :256-362), the depth resampling around it (interp_irregular_z :122-170),
// the driver DSurfTomo::smoothing (src/surftomo/surftomo.cpp:147-178) and the
// Laplacian rows appended to the Frechet matrix (add_regularization,
// src/shared/csr_matrix.cpp:268-328).
//
// Arrays are column-major (x fastest), float, in device memory.  Each
// diffusion step is one launch over the interior nodes of a padded copy (the
// ghost nodes keep the clamped edge values of the input, as in the
// reference); the stencil sums are taken in double in the reference's order
// with contraction off.  The finite-difference coefficients are the exact
// central-difference rationals rounded to float.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace {

struct Coefs9 {
  float c[9];
};
struct Coefs7 {
  float c[7];
};

// central second derivative, 8th order; second and first derivative, 6th order
Coefs9 cart_coefs() {
  const double r[9] = {-1.0 / 560, 8.0 / 315, -1.0 / 5, 8.0 / 5, -205.0 / 72,
                       8.0 / 5,    -1.0 / 5,  8.0 / 315, -1.0 / 560};
  Coefs9 k;
  for (int i = 0; i < 9; ++i) k.c[i] = (float)r[i];
  return k;
}
Coefs7 sph_coefs2() {
  const double r[7] = {1.0 / 90, -3.0 / 20, 3.0 / 2, -49.0 / 18, 3.0 / 2, -3.0 / 20, 1.0 / 90};
  Coefs7 k;
  for (int i = 0; i < 7; ++i) k.c[i] = (float)r[i];
  return k;
}
Coefs7 sph_coefs1() {
  const double r[7] = {-1.0 / 60, 3.0 / 20, -3.0 / 4, 0.0, 3.0 / 4, -3.0 / 20, 1.0 / 60};
  Coefs7 k;
  for (int i = 0; i < 7; ++i) k.c[i] = (float)r[i];
  return k;
}

__device__ __forceinline__ int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// p = q = the input with n2 ghost nodes a side holding the nearest edge value
__global__ void smooth_pad(const float *__restrict__ g, float *p, float *q, int nx, int ny, int nz,
                           int n2) {
  const int nx1 = nx + 2 * n2, ny1 = ny + 2 * n2, nz1 = nz + 2 * n2;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nx1 * ny1 * nz1) return;
  const int i = (int)(t % nx1), j = (int)((t / nx1) % ny1), k = (int)(t / ((long long)nx1 * ny1));
  const float v = g[((long long)clampi(k - n2, nz) * ny + clampi(j - n2, ny)) * nx +
                    clampi(i - n2, nx)];
  p[t] = v;
  q[t] = v;
}

// g = the interior of the padded array
__global__ void smooth_unpad(const float *__restrict__ p, float *g, int nx, int ny, int nz,
                             int n2) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nx * ny * nz) return;
  const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((long long)nx * ny));
  const int nx1 = nx + 2 * n2, ny1 = ny + 2 * n2;
  g[t] = p[((long long)(k + n2) * ny1 + (j + n2)) * nx1 + (i + n2)];
}

// one step of smooth_cart_pde (smoothing.cpp:228-241) at every interior node
__global__ void __launch_bounds__(256)
smooth_cart_step(const float *__restrict__ old, float *__restrict__ nw, int nx, int ny, int nz,
                 float Ch, float Cv, Coefs9 K) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nx * ny * nz) return;
  const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((long long)nx * ny));
  const long long sy = nx + 8, sz = (long long)(nx + 8) * (ny + 8);
  const long long o = (k + 4) * sz + (j + 4) * sy + (i + 4);
  double tx = 0., ty = 0., tz = 0.;
#pragma unroll
  for (int l = -4; l <= 4; ++l) {
    tx += (double)(K.c[l + 4] * old[o + l]);
    ty += (double)(K.c[l + 4] * old[o + l * sy]);
    tz += (double)(K.c[l + 4] * old[o + l * sz]);
  }
  double r = (double)old[o] + (double)Ch * tx;
  r = r + (double)Ch * ty;
  r = r + (double)Cv * tz;
  nw[o] = (float)r;
}

struct SphGeo {
  float Ch, Cv, dr, dth, dphi;
};

// one step of smooth_sph_pde (smoothing.cpp:328-353); radi over the padded
// depths, sth / cth over the padded colatitudes
__global__ void __launch_bounds__(256)
smooth_sph_step(const float *__restrict__ old, float *__restrict__ nw, int nx, int ny, int nz,
                SphGeo G, const float *__restrict__ radi, const float *__restrict__ sth,
                const float *__restrict__ cth, Coefs7 K2, Coefs7 K1) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nx * ny * nz) return;
  const int i0 = (int)(t % nx), j0 = (int)((t / nx) % ny), k0 = (int)(t / ((long long)nx * ny));
  const int i = i0 + 3, k = k0 + 3;
  const long long sy = nx + 6, sz = (long long)(nx + 6) * (ny + 6);
  const long long o = k * sz + (j0 + 3) * sy + i;
  double tx1 = 0., tx2 = 0., ty = 0., tz = 0.;
#pragma unroll
  for (int l = -3; l <= 3; ++l) {
    tx2 += (double)(K2.c[l + 3] * old[o + l]);
    tx1 += (double)(K1.c[l + 3] * old[o + l]);
    ty += (double)(K2.c[l + 3] * old[o + l * sy]);
    tz += (double)(K2.c[l + 3] * radi[k + l] * old[o + l * sz]);
  }
  const float rk = radi[k], s = sth[i];
  const float rr = rk * rk, dth2 = G.dth * G.dth, dphi2 = G.dphi * G.dphi, dr2 = G.dr * G.dr;
  const float rrs = rk * rk * s, rs = rk * s;
  const double t1 = (double)G.Ch * tx2 / (double)rr / (double)dth2;
  const double t2 = (double)G.Ch * tx1 / (double)G.dth * (double)cth[i] / (double)rrs;
  const double t3 = (double)G.Ch * ty / (double)dphi2 / ((double)rs * (double)rs);
  const double t4 = (double)G.Cv * tz / (double)rk / (double)dr2;
  double r = (double)old[o] + t1;
  r = r + t2;
  r = r + t3;
  r = r + t4;
  nw[o] = (float)r;
}

// out[k] = in[k0[k]] + coef[k] (in[k0[k]+1] - in[k0[k]]); k0[k] = -1 - s
// marks an end plane, a copy of in[s]
__global__ void smooth_lerp_z(const float *__restrict__ in, float *out, long long nxy, int nzo,
                              const int *__restrict__ k0, const float *__restrict__ coef) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nxy * nzo) return;
  const int k = (int)(t / nxy);
  const long long q = t % nxy;
  if (k0[k] < 0) {
    out[t] = in[(long long)(-1 - k0[k]) * nxy + q];
    return;
  }
  const float a = in[(long long)k0[k] * nxy + q];
  const float b = in[(long long)(k0[k] + 1) * nxy + q];
  const float d = b - a, e = coef[k] * d;
  out[t] = a + e;
}

// interior nodes (no index on a face) of the box before node (i, j, k) in
// x-fastest order
__device__ __forceinline__ int clamp0(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

__device__ __forceinline__ long long interior_before(int i, int j, int k, int nx, int ny, int nz) {
  const long long ix = nx > 2 ? nx - 2 : 0, iy = ny > 2 ? ny - 2 : 0;
  long long c = (long long)clamp0(k - 1, nz > 2 ? nz - 2 : 0) * ix * iy;
  if (k >= 1 && k <= nz - 2) {
    c += (long long)clamp0(j - 1, ny > 2 ? ny - 2 : 0) * ix;
    if (j >= 1 && j <= ny - 2) c += clamp0(i - 1, nx > 2 ? nx - 2 : 0);
  }
  return c;
}

// the regularisation row of model node r (csr_matrix.cpp:277-327): one entry
// 2w on a face node, else 6w at the node and -w at its six neighbours
__global__ void reg_rows(int m, int *ptr, int *col, float *val, float w, int nx, int ny, int nz,
                         long long base) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)nx * ny * nz;
  if (r >= n) return;
  const int i = (int)(r % nx), j = (int)((r / nx) % ny), k = (int)(r / ((long long)nx * ny));
  const long long off = base + r + 6 * interior_before(i, j, k, nx, ny, nz);
  const int clc = (int)r;
  const bool edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1;
  if (edge) {
    val[off] = 2.0f * w;
    col[off] = clc;
    ptr[m + r + 1] = (int)(off + 1);
  } else {
    const int nxy = nx * ny;
    const int cc[7] = {clc, clc - 1, clc + 1, clc - nx, clc + nx, clc - nxy, clc + nxy};
    for (int q = 0; q < 7; ++q) {
      col[off + q] = cc[q];
      val[off + q] = q == 0 ? 6.0f * w : -w;
    }
    ptr[m + r + 1] = (int)(off + 7);
  }
}

struct Bufs {
  std::vector<void *> p;
  int err = 0;
  template <class T>
  T *get(size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      err = 1;
      return nullptr;
    }
    p.push_back(q);
    return (T *)q;
  }
  ~Bufs() {
    for (void *q : p) (void)hipFree(q);
  }
};

thread_local std::string g_err;
int fail(const char *m) {
  g_err = m;
  return 1;
}

dim3 grid1(long long n) { return dim3((unsigned)std::max<long long>(1, (n + 255) / 256)); }

// the explicit diffusion loop shared by both smoothers: pad, nstep steps
// (ping-pong), unpad; step(old, new) launches one step
template <class Step>
int diffuse(float *grad, int nx, int ny, int nz, int n2, int nstep, hipStream_t st, Step step) {
  const long long tot = (long long)(nx + 2 * n2) * (ny + 2 * n2) * (nz + 2 * n2);
  Bufs B;
  float *a = B.get<float>(tot), *b = B.get<float>(tot);
  if (B.err) return fail("smooth: out of device memory");
  hipLaunchKernelGGL(smooth_pad, grid1(tot), dim3(256), 0, st, grad, a, b, nx, ny, nz, n2);
  for (int it = 0; it < nstep; ++it) {
    step(a, b);
    std::swap(a, b);
  }
  hipLaunchKernelGGL(smooth_unpad, grid1((long long)nx * ny * nz), dim3(256), 0, st, a, grad, nx,
                     ny, nz, n2);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return fail("smooth: kernel launch failed");
  return 0;
}

int cart_pde(float *grad, int nx, int ny, int nz, float sigma_h, float sigma_v, hipStream_t st) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !grad) return fail("smooth: empty grid or null pointer");
  const Coefs9 K = cart_coefs();
  const float eta = (float)(1. / 12);
  const float sh2 = sigma_h * sigma_h, sv2 = sigma_v * sigma_v, mx = std::max(sh2, sv2);
  const int nstep = mx > 0.f ? (int)(mx / (2. * eta)) : 0;
  float Ch = mx > 0.f ? sh2 / mx : 0.f, Cv = mx > 0.f ? sv2 / mx : 0.f;
  Ch *= eta;
  Cv *= eta;
  const long long n = (long long)nx * ny * nz;
  return diffuse(grad, nx, ny, nz, 4, nstep, st, [&](const float *o, float *w) {
    hipLaunchKernelGGL(smooth_cart_step, grid1(n), dim3(256), 0, st, o, w, nx, ny, nz, Ch, Cv,
                       K);
  });
}

int sph_pde(float *grad, int nx, int ny, int nz, float dx, float dy, float dz, float lat0,
            float lon0, float z0, float sigma_h, float sigma_v, hipStream_t st) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !grad) return fail("smooth: empty grid or null pointer");
  const float R = 6371.f, deg2rad = (float)(M_PI / 180.);
  const int n2 = 3, nx1 = nx + 6, nz1 = nz + 6;
  std::vector<float> radi(nz1), theta(nx1), sth(nx1), cth(nx1);
  for (int iz = -n2; iz < nz + n2; ++iz) radi[iz + n2] = R - (iz * dz + z0);
  const float phi0 = deg2rad * (lon0 + (-n2 * dy)), phi1 = deg2rad * (lon0 + ((1 - n2) * dy));
  for (int ix = -n2; ix < nx + n2; ++ix) {
    // 90. is double: the colatitude is formed in double, stored as float
    theta[ix + n2] = (float)((double)deg2rad * (90. - (double)lat0 + (double)(ix * dx)));
    sth[ix + n2] = std::sin(theta[ix + n2]);
    cth[ix + n2] = std::cos(theta[ix + n2]);
  }
  SphGeo G;
  G.dr = radi[1] - radi[0];
  G.dth = theta[1] - theta[0];
  G.dphi = phi1 - phi0;
  float h2 = R * 1000;
  for (int k = 0; k < nz1; ++k)
    for (int i = 0; i < nx1; ++i) {
      const float d1 = std::fabs(G.dr), d2 = radi[k] * G.dth, d3 = radi[k] * sth[i] * G.dphi;
      h2 = std::min(h2, d1);
      h2 = std::min(h2, d2);
      h2 = std::min(h2, d3);
    }
  h2 = h2 * h2;
  const float eta = (float)(1. / 7);
  const float sh2 = sigma_h * sigma_h, sv2 = sigma_v * sigma_v, mx = std::max(sh2, sv2);
  const int nstep = mx > 0.f ? (int)(mx / (2. * eta * h2)) : 0;
  float Ch = mx > 0.f ? sh2 / mx : 0.f, Cv = mx > 0.f ? sv2 / mx : 0.f;
  Ch *= eta * h2;
  Cv *= eta * h2;
  G.Ch = Ch;
  G.Cv = Cv;
  Bufs B;
  float *d_r = B.get<float>(nz1), *d_s = B.get<float>(nx1), *d_c = B.get<float>(nx1);
  if (B.err) return fail("smooth: out of device memory");
  if (hipMemcpyAsync(d_r, radi.data(), sizeof(float) * nz1, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(d_s, sth.data(), sizeof(float) * nx1, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(d_c, cth.data(), sizeof(float) * nx1, hipMemcpyHostToDevice, st))
    return fail("smooth: copy failed");
  const Coefs7 K2 = sph_coefs2(), K1 = sph_coefs1();
  const long long n = (long long)nx * ny * nz;
  return diffuse(grad, nx, ny, nz, n2, nstep, st, [&](const float *o, float *w) {
    hipLaunchKernelGGL(smooth_sph_step, grid1(n), dim3(256), 0, st, o, w, nx, ny, nz, G, d_r, d_s,
                       d_c, K2, K1);
  });
}

int interp_z(float *grad, float *gradr, int nx, int ny, int nz, int nz1, const float *dep,
             bool fwd, hipStream_t st) {
  if (nx <= 0 || ny <= 0 || nz < 2 || nz1 < 2 || !grad || !gradr || !dep)
    return fail("interp_irregular_z: bad sizes or null pointer");
  const float dz = (dep[nz - 1] - dep[0]) / (nz1 - 1);
  const int nzo = fwd ? nz1 : nz;
  std::vector<int> k0(nzo);
  std::vector<float> coef(nzo);
  for (int k = 0; k < nzo; ++k) {
    const int last = fwd ? nz1 - 1 : nz - 1;
    if (k == 0 || k == last) {  // copy an end plane
      k0[k] = -1 - (k == 0 ? 0 : (fwd ? nz - 1 : nz1 - 1));
      coef[k] = 0.f;
      continue;
    }
    if (fwd) {
      const float z = dep[0] + k * dz;
      int k1 = 0;
      for (; k1 < nz - 1; ++k1)
        if (dep[k1] <= z && z < dep[k1 + 1]) break;
      k1 = std::min(k1, nz - 2);  // no interval found: the reference reads past dep
      k0[k] = k1;
      coef[k] = (z - dep[k1]) / (dep[k1 + 1] - dep[k1]);
    } else {
      int k1 = (int)((dep[k] - dep[0]) / dz);
      k1 = std::min(std::max(k1, 0), nz1 - 2);  // rounding at the top: the reference reads past
      const float base = dep[0] + k1 * dz;
      k0[k] = k1;
      coef[k] = (dep[k] - base) / dz;
    }
  }
  Bufs B;
  int *d_k = B.get<int>(nzo);
  float *d_c = B.get<float>(nzo);
  if (B.err) return fail("interp_irregular_z: out of device memory");
  if (hipMemcpyAsync(d_k, k0.data(), sizeof(int) * nzo, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(d_c, coef.data(), sizeof(float) * nzo, hipMemcpyHostToDevice, st))
    return fail("interp_irregular_z: copy failed");
  const long long nxy = (long long)nx * ny;
  hipLaunchKernelGGL(smooth_lerp_z, grid1(nxy * nzo), dim3(256), 0, st, fwd ? grad : gradr,
                     fwd ? gradr : grad, nxy, nzo, d_k, d_c);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return fail("interp_irregular_z: kernel launch failed");
  return 0;
}

}  // namespace

extern "C" const char *jdsg_smooth_last_error(void) { return g_err.c_str(); }

extern "C" int jdsg_smooth_cart_pde(float *grad, int nx, int ny, int nz, float sigma_h,
                                    float sigma_v, void *stream) {
  return cart_pde(grad, nx, ny, nz, sigma_h, sigma_v, (hipStream_t)stream);
}

extern "C" int jdsg_smooth_sph_pde(float *grad, int nx, int ny, int nz, float dx, float dy,
                                   float dz, float lat0, float lon0, float z0, float sigma_h,
                                   float sigma_v, void *stream) {
  return sph_pde(grad, nx, ny, nz, dx, dy, dz, lat0, lon0, z0, sigma_h, sigma_v,
                 (hipStream_t)stream);
}

extern "C" int jdsg_interp_irregular_z(float *grad, float *gradr, int nx, int ny, int nz, int nz1,
                                       const float *dep, int fwd, void *stream) {
  return interp_z(grad, gradr, nx, ny, nz, nz1, dep, fwd != 0, (hipStream_t)stream);
}

extern "C" int jdsg_smooth_gradient(float *grad, int nx, int ny, int nz, const float *dep,
                                    const float *lat, const float *lon, int smooth_in_km,
                                    float sigma_h, float sigma_v, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (nx < 3 || ny < 3 || nz < 3 || !grad || !dep || !lat || !lon)
    return fail("smooth_gradient: grid smaller than 3 nodes a side or null pointer");
  // surftomo.cpp:152-161: the finest depth step, a regular depth grid on it
  float dz = 1000 * (dep[nz - 2] - dep[0]);
  for (int i = 0; i < nz - 2; ++i) dz = std::min(dz, dep[i + 1] - dep[i]);
  if (!(dz > 0.f)) return fail("smooth_gradient: depths must increase");
  const int nz1 = (int)((dep[nz - 2] - dep[0]) / (dz * 0.99));
  if (nz1 < 2) return fail("smooth_gradient: fewer than 2 regular depths");
  dz = (dep[nz - 2] - dep[0]) / (nz1 - 1);
  const int mx = nx - 2, my = ny - 2;
  Bufs B;
  float *gr = B.get<float>((size_t)mx * my * nz1);
  if (B.err) return fail("smooth_gradient: out of device memory");
  if (interp_z(grad, gr, mx, my, nz - 1, nz1, dep, true, st)) return 1;
  int rc;
  if (smooth_in_km) {
    const float dx = std::fabs(lat[1] - lat[0]), dy = std::fabs(lon[1] - lon[0]);
    rc = sph_pde(gr, mx, my, nz1, dx, dy, dz, lat[1], lon[1], dep[0], sigma_h, sigma_v, st);
  } else {
    rc = cart_pde(gr, mx, my, nz1, sigma_h, sigma_v, st);
  }
  if (rc) return rc;
  return interp_z(grad, gr, mx, my, nz - 1, nz1, dep, false, st);
}

extern "C" int jdsg_add_regularization(int m, int *ptr, int *col, float *val, long long cap,
                                       float weight, int nx, int ny, int nz, long long *nnz,
                                       void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (m < 0 || nx <= 0 || ny <= 0 || nz <= 0 || !ptr || !col || !val)
    return fail("add_regularization: bad sizes or null pointer");
  int base = 0;
  if (hipMemcpyAsync(&base, ptr + m, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail("add_regularization: cannot read ptr[m]");
  const long long n = (long long)nx * ny * nz;
  const long long inner = (long long)std::max(nx - 2, 0) * std::max(ny - 2, 0) * std::max(nz - 2, 0);
  const long long total = base + n + 6 * inner;
  if (nnz) *nnz = total;
  if (base < 0 || total > cap) return fail("add_regularization: please increase sparse ratio");
  if (total > 0x7fffffffLL) return fail("add_regularization: more than 2^31 entries");
  hipLaunchKernelGGL(reg_rows, grid1(n), dim3(256), 0, st, m, ptr, col, val, weight, nx, ny, nz,
                     (long long)base);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return fail("add_regularization: kernel launch failed");
  return 0;
}
