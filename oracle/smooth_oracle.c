/* smooth_oracle.c -- test infrastructure (the oracle; never part of the
 * product): a plain-C restatement of JDSurfG's model smoothing,
 * src/shared/smoothing.cpp (fdcoefs :84-110, interp_irregular_z :122-170,
 * smooth_cart_pde :177-249, smooth_sph_pde :256-362) and the regularisation
 * rows of src/shared/csr_matrix.cpp:268-328 (add_regularization).
 *
 * Arrays are column-major (x fastest), float, as in the reference.  The
 * finite-difference coefficients are the exact central-difference rationals
 * rounded to float; the reference gets them from a float LU solve (Eigen),
 * which is absent here, so its coefficient bits are not reproduced. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "swd_oracle.h"

/* smoothing.cpp:84-110: central coefficients of the deriv-th derivative on
 * acc+1 points, from sum_j c_j (j-n)^i / i! = delta(i, deriv); solved in
 * double by Gaussian elimination with partial pivoting, then rounded */
/* test hook: coefficient sets used in place of the rationals by the
 * smoothers (acc 8 / deriv 2, acc 6 / deriv 2, acc 6 / deriv 1; null pointers
 * restore the rationals), so a test can run the reference's float-LU
 * coefficients through the same smoother */
static const float *g_fd_override[3];
void or_fdcoefs_override(const float *c82, const float *c62, const float *c61) {
  g_fd_override[0] = c82;
  g_fd_override[1] = c62;
  g_fd_override[2] = c61;
}

void or_fdcoefs(int acc, int deriv, float *coefs) {
  const float *ov = acc == 8 && deriv == 2 ? g_fd_override[0]
                    : acc == 6 && deriv == 2 ? g_fd_override[1]
                    : acc == 6 && deriv == 1 ? g_fd_override[2] : NULL;
  if (ov) {
    memcpy(coefs, ov, sizeof(float) * (acc + 1));
    return;
  }
  const int N = acc + 1, n = acc / 2;
  double A[17][18];
  for (int i = 0; i < N; ++i) {
    double f = 1.0;
    for (int t = 2; t <= i; ++t) f *= t;
    for (int j = 0; j < N; ++j) A[i][j] = pow((double)(j - n), i) / f;
    A[i][N] = i == deriv ? 1.0 : 0.0;
  }
  for (int c = 0; c < N; ++c) {
    int p = c;
    for (int r = c + 1; r < N; ++r)
      if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
    for (int j = 0; j <= N; ++j) {
      double t = A[c][j];
      A[c][j] = A[p][j];
      A[p][j] = t;
    }
    for (int r = 0; r < N; ++r) {
      if (r == c) continue;
      double f = A[r][c] / A[c][c];
      for (int j = c; j <= N; ++j) A[r][j] -= f * A[c][j];
    }
  }
  for (int i = 0; i < N; ++i) coefs[i] = (float)(A[i][N] / A[i][i]);
  if (deriv % 2 == 1) coefs[n] = 0.0f; /* :107 */
}

/* smoothing.cpp:122-170: between the irregular depths dep[0..nz) and nz1
 * regular ones; fwd fills gradr (nx*ny x nz1) from grad, else grad from gradr */
void or_interp_irregular_z(float *grad, float *gradr, int nx, int ny, int nz, int nz1,
                           const float *dep, int fwd) {
  const long long nxy = (long long)nx * ny;
  const float dz = (dep[nz - 1] - dep[0]) / (nz1 - 1);
  if (fwd) {
    for (int k = 0; k < nz1; ++k) {
      if (k == 0 || k == nz1 - 1) {
        int k1 = k == 0 ? 0 : nz - 1;
        memcpy(gradr + k * nxy, grad + k1 * nxy, sizeof(float) * nxy);
        continue;
      }
      float z = dep[0] + k * dz;
      int k1 = 0;
      for (; k1 < nz - 1; ++k1)
        if (dep[k1] <= z && z < dep[k1 + 1]) break;
      if (k1 > nz - 2) k1 = nz - 2; /* no interval found: the reference reads past dep */
      float coef = (z - dep[k1]) / (dep[k1 + 1] - dep[k1]);
      for (long long q = 0; q < nxy; ++q) {
        float a = grad[k1 * nxy + q], b = grad[(k1 + 1) * nxy + q];
        float d = b - a, e = coef * d;
        gradr[k * nxy + q] = a + e;
      }
    }
  } else {
    for (int k = 0; k < nz; ++k) {
      if (k == 0 || k == nz - 1) {
        int k1 = k == 0 ? 0 : nz1 - 1;
        memcpy(grad + k * nxy, gradr + k1 * nxy, sizeof(float) * nxy);
        continue;
      }
      int k1 = (int)((dep[k] - dep[0]) / dz);
      if (k1 > nz1 - 2) k1 = nz1 - 2; /* rounding at the top: the reference reads past gradr */
      float base = dep[0] + k1 * dz;
      float coef = (dep[k] - base) / dz;
      for (long long q = 0; q < nxy; ++q) {
        float a = gradr[k1 * nxy + q], b = gradr[(k1 + 1) * nxy + q];
        float d = b - a, e = coef * d;
        grad[k * nxy + q] = a + e;
      }
    }
  }
}

static int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

/* the padded copy both PDE smoothers start from: n2 ghost nodes a side
 * holding the nearest edge value (smoothing.cpp:199-207, :284-292) */
static float *padded(const float *g, int nx, int ny, int nz, int n2) {
  const int nx1 = nx + 2 * n2, ny1 = ny + 2 * n2, nz1 = nz + 2 * n2;
  float *p = malloc(sizeof(float) * (size_t)nx1 * ny1 * nz1);
  for (int k = -n2; k < nz + n2; ++k)
    for (int j = -n2; j < ny + n2; ++j)
      for (int i = -n2; i < nx + n2; ++i)
        p[((size_t)(k + n2) * ny1 + (j + n2)) * nx1 + (i + n2)] =
            g[((size_t)clampi(k, nz) * ny + clampi(j, ny)) * nx + clampi(i, nx)];
  return p;
}

/* smoothing.cpp:177-249 */
void or_smooth_cart_pde(float *grad, int nx, int ny, int nz, float sigma_h, float sigma_v) {
  float c[9];
  or_fdcoefs(8, 2, c);
  const int n2 = 4, nx1 = nx + 8, ny1 = ny + 8, nz1 = nz + 8;
  const size_t tot = (size_t)nx1 * ny1 * nz1, sx = 1, sy = nx1, sz = (size_t)nx1 * ny1;
  float *old = padded(grad, nx, ny, nz, n2), *nw = malloc(sizeof(float) * tot);
  memcpy(nw, old, sizeof(float) * tot);
  const float eta = (float)(1. / 12);
  const float sh2 = sigma_h * sigma_h, sv2 = sigma_v * sigma_v;
  const float mx = sh2 > sv2 ? sh2 : sv2;
  const int nstep = (int)(mx / (2. * eta));
  float Ch = sh2 / mx, Cv = sv2 / mx;
  Ch *= eta;
  Cv *= eta;
  for (int it = 0; it < nstep; ++it) {
    for (int k = n2; k < nz1 - n2; ++k)
      for (int j = n2; j < ny1 - n2; ++j)
        for (int i = n2; i < nx1 - n2; ++i) {
          const size_t o = k * sz + j * sy + i * sx;
          double tx = 0., ty = 0., tz = 0.;
          for (int l = -n2; l <= n2; ++l) {
            tx += (double)(c[l + n2] * old[o + l * (long)sx]);
            ty += (double)(c[l + n2] * old[o + l * (long)sy]);
            tz += (double)(c[l + n2] * old[o + l * (long)sz]);
          }
          double r = (double)old[o] + (double)Ch * tx;
          r = r + (double)Ch * ty;
          r = r + (double)Cv * tz;
          nw[o] = (float)r;
        }
    memcpy(old, nw, sizeof(float) * tot);
  }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        grad[((size_t)k * ny + j) * nx + i] = old[(k + n2) * sz + (j + n2) * sy + (i + n2)];
  free(old);
  free(nw);
}

/* the grid geometry of smooth_sph_pde (smoothing.cpp:294-316): radius,
 * colatitude and their sines/cosines on the padded grid, the smallest
 * node spacing squared and the step count; returns nstep */
int or_sph_geometry(int nx, int ny, int nz, float dx, float dy, float dz, float lat0, float lon0,
                    float z0, float sigma_h, float sigma_v, float *radi, float *theta,
                    float *sinth, float *costh, float *scal) {
  const float R = 6371.f, deg2rad = (float)(M_PI / 180.);
  const int n2 = 3, nx1 = nx + 6, ny1 = ny + 6, nz1 = nz + 6;
  float phi0 = 0.f, phi1 = 0.f;
  for (int iz = -n2; iz < nz + n2; ++iz) radi[iz + n2] = R - (iz * dz + z0);
  for (int iy = -n2; iy < -n2 + 2; ++iy) {
    float p = deg2rad * (lon0 + (iy * dy));
    if (iy == -n2) phi0 = p; else phi1 = p;
  }
  for (int ix = -n2; ix < nx + n2; ++ix) {
    /* 90. is double: the colatitude is formed in double, stored as float */
    theta[ix + n2] = (float)((double)deg2rad * (90. - (double)lat0 + (double)(ix * dx)));
    sinth[ix + n2] = sinf(theta[ix + n2]);
    costh[ix + n2] = cosf(theta[ix + n2]);
  }
  const float dr = radi[1] - radi[0], dth = theta[1] - theta[0], dphi = phi1 - phi0;
  float h2 = R * 1000;
  for (int k = 0; k < nz1; ++k)
    for (int i = 0; i < nx1; ++i) {
      float d1 = fabsf(dr), d2 = radi[k] * dth, d3 = radi[k] * sinth[i] * dphi;
      h2 = fminf(h2, d1);
      h2 = fminf(h2, d2);
      h2 = fminf(h2, d3);
    }
  (void)ny1;
  h2 = h2 * h2;
  const float eta = (float)(1. / 7);
  const float sh2 = sigma_h * sigma_h, sv2 = sigma_v * sigma_v;
  const float mx = sh2 > sv2 ? sh2 : sv2;
  const int nstep = (int)(mx / (2. * eta * h2));
  float Ch = sh2 / mx, Cv = sv2 / mx;
  Ch *= eta * h2;
  Cv *= eta * h2;
  scal[0] = Ch;
  scal[1] = Cv;
  scal[2] = dr;
  scal[3] = dth;
  scal[4] = dphi;
  return nstep;
}

/* smoothing.cpp:256-362 */
void or_smooth_sph_pde(float *grad, int nx, int ny, int nz, float dx, float dy, float dz,
                       float lat0, float lon0, float z0, float sigma_h, float sigma_v) {
  float c1[7], c2[7];
  or_fdcoefs(6, 2, c2);
  or_fdcoefs(6, 1, c1);
  const int n2 = 3, nx1 = nx + 6, ny1 = ny + 6, nz1 = nz + 6;
  float *radi = malloc(sizeof(float) * nz1), *theta = malloc(sizeof(float) * nx1),
        *sth = malloc(sizeof(float) * nx1), *cth = malloc(sizeof(float) * nx1), sc[5];
  const int nstep = or_sph_geometry(nx, ny, nz, dx, dy, dz, lat0, lon0, z0, sigma_h, sigma_v,
                                    radi, theta, sth, cth, sc);
  const float Ch = sc[0], Cv = sc[1], dr = sc[2], dth = sc[3], dphi = sc[4];
  const size_t tot = (size_t)nx1 * ny1 * nz1, sy = nx1, sz = (size_t)nx1 * ny1;
  float *old = padded(grad, nx, ny, nz, n2), *nw = malloc(sizeof(float) * tot);
  memcpy(nw, old, sizeof(float) * tot);
  for (int it = 0; it < nstep; ++it) {
    for (int k = n2; k < nz1 - n2; ++k)
      for (int j = n2; j < ny1 - n2; ++j)
        for (int i = n2; i < nx1 - n2; ++i) {
          const size_t o = k * sz + j * sy + i;
          double tx1 = 0., tx2 = 0., ty = 0., tz = 0.;
          for (int l = -n2; l <= n2; ++l) {
            tx2 += (double)(c2[l + n2] * old[o + l]);
            tx1 += (double)(c1[l + n2] * old[o + l]);
            ty += (double)(c2[l + n2] * old[o + l * (long)sy]);
            tz += (double)(c2[l + n2] * radi[k + l] * old[o + l * (long)sz]);
          }
          const float rr = radi[k] * radi[k], dth2 = dth * dth, dphi2 = dphi * dphi,
                      dr2 = dr * dr, rrs = radi[k] * radi[k] * sth[i], rs = radi[k] * sth[i];
          const double t1 = (double)Ch * tx2 / (double)rr / (double)dth2;
          const double t2 = (double)Ch * tx1 / (double)dth * (double)cth[i] / (double)rrs;
          const double t3 = (double)Ch * ty / (double)dphi2 / ((double)rs * (double)rs);
          const double t4 = (double)Cv * tz / (double)radi[k] / (double)dr2;
          double r = (double)old[o] + t1;
          r = r + t2;
          r = r + t3;
          r = r + t4;
          nw[o] = (float)r;
        }
    memcpy(old, nw, sizeof(float) * tot);
  }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        grad[((size_t)k * ny + j) * nx + i] = old[(k + n2) * sz + (j + n2) * sy + (i + n2)];
  free(old);
  free(nw);
  free(radi);
  free(theta);
  free(sth);
  free(cth);
}

/* csr_matrix.cpp:268-328: the n = nx*ny*nz regularisation rows appended
 * after m data rows; ptr[0..m] given, ptr[m+1..m+n] written, entries from
 * ptr[m] on; returns the new entry count */
long long or_add_regularization(int m, int *ptr, int *col, float *val, float weight, int nx,
                                int ny, int nz) {
  long long nar = ptr[m];
  int count = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int rwc = count + m, clc = k * nx * ny + j * nx + i;
        if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1) {
          val[nar] = 2.0f * weight;
          col[nar] = clc;
          ptr[rwc + 1] = ptr[rwc] + 1;
          nar += 1;
        } else {
          ptr[rwc + 1] = ptr[rwc] + 7;
          const int cc[7] = {clc, clc - 1, clc + 1, clc - nx, clc + nx, clc - nx * ny,
                             clc + nx * ny};
          for (int t = 0; t < 7; ++t) {
            val[nar + t] = t == 0 ? 6.0f * weight : -weight;
            col[nar + t] = cc[t];
          }
          nar += 7;
        }
        count += 1;
      }
  return nar;
}

/* surftomo.cpp:147-178 (DSurfTomo::smoothing): grad of shape (nx-2, ny-2,
 * nz-1) resampled to a regular depth grid, smoothed, resampled back; dep
 * (nz), lat (nx), lon (ny) are the model's axes; returns 0, or 1 when the
 * depths do not give a regular grid */
int or_smooth_gradient(float *grad, int nx, int ny, int nz, const float *dep, const float *lat,
                       const float *lon, int smooth_in_km, float sigma_h, float sigma_v) {
  float dz = 1000 * (dep[nz - 2] - dep[0]);
  for (int i = 0; i < nz - 2; ++i) dz = fminf(dz, dep[i + 1] - dep[i]);
  if (!(dz > 0.f)) return 1;
  const int nz1 = (int)((dep[nz - 2] - dep[0]) / (dz * 0.99));
  if (nz1 < 2) return 1;
  dz = (dep[nz - 2] - dep[0]) / (nz1 - 1);
  const int mx = nx - 2, my = ny - 2;
  float *gr = malloc(sizeof(float) * (size_t)mx * my * nz1);
  or_interp_irregular_z(grad, gr, mx, my, nz - 1, nz1, dep, 1);
  if (smooth_in_km) {
    const float dx = fabsf(lat[1] - lat[0]), dy = fabsf(lon[1] - lon[0]);
    or_smooth_sph_pde(gr, mx, my, nz1, dx, dy, dz, lat[1], lon[1], dep[0], sigma_h, sigma_v);
  } else {
    or_smooth_cart_pde(gr, mx, my, nz1, sigma_h, sigma_v);
  }
  or_interp_irregular_z(grad, gr, mx, my, nz - 1, nz1, dep, 0);
  free(gr);
  return 0;
}

/* spherical.cpp:9-26 delsph: great-circle distance (km, R = 6371) between
 * two (colatitude, longitude) points in radians, by the haversine formula;
 * in double here */
static double delsph_d(double c1, double l1, double c2, double l2) {
  const double dlat = c2 - c1, dlon = l2 - l1, la1 = M_PI * 0.5 - c1, la2 = M_PI * 0.5 - c2;
  const double a = sin(dlat * 0.5) * sin(dlat * 0.5) +
                   sin(dlon * 0.5) * sin(dlon * 0.5) * cos(la1) * cos(la2);
  return 6371.0 * 2.0 * atan2(sqrt(a), sqrt(1 - a));
}

/* smoothing.cpp:21-63 smooth_sph: the direct Gaussian convolution the PDE
 * smoother approximates, grad[n] = sum_m g[m] w(n, m) / sum_m w(n, m) with
 * w = exp(-(delta r_m)^2 / 2 sigma_h^2) exp(-(r_n - r_m)^2 / 2 sigma_v^2),
 * delta the angular distance; in double (a yardstick, not the reference's
 * float bits), the horizontal factor taken once per pair of columns */
void or_smooth_sph_conv(float *grad, int nx, int ny, int nz, float dx, float dy, float dz,
                        float lat0, float lon0, float z0, float sigma_h, float sigma_v) {
  const double R = 6371.0, d2r = M_PI / 180.0;
  const size_t nxy = (size_t)nx * ny, n = nxy * nz;
  float *g1 = malloc(sizeof(float) * n);
  memcpy(g1, grad, sizeof(float) * n);
  double *r = malloc(sizeof(double) * nz), *V = malloc(sizeof(double) * nz * nz);
  double *sk = malloc(sizeof(double) * nz), *sn = malloc(sizeof(double) * nz);
  double *h = malloc(sizeof(double) * nz);
  for (int k = 0; k < nz; ++k) r[k] = R - (z0 + k * dz);
  for (int k = 0; k < nz; ++k)
    for (int k1 = 0; k1 < nz; ++k1)
      V[k * nz + k1] = exp(-(r[k] - r[k1]) * (r[k] - r[k1]) * 0.5 / ((double)sigma_v * sigma_v));
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double c = d2r * (90. - lat0 + i * dx), l = d2r * (lon0 + j * dy);
      for (int k = 0; k < nz; ++k) sk[k] = sn[k] = 0.0;
      for (int j1 = 0; j1 < ny; ++j1)
        for (int i1 = 0; i1 < nx; ++i1) {
          const double del = delsph_d(c, l, d2r * (90. - lat0 + i1 * dx), d2r * (lon0 + j1 * dy)) / R;
          for (int k1 = 0; k1 < nz; ++k1)
            h[k1] = exp(-(del * r[k1]) * (del * r[k1]) * 0.5 / ((double)sigma_h * sigma_h));
          for (int k = 0; k < nz; ++k)
            for (int k1 = 0; k1 < nz; ++k1) {
              const double w = h[k1] * V[k * nz + k1];
              sk[k] += g1[k1 * nxy + (size_t)j1 * nx + i1] * w;
              sn[k] += w;
            }
        }
      for (int k = 0; k < nz; ++k) grad[k * nxy + (size_t)j * nx + i] = (float)(sk[k] / sn[k]);
    }
  free(g1);
  free(r);
  free(V);
  free(sk);
  free(sn);
  free(h);
}
