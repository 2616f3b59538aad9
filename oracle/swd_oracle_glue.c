/*
 * swd_oracle_glue.c -- CPU restatement of the SurfTime column glue in
 * src/SWD/SurfaceWave.cpp (grid -> layers, per-column call sequence,
 * layer -> grid kernels, Brocher chain rule, OpenMP column loop) and of the
 * groupvel_* wrappers in src/SWD/surfdisp.cpp.  TEST INFRASTRUCTURE ONLY.
 *
 * The glue calls its numerics through an or_backend table so the same glue
 * drives either this restatement or the compiled reference (oracle/_ref).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "swd_oracle.h"

static int rest_surfdisp(const float *thk, const float *vp, const float *vs,
                         const float *rho, int nlayer, const double *t,
                         double *cg, int kmax, int swdtp, int mode, int sphere,
                         int keep_flat) {
  return or_surfdisp(thk, vp, vs, rho, nlayer, t, cg, kmax, swdtp, mode,
                     sphere, keep_flat);
}

static const or_backend REST = {rest_surfdisp, or_sregn96, or_sregnpu,
                                or_slegn96, or_slegnpu};

const or_backend *or_restatement_backend(void) { return &REST; }

/* SurfaceWave.cpp:113-155 grid2LayerModel */
int or_grid2layer(const float *dep, const float *vp, const float *vs,
                  const float *rho, int nz, int sublayers, float *rthk,
                  float *rvp, float *rvs, float *rrho) {
  int n = sublayers + 1;
  int rmax, istart;
  if (vs[0] == 0.0f) {
    rmax = nz - 1 + (nz - 2) * sublayers + 1;
    istart = 1;
    rthk[0] = dep[1] - dep[0];
    rrho[0] = rho[0];
    rvs[0] = 0.0f;
    rvp[0] = vp[0];
  } else {
    rmax = nz + (nz - 1) * sublayers;
    istart = 0;
  }
  int k = istart;
  for (int i = istart; i < nz - 1; ++i) {
    float thk = (dep[i + 1] - dep[i]) / n;
    for (int j = 0; j < n; ++j) {
      rthk[k] = thk;
      rvp[k] = (float)(vp[i] + (float)(2 * j + 1) * (vp[i + 1] - vp[i]) / (2.0 * n));
      rvs[k] = (float)(vs[i] + (float)(2 * j + 1) * (vs[i + 1] - vs[i]) / (2.0 * n));
      rrho[k] = (float)(rho[i] + (float)(2 * j + 1) * (rho[i + 1] - rho[i]) / (2.0 * n));
      k++;
    }
  }
  k = rmax - 1;
  rthk[k] = 0.0f;
  rvp[k] = vp[nz - 1];
  rvs[k] = vs[nz - 1];
  rrho[k] = rho[nz - 1];
  return rmax;
}

/* SurfaceWave.cpp:59-101 _ParamConvert (only vs[0] of the grid is read) */
void or_param_convert(const float *vs, int nz, int sublayer, int rmax,
                      const double *rdcdm, double *dcdm) {
  int n = sublayer + 1;
  if (vs[0] != 0.0f) {
    for (int i = 0; i < nz; ++i) {
      double tup = 0.0, tdwn = 0.0;
      int idup = (i - 1) * n, iddw = i * n;
      dcdm[i] = 0.0;
      for (int j = 0; j < n; ++j) {
        double tmp = (2. * j + 1.0) / (2. * n);
        if (idup + j >= 0) tup += rdcdm[idup + j] * tmp;
        if (iddw + j < rmax) tdwn += rdcdm[iddw + j] * (1.0 - tmp);
        if (iddw == rmax - 1) tdwn += rdcdm[rmax - 1];
      }
      dcdm[i] += tup + tdwn;
    }
  } else {
    dcdm[0] = rdcdm[0];
    for (int i = 1; i < nz; ++i) {
      dcdm[i] = 0.0;
      double tup = 0.0, tdwn = 0.0;
      int idup = (i - 2) * n + 1, iddw = (i - 1) * n + 1;
      for (int j = 0; j < n; ++j) {
        double tmp = (2. * j + 1.0) / (2. * n);
        if (idup + j >= 1) tup += rdcdm[idup + j] * tmp;
        if (iddw + j < rmax) tdwn += rdcdm[iddw + j] * (1.0 - tmp);
        if (iddw == rmax - 1) tdwn += rdcdm[rmax - 1];
      }
      dcdm[i] += tup + tdwn;
    }
  }
}

/* SurfaceWave.cpp:157-168 get_nt_row */
static int nt_row(const or_tables *tb, int w, int mode) {
  const double *col = tb->t[w] + (size_t)mode * tb->ntmax[w];
  int n;
  for (n = 0; n < tb->ntmax[w]; ++n)
    if (col[n] == -1) break;
  return n;
}

static int tables_kmax(const or_tables *tb) {
  int k = 0;
  for (int w = 0; w < 4; ++w)
    for (int m = 0; m < tb->nmode[w]; ++m) k += nt_row(tb, w, m);
  return k;
}

/* surfdisp.cpp:156-174 groupvel_r */
static int groupvel_r(const or_backend *be, const float *thk, const float *vp,
                      const float *vs, const float *rho, int nlayer,
                      const double *t, double *cg, int kmax, int mode,
                      int sphere) {
  int iflsph = sphere ? 1 : 0;
  double *cp = (double *)malloc(sizeof(double) * (kmax + 8 * (size_t)nlayer + 1));
  double *w = cp + kmax;
  int err = be->surfdisp(thk, vp, vs, rho, nlayer, t, cp, kmax, 0, mode, sphere, 1);
  for (int i = 0; i < kmax; ++i)
    be->sregn96(thk, vp, vs, rho, nlayer, t + i, cp + i, cg + i, w, w + nlayer,
                w + 2 * nlayer, w + 3 * nlayer, w + 4 * nlayer, w + 5 * nlayer,
                w + 6 * nlayer, w + 7 * nlayer, iflsph);
  free(cp);
  return err;
}

/* surfdisp.cpp:126-154 groupvel_l (Vp is replaced by 1.732*Vs) */
static int groupvel_l(const or_backend *be, const float *thk, const float *vs,
                      const float *rho, int nlayer, const double *t,
                      double *cg, int kmax, int mode, int sphere) {
  int iflsph = sphere ? 1 : 0;
  float *vp = (float *)malloc(sizeof(float) * (size_t)nlayer);
  double *cp = (double *)malloc(sizeof(double) * (kmax + 5 * (size_t)nlayer + 1));
  double *w = cp + kmax;
  for (int i = 0; i < nlayer; ++i) vp[i] = (float)(1.732 * vs[i]);
  int err = be->surfdisp(thk, vp, vs, rho, nlayer, t, cp, kmax, 2, mode, sphere, 1);
  for (int i = 0; i < kmax; ++i)
    be->slegn96(thk, vs, rho, nlayer, &t[i], cp + i, cg + i, w, w + nlayer,
                w + 2 * nlayer, w + 3 * nlayer, w + 4 * nlayer, iflsph);
  free(cp);
  free(vp);
  return err;
}

/* SurfaceWave.cpp:177-256 disp1D */
int or_disp1d(const or_backend *be, const float *vs, const float *vp,
              const float *rho, const float *dep, int nz, const or_tables *tb,
              double *vc, double *vout) {
  const int sublayer = 3;
  int rmax = nz + (nz - 1) * sublayer;
  float *r = (float *)malloc(sizeof(float) * 4 * (size_t)rmax);
  float *rvp = r, *rvs = r + rmax, *rthk = r + 2 * rmax, *rrho = r + 3 * rmax;
  rmax = or_grid2layer(dep, vp, vs, rho, nz, sublayer, rthk, rvp, rvs, rrho);
  int err = 0, idx = 0;
  for (int w = 0; w < 4; ++w) {
    for (int mode = 0; mode < tb->nmode[w]; ++mode) {
      int nt = nt_row(tb, w, mode);
      if (nt <= 0) continue;
      const double *t = tb->t[w] + (size_t)mode * tb->ntmax[w];
      int code = (w < 2) ? 0 : 2;
      err |= be->surfdisp(rthk, rvp, rvs, rrho, rmax, t, vc + idx, nt, code,
                          mode, 1, 0);
      if (w == 0 || w == 2)
        memcpy(vout + idx, vc + idx, nt * sizeof(double));
      else if (w == 1)
        err |= groupvel_r(be, rthk, rvp, rvs, rrho, rmax, t, vout + idx, nt,
                          mode, 1);
      else
        err |= groupvel_l(be, rthk, rvs, rrho, rmax, t, vout + idx, nt, mode, 1);
      idx += nt;
    }
  }
  free(r);
  return err;
}

/* SurfaceWave.cpp:266-412 kernel1D; kvs/kvp/krho are (nz, kmax) col-major */
int or_kernel1d(const or_backend *be, const float *vs, const float *vp,
                const float *rho, const float *dep, int nz,
                const or_tables *tb, double *vc, double *vout, double *kvs,
                double *kvp, double *krho) {
  const int sublayer = 3;
  int rmax = nz + (nz - 1) * sublayer;
  float *r = (float *)malloc(sizeof(float) * 4 * (size_t)rmax);
  float *rvp = r, *rvs = r + rmax, *rthk = r + 2 * rmax, *rrho = r + 3 * rmax;
  rmax = or_grid2layer(dep, vp, vs, rho, nz, sublayer, rthk, rvp, rvs, rrho);
  double *wk = (double *)malloc(sizeof(double) * 12 * (size_t)rmax);
  double *ur = wk, *uz = wk + rmax, *tr = wk + 2 * rmax, *tz = wk + 3 * rmax;
  double *dudar = wk + 4 * rmax, *dudbr = wk + 5 * rmax, *dudrr = wk + 6 * rmax,
         *dudhr = wk + 7 * rmax;
  double *dcdar = wk + 8 * rmax, *dcdbr = wk + 9 * rmax, *dcdrr = wk + 10 * rmax,
         *dcdhr = wk + 11 * rmax;
  const int iflsph = 1;
  int err = 0, idx = 0;
  /* Rc */
  for (int mode = 0; mode < tb->nmode[0]; ++mode) {
    int nt = nt_row(tb, 0, mode);
    if (nt <= 0) continue;
    const double *t = tb->t[0] + (size_t)mode * tb->ntmax[0];
    err |= be->surfdisp(rthk, rvp, rvs, rrho, rmax, t, vc + idx, nt, 0, mode, 1, 1);
    for (int i = 0; i < nt; ++i) {
      int k = i + idx;
      double cg;
      be->sregn96(rthk, rvp, rvs, rrho, rmax, t + i, vc + k, &cg, ur, uz, tr,
                  tz, dcdar, dcdbr, dcdhr, dcdrr, iflsph);
      or_param_convert(vs, nz, sublayer, rmax, dcdar, kvp + (size_t)k * nz);
      or_param_convert(vs, nz, sublayer, rmax, dcdbr, kvs + (size_t)k * nz);
      or_param_convert(vs, nz, sublayer, rmax, dcdrr, krho + (size_t)k * nz);
      vout[k] = vc[k];
    }
    idx += nt;
  }
  /* Rg and Lg share the three-chain form */
  for (int pass = 0; pass < 3; ++pass) {
    int w = (pass == 0) ? 1 : (pass == 1 ? 2 : 3);
    for (int mode = 0; mode < tb->nmode[w]; ++mode) {
      int nt = nt_row(tb, w, mode);
      if (nt <= 0) continue;
      const double *t = tb->t[w] + (size_t)mode * tb->ntmax[w];
      if (w == 2) { /* Lc */
        err |= be->surfdisp(rthk, rvp, rvs, rrho, rmax, t, vc + idx, nt, 2, mode, 1, 1);
        for (int i = 0; i < nt; ++i) {
          int k = i + idx;
          double cg;
          be->slegn96(rthk, rvs, rrho, rmax, t + i, vc + k, &cg, uz, tz, dcdbr,
                      dcdhr, dcdrr, iflsph);
          or_param_convert(vs, nz, sublayer, rmax, dcdbr, kvs + (size_t)k * nz);
          or_param_convert(vs, nz, sublayer, rmax, dcdrr, krho + (size_t)k * nz);
          vout[k] = vc[k];
          for (int iz = 0; iz < nz; ++iz) kvp[(size_t)k * nz + iz] = 0.;
        }
        idx += nt;
        continue;
      }
      double *cp = (double *)malloc(sizeof(double) * 5 * (size_t)nt);
      double *t1 = cp + nt, *t2 = cp + 2 * nt, *c1 = cp + 3 * nt, *c2 = cp + 4 * nt;
      double dt = 0.01;
      for (int i = 0; i < nt; ++i) {
        t1[i] = t[i] * (1.0 + 0.5 * dt);
        t2[i] = t[i] * (1.0 - 0.5 * dt);
      }
      int code = (w == 1) ? 0 : 2;
      err |= be->surfdisp(rthk, rvp, rvs, rrho, rmax, t, cp, nt, code, mode, 1, 1);
      err |= be->surfdisp(rthk, rvp, rvs, rrho, rmax, t1, c1, nt, code, mode, 1, 1);
      err |= be->surfdisp(rthk, rvp, rvs, rrho, rmax, t2, c2, nt, code, mode, 1, 1);
      for (int i = 0; i < nt; ++i) {
        int k = i + idx;
        if (w == 1) {
          be->sregnpu(rthk, rvp, rvs, rrho, rmax, t + i, cp + i, vc + k, ur, uz,
                      tr, tz, t1 + i, c1 + i, t2 + i, c2 + i, dcdar, dcdbr,
                      dcdhr, dcdrr, dudar, dudbr, dudhr, dudrr, iflsph);
          or_param_convert(vs, nz, sublayer, rmax, dudar, kvp + (size_t)k * nz);
        } else {
          be->slegnpu(rthk, rvs, rrho, rmax, t + i, cp + i, vc + k, uz, tz,
                      t1 + i, c1 + i, t2 + i, c2 + i, dcdbr, dcdhr, dcdrr,
                      dudbr, dudhr, dudrr, iflsph);
          for (int iz = 0; iz < nz; ++iz) kvp[(size_t)k * nz + iz] = 0.;
        }
        or_param_convert(vs, nz, sublayer, rmax, dudbr, kvs + (size_t)k * nz);
        or_param_convert(vs, nz, sublayer, rmax, dudrr, krho + (size_t)k * nz);
        vout[k] = vc[k];
        vc[k] = cp[i];
      }
      free(cp);
      idx += nt;
    }
  }
  free(wk);
  free(r);
  return err;
}

/* SurfaceWave.cpp:458-502 get_1d_kernel */
int or_get_1d_kernel(const or_backend *be, const float *vs, int nx, int ny,
                     int nz, const float *dep, const or_tables *tb,
                     double *vc, double *vout, double *kernel, int nthreads) {
  int nxy = nx * ny;
  int nt = tables_kmax(tb);
  int nerr = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : nerr)
#endif
  for (int n = 0; n < nxy; ++n) {
    int i = n % nx, j = n / nx;
    float *v = (float *)malloc(sizeof(float) * 3 * (size_t)nz);
    float *vp = v + nz, *rho = v + 2 * nz;
    for (int k = 0; k < nz; ++k) {
      v[k] = vs[i + (size_t)j * nx + (size_t)k * nxy];
      or_empirical_relation(v[k], &vp[k], &rho[k]);
    }
    double *buf = (double *)calloc((size_t)nz * nt * 3 + 2 * (size_t)nt, sizeof(double));
    double *kvs = buf, *kvp = buf + (size_t)nz * nt, *krho = buf + 2 * (size_t)nz * nt;
    double *cp = buf + 3 * (size_t)nz * nt, *cg = cp + nt;
    nerr += or_kernel1d(be, v, vp, rho, dep, nz, tb, cp, cg, kvs, kvp, krho) ? 1 : 0;
    for (int ii = 0; ii < nt; ++ii) {
      vc[n + (size_t)ii * nxy] = cp[ii];
      vout[n + (size_t)ii * nxy] = cg[ii];
    }
    for (int ii = 0; ii < nt; ++ii) {
      for (int jj = 0; jj < nz; ++jj) {
        if (v[jj] == 0.0f && jj == 0) continue;
        float drda, dadb;
        or_empirical_deriv(vp[jj], v[jj], &drda, &dadb);
        size_t q = (size_t)jj + (size_t)ii * nz;
        kernel[n + (size_t)jj * nxy + (size_t)ii * nxy * nz] =
            kvs[q] + kvp[q] * dadb + krho[q] * drda * dadb;
      }
    }
    free(buf);
    free(v);
  }
  return nerr;
}

/* SurfaceWave.cpp:420-450 get_2d_map */
int or_get_2d_map(const or_backend *be, const float *vs, int nx, int ny,
                  int nz, const float *dep, const or_tables *tb, double *vc,
                  double *vout, int nthreads) {
  int nxy = nx * ny;
  int nt = tables_kmax(tb);
  int nerr = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : nerr)
#endif
  for (int n = 0; n < nxy; ++n) {
    int i = n % nx, j = n / nx;
    float *v = (float *)malloc(sizeof(float) * 3 * (size_t)nz);
    float *vp = v + nz, *rho = v + 2 * nz;
    for (int k = 0; k < nz; ++k) {
      v[k] = vs[i + (size_t)j * nx + (size_t)k * nxy];
      or_empirical_relation(v[k], &vp[k], &rho[k]);
    }
    double *c = (double *)calloc(2 * (size_t)nt, sizeof(double));
    nerr += or_disp1d(be, v, vp, rho, dep, nz, tb, c, c + nt) ? 1 : 0;
    for (int ii = 0; ii < nt; ++ii) {
      vc[n + (size_t)ii * nxy] = c[ii];
      vout[n + (size_t)ii * nxy] = c[nt + ii];
    }
    free(c);
    free(v);
  }
  return nerr;
}

/* ---------------- layer-stack batch ---------------- */

void or_batch_offsets(const or_batch *b, long long *ent_off,
                      long long *kern_off) {
  long long e = 0, q = 0;
  for (int c = 0; c < b->ncol; ++c) {
    ent_off[c] = e;
    kern_off[c] = q;
    long long K = b->per_off[c + 1] - b->per_off[c];
    long long N = b->layer_off[c + 1] - b->layer_off[c];
    long long ne = (long long)(b->nmode[2 * c] + b->nmode[2 * c + 1]) * K;
    e += ne;
    q += ne * N;
  }
  ent_off[b->ncol] = e;
  kern_off[b->ncol] = q;
}

static void zero_entry(const or_batch_out *o, long long e, long long q, int N,
                       int status) {
  o->c[e] = 0.0;
  o->u[e] = 0.0;
  o->status[e] = status;
  double *arr[8] = {o->dcda, o->dcdb, o->dcdr, o->dcdh,
                    o->duda, o->dudb, o->dudr, o->dudh};
  for (int a = 0; a < 8; ++a)
    if (arr[a]) memset(arr[a] + q, 0, sizeof(double) * (size_t)N);
}

/* One column of or_batch_solve.  With roots non-null ([3*nent], T, T(1+.005),
 * T(1-.005) flat roots as or_batch_out.cflat holds them) the root chains are
 * not run: the eigen stage is evaluated at those roots (status from the
 * roots alone: 0 where all needed roots are nonzero, 1 where the T root is 0,
 * 2 where a T(1+-.005) root is 0). */
static void solve_column(const or_backend *be, const or_batch *b,
                         const or_batch_out *o, int col, const double *roots) {
  int l0 = b->layer_off[col], N = b->layer_off[col + 1] - l0;
  int p0 = b->per_off[col], K = b->per_off[col + 1] - p0;
  const float *thk = b->thk + l0, *vp = b->vp + l0, *vs = b->vs + l0,
              *rho = b->rho + l0;
  const double *T = b->periods + p0;
  int nR = b->nmode[2 * col], nL = b->nmode[2 * col + 1];
  double *cbuf = (double *)malloc(sizeof(double) * 5 * (size_t)(K + 1));
  double *cp = cbuf, *c1 = cbuf + K, *c2 = cbuf + 2 * K, *t1 = cbuf + 3 * K,
         *t2 = cbuf + 4 * K;
  double *wk = (double *)malloc(sizeof(double) * 16 * (size_t)N);
  double *ur = wk, *uz = wk + N, *tr = wk + 2 * N, *tz = wk + 3 * N;
  double *k8 = wk + 4 * N; /* 8 kernel arrays of N */
  for (int i = 0; i < K; ++i) {
    t1[i] = T[i] * (1.0 + 0.5 * 0.01);
    t2[i] = T[i] * (1.0 - 0.5 * 0.01);
  }
  for (int w = 0; w < 2; ++w) {
    int M = (w == 0) ? nR : nL;
    int code = (w == 0) ? 0 : 2;
    for (int m = 0; m < M; ++m) {
      long long ebase = o->ent_off[col] + (long long)((w == 0 ? 0 : nR) + m) * K;
      int err = 0, errpm = 0;
      if (roots) {
        long long ne = o->ent_off[b->ncol];
        for (int k = 0; k < K; ++k) {
          cp[k] = roots[ebase + k];
          c1[k] = b->group ? roots[ne + ebase + k] : 0.0;
          c2[k] = b->group ? roots[2 * ne + ebase + k] : 0.0;
        }
      } else {
        err = be->surfdisp(thk, vp, vs, rho, N, T, cp, K, code, m, 1, 1);
        if (b->group) {
          errpm |= be->surfdisp(thk, vp, vs, rho, N, t1, c1, K, code, m, 1, 1);
          errpm |= be->surfdisp(thk, vp, vs, rho, N, t2, c2, K, code, m, 1, 1);
        }
      }
      for (int k = 0; k < K; ++k) {
        long long e = ebase + k;
        long long q = o->kern_off[col] + (e - o->ent_off[col]) * N;
        if (o->cflat) {
          long long ne = o->ent_off[b->ncol];
          o->cflat[e] = cp[k];
          o->cflat[ne + e] = b->group ? c1[k] : 0.0;
          o->cflat[2 * ne + e] = b->group ? c2[k] : 0.0;
        }
        if (err) { zero_entry(o, e, q, N, 4); continue; }
        if (cp[k] == 0.0) { zero_entry(o, e, q, N, 1); continue; }
        /* without the T*(1+-0.005) roots the T outputs stand (phase form)
         * and dU/dm is zero: status 2, or 8 on the exit path */
        int pm_bad = b->group && (errpm || c1[k] == 0.0 || c2[k] == 0.0);
        int st = pm_bad ? (errpm ? 8 : 2) : 0;
        double c = cp[k], u = 0.0;
        double *da = k8, *db = k8 + N, *dr = k8 + 2 * N, *dh = k8 + 3 * N;
        double *ua = k8 + 4 * N, *ub = k8 + 5 * N, *ur_ = k8 + 6 * N,
               *uh = k8 + 7 * N;
        memset(k8, 0, sizeof(double) * 8 * (size_t)N);
        if (w == 0 && b->group && !pm_bad)
          be->sregnpu(thk, vp, vs, rho, N, T + k, &c, &u, ur, uz, tr, tz,
                      t1 + k, c1 + k, t2 + k, c2 + k, da, db, dh, dr, ua, ub,
                      uh, ur_, 1);
        else if (w == 0)
          be->sregn96(thk, vp, vs, rho, N, T + k, &c, &u, ur, uz, tr, tz, da,
                      db, dh, dr, 1);
        else if (b->group && !pm_bad)
          be->slegnpu(thk, vs, rho, N, T + k, &c, &u, uz, tz, t1 + k, c1 + k,
                      t2 + k, c2 + k, db, dh, dr, ub, uh, ur_, 1);
        else
          be->slegn96(thk, vs, rho, N, T + k, &c, &u, uz, tz, db, dh, dr, 1);
        o->c[e] = c;
        o->u[e] = u;
        o->status[e] = st;
        double *src[8] = {da, db, dr, dh, ua, ub, ur_, uh};
        double *dst[8] = {o->dcda, o->dcdb, o->dcdr, o->dcdh,
                          o->duda, o->dudb, o->dudr, o->dudh};
        for (int a = 0; a < 8; ++a)
          if (dst[a]) memcpy(dst[a] + q, src[a], sizeof(double) * (size_t)N);
      }
    }
  }
  free(wk);
  free(cbuf);
}

void or_batch_solve(const or_backend *be, const or_batch *b,
                    const or_batch_out *o, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
#endif
  for (int c = 0; c < b->ncol; ++c) solve_column(be, b, o, c, NULL);
}

void or_batch_eigen(const or_backend *be, const or_batch *b, const or_batch_out *o,
                    const double *roots, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
#endif
  for (int c = 0; c < b->ncol; ++c) solve_column(be, b, o, c, roots);
}

/* SurfaceWave.cpp:601-641 (frechet_matrix): row r of the Fréchet matrix from
 * the dense ray sensitivity fdm_r (nx*ny floats, fdm(ix,iy) at ix + iy*nx,
 * column-major like fmat2) and the 1-D kernels (nx*ny, nz, kmax) column-major
 * (dmat3).  frechet(c) = fdm(ix+1,iy+1) * kernel((iy+1)*nx+ix+1, iz, idx) is
 * stored in a float vector; its nonzero entries (std::abs > 0) are kept in
 * increasing c.  Writes the CSR rows up to `cap` entries and returns the
 * total number of nonzero entries. */
long long or_frechet_rows(int nx, int ny, int nz, int kmax, const double *kernel, int nrows,
                          const int *row_idx, const float *fdm, long long *row_off, int *col,
                          float *val, long long cap) {
  long long nnz = 0;
  long long nxy = (long long)nx * ny;
  (void)kmax;
  row_off[0] = 0;
  for (int r = 0; r < nrows; ++r) {
    const float *f = fdm + (long long)r * nxy;
    int idx = row_idx[r];
    int c = 0;
    for (int iz = 0; iz < nz - 1; ++iz)
      for (int iy = 0; iy < ny - 2; ++iy)
        for (int ix = 0; ix < nx - 2; ++ix) {
          long long ii = (long long)(iy + 1) * nx + ix + 1;
          float tmp = (float)(f[(ix + 1) + (long long)(iy + 1) * nx] *
                              kernel[ii + (long long)iz * nxy + (long long)idx * nxy * nz]);
          if (fabsf(tmp) > 0.0f) {
            if (nnz < cap) {
              col[nnz] = c;
              val[nnz] = tmp;
            }
            ++nnz;
          }
          c += 1;
        }
    row_off[r + 1] = nnz;
  }
  return nnz;
}

/* SurfaceWave.cpp:738-773 (compute_grad) with one thread: the rows summed in
 * row order in float, each times its residual (t - t_obs), then every cell
 * times vs(ix+1, iy+1, iz) (vs column-major (nx, ny, nz)). */
void or_frechet_grad(int nx, int ny, int nz, int kmax, const double *kernel, int nrows,
                     const int *row_idx, const float *fdm, const float *resid, const float *vs,
                     float *grad) {
  long long nxy = (long long)nx * ny;
  long long n = (long long)(nx - 2) * (ny - 2) * (nz - 1);
  (void)kmax;
  for (long long i = 0; i < n; ++i) grad[i] = 0.0f;
  for (int r = 0; r < nrows; ++r) {
    const float *f = fdm + (long long)r * nxy;
    int idx = row_idx[r];
    int c = 0;
    for (int iz = 0; iz < nz - 1; ++iz)
      for (int iy = 0; iy < ny - 2; ++iy)
        for (int ix = 0; ix < nx - 2; ++ix) {
          long long ii = (long long)(iy + 1) * nx + ix + 1;
          float fr = (float)(f[(ix + 1) + (long long)(iy + 1) * nx] *
                             kernel[ii + (long long)iz * nxy + (long long)idx * nxy * nz]);
          grad[c] += fr * resid[r];
          c += 1;
        }
  }
  for (int k = 0; k < nz - 1; ++k)
    for (int j = 0; j < ny - 2; ++j)
      for (int i = 0; i < nx - 2; ++i) {
        long long ic = (long long)k * (ny - 2) * (nx - 2) + (long long)j * (nx - 2) + i;
        grad[ic] *= vs[(i + 1) + (long long)(j + 1) * nx + (long long)k * nxy];
      }
}
