/* CPU oracle (test infrastructure only): a plain-C restatement of JDSurfG's
 * FMST travel-time and ray-Frechet path, src/SWD/fmst.f90 and src/SWD/fmm.f90:
 *   fmst_init      fmst.f90:9-82     grid constants, allocations
 *   fmst_run       fmst.f90:108-332  (fmst_travel) source-grid refinement,
 *                                    two fast-marching passes, receiver times
 *   fmst_reset     fmst.f90:334-348  group-velocity grid for the ray integral
 *   fmst_raypath   fmst.f90:350-976  ray from receiver to source down the
 *                                    traveltime gradient, B-spline Frechet
 *                                    weights, group traveltime along the ray
 *   travel         fmm.f90:138-329   fast marching with a binary-heap band
 *   fouds2         fmm.f90:437-608   mixed-order upwind update
 *   addtree/downtree/updtree fmm.f90:616-772
 *   gridder        fmm.f90:776-870   cubic B-spline dicing of the control grid
 *   bsplrefine     fmm.f90:877-944   the same on the refined source grid
 *   srtimes        fmm.f90:952-1080  receiver traveltimes
 *   bilinear       fmm.f90:1662-1682
 * Arithmetic is single precision in the reference's order (built without FMA
 * contraction); the Fortran module state becomes one struct per handle, with
 * every node array allocated at the larger of the coarse and refined shapes
 * (the reference reallocates to that shape when the refined grid is larger,
 * fmst.f90:204-215).  Where the reference would run past its heap
 * (NINT(snb*nnx*nnz) entries) or leave a ray unfinished, the call reports a
 * status instead.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FPI 3.1415926535898f /* globalp: REAL(sp), PARAMETER :: pi */

typedef struct {
  int px, pz;
} fo_bp;

typedef struct {
  int nvx, nvz, nnx, nnz, fom, gdx, gdz;
  int vnl, vnr, vnt, vnb, nrnx, nrnz, sgdl, rbint;
  int nnxr, nnzr, asgr;
  float gox, goz, dnx, dnz, dvx, dvz, snb, earth;
  float goxd, gozd, dvxd, dvzd, dnxd, dnzd;
  float drnx, drnz, gorx, gorz;
  float dnxr, dnzr, goxr, gozr;
  int ld, nc;     /* allocated node-array shape (rows = z, columns = x) */
  float *velv;    /* (0:nvz+1, 0:nvx+1) */
  float *veln, *velnb, *ttn, *ttnr, *velnew;
  int *nsts, *nstsr;
  fo_bp *btg;     /* 1-based heap */
  int maxbt, ntr;
  int status;     /* 0 ok; 1 source or receiver outside; 2 heap overflow; 3 ray unfinished */
} fo_state;

#define VELV(s, i, j) (s)->velv[(size_t)(j) * ((s)->nvz + 2) + (i)]
#define AT(s, a, iz, ix) (s)->a[(size_t)((ix)-1) * (s)->ld + ((iz)-1)]

static float cube(float x) { return x * (x * x); }

/* the four cubic B-spline basis values at u in [0, 1] (fmm.f90:826-830) */
static void basis4(float u, float b[4]) {
  b[0] = cube(1.0f - u) / 6.0f;
  b[1] = (4.0f - 6.0f * (u * u) + 3.0f * cube(u)) / 6.0f;
  b[2] = (1.0f + 3.0f * u + 3.0f * (u * u) - 3.0f * cube(u)) / 6.0f;
  b[3] = cube(u) / 6.0f;
}

void *fo_init(int nx, int ny, float goxdf, float gozdf, float dvxdf, float dvzdf) {
  fo_state *s = (fo_state *)calloc(1, sizeof(fo_state));
  s->gdx = 8;
  s->gdz = 8;
  s->asgr = 1;
  s->sgdl = 8;
  s->earth = 6371.0f;
  s->fom = 1;
  s->snb = 0.5f;
  s->nvx = nx - 2;
  s->nvz = ny - 2;
  s->goxd = goxdf - dvxdf;
  s->gozd = gozdf + dvzdf;
  s->dvxd = dvxdf;
  s->dvzd = dvzdf;
  s->velv = (float *)calloc((size_t)(s->nvz + 2) * (s->nvx + 2), sizeof(float));
  s->dvx = s->dvxd * FPI / 180.0f;
  s->dvz = s->dvzd * FPI / 180.0f;
  s->gox = (90.0f - s->goxd) * FPI / 180.0f;
  s->goz = s->gozd * FPI / 180.0f;
  s->nnx = (s->nvx - 1) * s->gdx + 1;
  s->nnz = (s->nvz - 1) * s->gdz + 1;
  s->dnx = s->dvx / (float)s->gdx;
  s->dnz = s->dvz / (float)s->gdz;
  s->dnxd = s->dvxd / (float)s->gdx;
  s->dnzd = s->dvzd / (float)s->gdz;
  /* the refined source grid spans at most 2*sgs coarse cells (sgs = 8) */
  const int rmax = 16 * s->sgdl + 1;
  s->ld = s->nnz > rmax ? s->nnz : rmax;
  s->nc = s->nnx > rmax ? s->nnx : rmax;
  size_t nn = (size_t)s->ld * s->nc;
  s->veln = (float *)calloc(nn, sizeof(float));
  s->velnb = (float *)calloc(nn, sizeof(float));
  s->ttn = (float *)calloc(nn, sizeof(float));
  s->ttnr = (float *)calloc(nn, sizeof(float));
  s->velnew = (float *)calloc(nn, sizeof(float));
  s->nsts = (int *)calloc(nn, sizeof(int));
  s->nstsr = (int *)calloc(nn, sizeof(int));
  s->maxbt = (int)roundf(s->snb * (float)s->nnx * (float)s->nnz); /* NINT */
  s->btg = (fo_bp *)calloc((size_t)s->maxbt + 2, sizeof(fo_bp));
  s->rbint = 0;
  return s;
}

void fo_finalize(void *h) {
  fo_state *s = (fo_state *)h;
  if (!s) return;
  free(s->velv);
  free(s->veln);
  free(s->velnb);
  free(s->ttn);
  free(s->ttnr);
  free(s->velnew);
  free(s->nsts);
  free(s->nstsr);
  free(s->btg);
  free(s);
}

/* gridder (fmm.f90:776-870): control values from the dispersion map (float
 * of the double map, column n = ix + iy*nx), then the diced grid into out */
static void gridder(fo_state *s, const double *pv, float *out) {
  for (int i = 0; i <= s->nvz + 1; ++i)
    for (int j = 0; j <= s->nvx + 1; ++j) VELV(s, i, j) = (float)pv[i * (s->nvx + 2) + j];
  float ui[17][4], vi[17][4];
  for (int i = 1; i <= s->gdx + 1; ++i) basis4((float)(i - 1) / (float)s->gdx, ui[i - 1]);
  for (int i = 1; i <= s->gdz + 1; ++i) basis4((float)(i - 1) / (float)s->gdz, vi[i - 1]);
  for (int i = 1; i <= s->nvz - 1; ++i) {
    int conz = (i == s->nvz - 1) ? s->gdz + 1 : s->gdz;
    for (int j = 1; j <= s->nvx - 1; ++j) {
      int conx = (j == s->nvx - 1) ? s->gdx + 1 : s->gdx;
      for (int l = 1; l <= conz; ++l) {
        int stz = s->gdz * (i - 1) + l;
        for (int m = 1; m <= conx; ++m) {
          int stx = s->gdx * (j - 1) + m;
          float sumi = 0.0f;
          for (int i1 = 1; i1 <= 4; ++i1) {
            float sumj = 0.0f;
            for (int j1 = 1; j1 <= 4; ++j1)
              sumj = sumj + ui[m - 1][j1 - 1] * VELV(s, i - 2 + i1, j - 2 + j1);
            sumi = sumi + vi[l - 1][i1 - 1] * sumj;
          }
          out[(size_t)(stx - 1) * s->ld + (stz - 1)] = sumi;
        }
      }
    }
  }
}

/* bsplrefine (fmm.f90:877-944) into veln over the refined box */
static void bsplrefine(fo_state *s) {
  const int nrxr = s->gdx * s->sgdl, nrzr = s->gdz * s->sgdl;
  float ub[129][4], vb[129][4];
  for (int j = 1; j <= nrxr + 1; ++j) basis4((float)(j - 1) / (float)nrxr, ub[j - 1]);
  for (int i = 1; i <= nrzr + 1; ++i) basis4((float)(i - 1) / (float)nrzr, vb[i - 1]);
  const int origx = (s->vnl - 1) * s->sgdl + 1, origz = (s->vnt - 1) * s->sgdl + 1;
  for (int i = 1; i <= s->nvz - 1; ++i) {
    int conz = (i == s->nvz - 1) ? nrzr + 1 : nrzr;
    for (int j = 1; j <= s->nvx - 1; ++j) {
      int conx = (j == s->nvx - 1) ? nrxr + 1 : nrxr;
      for (int k = 1; k <= conz; ++k) {
        int st1 = s->gdz * (i - 1) + (k - 1) / s->sgdl + 1;
        if (st1 < s->vnt || st1 > s->vnb) continue;
        st1 = nrzr * (i - 1) + k;
        for (int l = 1; l <= conx; ++l) {
          int st2 = s->gdx * (j - 1) + (l - 1) / s->sgdl + 1;
          if (st2 < s->vnl || st2 > s->vnr) continue;
          st2 = nrxr * (j - 1) + l;
          float sum[4];
          for (int i1 = 1; i1 <= 4; ++i1) {
            sum[i1 - 1] = 0.0f;
            for (int j1 = 1; j1 <= 4; ++j1)
              sum[i1 - 1] = sum[i1 - 1] + ub[l - 1][j1 - 1] * VELV(s, i - 2 + i1, j - 2 + j1);
            sum[i1 - 1] = vb[k - 1][i1 - 1] * sum[i1 - 1];
          }
          int idm1 = st1 - origz + 1, idm2 = st2 - origx + 1;
          if (idm1 < 1 || idm1 > s->nnz) continue;
          if (idm2 < 1 || idm2 > s->nnx) continue;
          AT(s, veln, idm1, idm2) = sum[0] + sum[1] + sum[2] + sum[3];
        }
      }
    }
  }
}

static void bilinear(const fo_state *s, const float nv[2][2], float dsx, float dsz, float *biv) {
  float b = 0.0f;
  for (int i = 1; i <= 2; ++i)
    for (int j = 1; j <= 2; ++j) {
      float produ = (1.0f - fabsf(((float)(i - 1) * s->dnx - dsx) / s->dnx)) *
                    (1.0f - fabsf(((float)(j - 1) * s->dnz - dsz) / s->dnz));
      b = b + nv[i - 1][j - 1] * produ;
    }
  *biv = b;
}

/* heap (fmm.f90:616-772); nsts holds each close node's heap position */
static int addtree(fo_state *s, int iz, int ix) {
  if (s->ntr + 1 > s->maxbt) {
    s->status = 2;
    return 1;
  }
  s->ntr += 1;
  AT(s, nsts, iz, ix) = s->ntr;
  s->btg[s->ntr].px = ix;
  s->btg[s->ntr].pz = iz;
  int tpc = s->ntr, tpp = tpc / 2;
  while (tpp > 0) {
    if (AT(s, ttn, iz, ix) < AT(s, ttn, s->btg[tpp].pz, s->btg[tpp].px)) {
      AT(s, nsts, iz, ix) = tpp;
      AT(s, nsts, s->btg[tpp].pz, s->btg[tpp].px) = tpc;
      fo_bp e = s->btg[tpc];
      s->btg[tpc] = s->btg[tpp];
      s->btg[tpp] = e;
      tpc = tpp;
      tpp = tpc / 2;
    } else {
      tpp = 0;
    }
  }
  return 0;
}

static void downtree(fo_state *s) {
  if (s->ntr == 1) {
    s->ntr -= 1;
    return;
  }
  AT(s, nsts, s->btg[s->ntr].pz, s->btg[s->ntr].px) = 1;
  s->btg[1] = s->btg[s->ntr];
  s->ntr -= 1;
  int tpp = 1, tpc = 2 * tpp;
  while (tpc < s->ntr) {
    float rd1 = AT(s, ttn, s->btg[tpc].pz, s->btg[tpc].px);
    float rd2 = AT(s, ttn, s->btg[tpc + 1].pz, s->btg[tpc + 1].px);
    if (rd1 > rd2) tpc = tpc + 1;
    rd1 = AT(s, ttn, s->btg[tpc].pz, s->btg[tpc].px);
    rd2 = AT(s, ttn, s->btg[tpp].pz, s->btg[tpp].px);
    if (rd1 < rd2) {
      AT(s, nsts, s->btg[tpp].pz, s->btg[tpp].px) = tpc;
      AT(s, nsts, s->btg[tpc].pz, s->btg[tpc].px) = tpp;
      fo_bp e = s->btg[tpc];
      s->btg[tpc] = s->btg[tpp];
      s->btg[tpp] = e;
      tpp = tpc;
      tpc = 2 * tpp;
    } else {
      tpc = s->ntr + 1;
    }
  }
  if (tpc == s->ntr) {
    float rd1 = AT(s, ttn, s->btg[tpc].pz, s->btg[tpc].px);
    float rd2 = AT(s, ttn, s->btg[tpp].pz, s->btg[tpp].px);
    if (rd1 < rd2) {
      AT(s, nsts, s->btg[tpp].pz, s->btg[tpp].px) = tpc;
      AT(s, nsts, s->btg[tpc].pz, s->btg[tpc].px) = tpp;
      fo_bp e = s->btg[tpc];
      s->btg[tpc] = s->btg[tpp];
      s->btg[tpp] = e;
    }
  }
}

static void updtree(fo_state *s, int iz, int ix) {
  int tpc = AT(s, nsts, iz, ix), tpp = tpc / 2;
  while (tpp > 0) {
    if (AT(s, ttn, iz, ix) < AT(s, ttn, s->btg[tpp].pz, s->btg[tpp].px)) {
      AT(s, nsts, iz, ix) = tpp;
      AT(s, nsts, s->btg[tpp].pz, s->btg[tpp].px) = tpc;
      fo_bp e = s->btg[tpc];
      s->btg[tpc] = s->btg[tpp];
      s->btg[tpp] = e;
      tpc = tpp;
      tpp = tpc / 2;
    } else {
      tpp = 0;
    }
  }
}

/* fouds2 (fmm.f90:437-608): mixed-order upwind trial time at (iz, ix) */
static void fouds2(fo_state *s, int iz, int ix) {
  int tsw1 = 0;
  float travm = 0.0f;
  const float slown = 1.0f / AT(s, veln, iz, ix);
  const float ri = s->earth;
  const float risti = ri * sinf(s->gox + (float)(ix - 1) * s->dnx);
  for (int j = ix - 1; j <= ix + 1; j += 2) {
    if (!(j >= 1 && j <= s->nnx)) continue;
    int swj = -1, j2;
    if (j == ix - 1) {
      j2 = j - 1;
      if (j2 >= 1)
        if (AT(s, nsts, iz, j2) == 0) swj = 0;
    } else {
      j2 = j + 1;
      if (j2 <= s->nnx)
        if (AT(s, nsts, iz, j2) == 0) swj = 0;
    }
    if (AT(s, nsts, iz, j) == 0 && swj == 0) {
      swj = -1;
      if (AT(s, ttn, iz, j) > AT(s, ttn, iz, j2)) swj = 0;
    } else {
      swj = -1;
    }
    for (int k = iz - 1; k <= iz + 1; k += 2) {
      if (!(k >= 1 && k <= s->nnz)) continue;
      int swk = -1, k2;
      if (k == iz - 1) {
        k2 = k - 1;
        if (k2 >= 1)
          if (AT(s, nsts, k2, ix) == 0) swk = 0;
      } else {
        k2 = k + 1;
        if (k2 <= s->nnz)
          if (AT(s, nsts, k2, ix) == 0) swk = 0;
      }
      if (AT(s, nsts, k, ix) == 0 && swk == 0) {
        swk = -1;
        if (AT(s, ttn, k, ix) > AT(s, ttn, k2, ix)) swk = 0;
      } else {
        swk = -1;
      }
      int swsol = 0;
      float a = 0, b = 0, c = 0, u, v, em, tref = 0, tdiv = 1;
      if (swj == 0) {
        swsol = 1;
        if (swk == 0) {
          u = 2.0f * ri * s->dnx;
          v = 2.0f * risti * s->dnz;
          em = 4.0f * AT(s, ttn, iz, j) - AT(s, ttn, iz, j2) - 4.0f * AT(s, ttn, k, ix);
          em = em + AT(s, ttn, k2, ix);
          a = v * v + u * u;
          b = 2.0f * em * (u * u);
          c = (u * u) * (em * em - (slown * slown) * (v * v));
          tref = 4.0f * AT(s, ttn, iz, j) - AT(s, ttn, iz, j2);
          tdiv = 3.0f;
        } else if (AT(s, nsts, k, ix) == 0) {
          u = risti * s->dnz;
          v = 2.0f * ri * s->dnx;
          em = 3.0f * AT(s, ttn, k, ix) - 4.0f * AT(s, ttn, iz, j) + AT(s, ttn, iz, j2);
          a = v * v + 9.0f * (u * u);
          b = 6.0f * em * (u * u);
          c = (u * u) * (em * em - (slown * slown) * (v * v));
          tref = AT(s, ttn, k, ix);
          tdiv = 1.0f;
        } else {
          u = 2.0f * ri * s->dnx;
          a = 1.0f;
          b = 0.0f;
          c = -((u * u) * (slown * slown));
          tref = 4.0f * AT(s, ttn, iz, j) - AT(s, ttn, iz, j2);
          tdiv = 3.0f;
        }
      } else if (AT(s, nsts, iz, j) == 0) {
        swsol = 1;
        if (swk == 0) {
          u = ri * s->dnx;
          v = 2.0f * risti * s->dnz;
          em = 3.0f * AT(s, ttn, iz, j) - 4.0f * AT(s, ttn, k, ix) + AT(s, ttn, k2, ix);
          a = v * v + 9.0f * (u * u);
          b = 6.0f * em * (u * u);
          c = (u * u) * (em * em - (v * v) * (slown * slown));
          tref = AT(s, ttn, iz, j);
          tdiv = 1.0f;
        } else if (AT(s, nsts, k, ix) == 0) {
          u = ri * s->dnx;
          v = risti * s->dnz;
          em = AT(s, ttn, k, ix) - AT(s, ttn, iz, j);
          a = u * u + v * v;
          b = -(2.0f * (u * u) * em);
          c = (u * u) * (em * em - (v * v) * (slown * slown));
          tref = AT(s, ttn, iz, j);
          tdiv = 1.0f;
        } else {
          a = 1.0f;
          b = 0.0f;
          c = -((slown * slown) * (ri * ri) * (s->dnx * s->dnx));
          tref = AT(s, ttn, iz, j);
          tdiv = 1.0f;
        }
      } else {
        if (swk == 0) {
          swsol = 1;
          u = 2.0f * risti * s->dnz;
          a = 1.0f;
          b = 0.0f;
          c = -((u * u) * (slown * slown));
          tref = 4.0f * AT(s, ttn, k, ix) - AT(s, ttn, k2, ix);
          tdiv = 3.0f;
        } else if (AT(s, nsts, k, ix) == 0) {
          swsol = 1;
          a = 1.0f;
          b = 0.0f;
          c = -((slown * slown) * (risti * risti) * (s->dnz * s->dnz));
          tref = AT(s, ttn, k, ix);
          tdiv = 1.0f;
        }
      }
      if (swsol == 1) {
        float rd1 = b * b - 4.0f * a * c;
        if (rd1 < 0.0f) rd1 = 0.0f;
        float tdsh = (-b + sqrtf(rd1)) / (2.0f * a);
        float trav = (tref + tdsh) / tdiv;
        if (tsw1 == 1) {
          travm = fminf(trav, travm);
        } else {
          travm = trav;
          tsw1 = 1;
        }
      }
    }
  }
  AT(s, ttn, iz, ix) = travm;
}

/* travel (fmm.f90:138-329) */
static int travel(fo_state *s, float scx, float scz, int urg) {
  int isx = (int)((scx - s->gox) / s->dnx) + 1;
  int isz = (int)((scz - s->goz) / s->dnz) + 1;
  if (isx < 1 || isx > s->nnx || isz < 1 || isz > s->nnz) {
    s->status = 1;
    return 1;
  }
  if (isx == s->nnx) isx -= 1;
  if (isz == s->nnz) isz -= 1;
  const size_t nn = (size_t)s->ld * s->nc;
  if (urg != 2)
    for (size_t q = 0; q < nn; ++q) s->nsts[q] = -1;
  s->ntr = 0;
  if (urg == 2) {
    for (int i = 1; i <= s->nnx; ++i)
      for (int j = 1; j <= s->nnz; ++j)
        if (AT(s, nsts, j, i) > 0)
          if (addtree(s, j, i)) return 1;
  } else {
    float vss[2][2], vsrc;
    for (int i = 1; i <= 2; ++i)
      for (int j = 1; j <= 2; ++j) vss[i - 1][j - 1] = AT(s, veln, isz - 1 + j, isx - 1 + i);
    float dsx = (scx - s->gox) - (float)(isx - 1) * s->dnx;
    float dsz = (scz - s->goz) - (float)(isz - 1) * s->dnz;
    bilinear(s, (const float(*)[2])vss, dsx, dsz, &vsrc);
    for (int i = 1; i <= 2; ++i)
      for (int j = 1; j <= 2; ++j) {
        float ex = dsx - (float)(i - 1) * s->dnx, ez = dsz - (float)(j - 1) * s->dnz;
        float ds = sqrtf(ex * ex + ez * ez);
        AT(s, ttn, isz - 1 + j, isx - 1 + i) = 2.0f * ds / (vss[i - 1][j - 1] + vsrc);
        if (addtree(s, isz - 1 + j, isx - 1 + i)) return 1;
      }
  }
  while (s->ntr > 0) {
    if (urg == 1) {
      int ix = s->btg[1].px, iz = s->btg[1].pz, swrg = 0;
      /* the refined run compares its own node counts with the coarse box
       * bounds, as written (fmst.f90 swaps nnx/nnz before this call) */
      if (ix == 1 && s->vnl != 1) swrg = 1;
      if (ix == s->nnx && s->vnr != s->nnx) swrg = 1;
      if (iz == 1 && s->vnt != 1) swrg = 1;
      if (iz == s->nnz && s->vnb != s->nnz) swrg = 1;
      if (swrg == 1) {
        AT(s, nsts, iz, ix) = 0;
        break;
      }
    }
    int ix = s->btg[1].px, iz = s->btg[1].pz;
    AT(s, nsts, iz, ix) = 0;
    downtree(s);
    for (int i = ix - 1; i <= ix + 1; i += 2) {
      if (i >= 1 && i <= s->nnx) {
        if (AT(s, nsts, iz, i) == -1) {
          fouds2(s, iz, i);
          if (addtree(s, iz, i)) return 1;
        } else if (AT(s, nsts, iz, i) > 0) {
          fouds2(s, iz, i);
          updtree(s, iz, i);
        }
      }
    }
    for (int i = iz - 1; i <= iz + 1; i += 2) {
      if (i >= 1 && i <= s->nnz) {
        if (AT(s, nsts, i, ix) == -1) {
          fouds2(s, i, ix);
          if (addtree(s, i, ix)) return 1;
        } else if (AT(s, nsts, i, ix) > 0) {
          fouds2(s, i, ix);
          updtree(s, i, ix);
        }
      }
    }
  }
  return 0;
}

/* srtimes (fmm.f90:952-1080) for one receiver */
static float srtimes(fo_state *s, float scx, float scz, float rcx1, float rcz1) {
  int irx = (int)((rcx1 - s->gox) / s->dnx) + 1;
  int irz = (int)((rcz1 - s->goz) / s->dnz) + 1;
  int sw = 0;
  if (irx < 1 || irx > s->nnx || irz < 1 || irz > s->nnz) {
    s->status = 1;
    return 0.0f;
  }
  if (irx == s->nnx) irx -= 1;
  if (irz == s->nnz) irz -= 1;
  int isx = (int)((scx - s->gox) / s->dnx) + 1;
  int isz = (int)((scz - s->goz) / s->dnz) + 1;
  float dpl = s->dnx * s->earth;
  float rd1 = s->dnz * s->earth * sinf(s->gox);
  if (rd1 < dpl) dpl = rd1;
  rd1 = s->dnz * s->earth * sinf(s->gox + (float)(s->nnx - 1) * s->dnx);
  if (rd1 < dpl) dpl = rd1;
  float e1 = (scx - rcx1) * s->earth, e2 = (scz - rcz1) * s->earth * sinf(rcx1);
  float sred = sqrtf(e1 * e1 + e2 * e2);
  if (sred < dpl) sw = 1;
  if (isx == irx && isz == irz) sw = 1;
  float trr;
  if (sw == 1) {
    float vss[2][2], vels, velr;
    for (int k = 1; k <= 2; ++k)
      for (int l = 1; l <= 2; ++l) vss[k - 1][l - 1] = AT(s, veln, isz - 1 + l, isx - 1 + k);
    float drx = (scx - s->gox) - (float)(isx - 1) * s->dnx;
    float drz = (scz - s->goz) - (float)(isz - 1) * s->dnz;
    bilinear(s, (const float(*)[2])vss, drx, drz, &vels);
    for (int k = 1; k <= 2; ++k)
      for (int l = 1; l <= 2; ++l) vss[k - 1][l - 1] = AT(s, veln, irz - 1 + l, irx - 1 + k);
    drx = (rcx1 - s->gox) - (float)(irx - 1) * s->dnx;
    drz = (rcz1 - s->goz) - (float)(irz - 1) * s->dnz;
    bilinear(s, (const float(*)[2])vss, drx, drz, &velr);
    trr = 2.0f * sred / (vels + velr);
  } else {
    float drx = (rcx1 - s->gox) - (float)(irx - 1) * s->dnx;
    float drz = (rcz1 - s->goz) - (float)(irz - 1) * s->dnz;
    trr = 0.0f;
    for (int k = 1; k <= 2; ++k)
      for (int l = 1; l <= 2; ++l) {
        float produ = (1.0f - fabsf(((float)(l - 1) * s->dnz - drz) / s->dnz)) *
                      (1.0f - fabsf(((float)(k - 1) * s->dnx - drx) / s->dnx));
        trr = trr + AT(s, ttn, irz - 1 + l, irx - 1 + k) * produ;
      }
  }
  return trr;
}

/* fmst_travel (fmst.f90:108-332) */
int fo_run(void *h, const double *velf, float srcx, float srcz, const float *rcx,
           const float *rcz, int nr, float *ttime) {
  fo_state *s = (fo_state *)h;
  const int sgs = 8;
  s->status = 0;
  gridder(s, velf, s->veln);
  float x = srcx, z = srcz;
  const size_t nn = (size_t)s->ld * s->nc;
  /* asgr = 1: back up the coarse grid */
  memcpy(s->velnb, s->veln, nn * sizeof(float));
  const int nnxb = s->nnx, nnzb = s->nnz;
  const float dnxb = s->dnx, dnzb = s->dnz, goxb = s->gox, gozb = s->goz;
  int isx = (int)((x - s->gox) / s->dnx) + 1;
  int isz = (int)((z - s->goz) / s->dnz) + 1;
  if (isx < 1 || isx > s->nnx || isz < 1 || isz > s->nnz) {
    s->status = 1;
    return 1;
  }
  if (isx == s->nnx) isx -= 1;
  if (isz == s->nnz) isz -= 1;
  s->vnl = isx - sgs;
  if (s->vnl < 1) s->vnl = 1;
  s->vnr = isx + sgs;
  if (s->vnr > s->nnx) s->vnr = s->nnx;
  s->vnt = isz - sgs;
  if (s->vnt < 1) s->vnt = 1;
  s->vnb = isz + sgs;
  if (s->vnb > s->nnz) s->vnb = s->nnz;
  s->nrnx = (s->vnr - s->vnl) * s->sgdl + 1;
  s->nrnz = (s->vnb - s->vnt) * s->sgdl + 1;
  s->drnx = s->dvx / (float)(s->gdx * s->sgdl);
  s->drnz = s->dvz / (float)(s->gdz * s->sgdl);
  s->gorx = s->gox + s->dnx * (float)(s->vnl - 1);
  s->gorz = s->goz + s->dnz * (float)(s->vnt - 1);
  s->nnx = s->nrnx;
  s->nnz = s->nrnz;
  s->dnx = s->drnx;
  s->dnz = s->drnz;
  s->gox = s->gorx;
  s->goz = s->gorz;
  {
    /* reallocation rule of fmst.f90:204-215: the heap grows with the grid */
    int idm1 = s->nnx > nnxb ? s->nnx : nnxb, idm2 = s->nnz > nnzb ? s->nnz : nnzb;
    int maxbt = (int)roundf(s->snb * (float)idm1 * (float)idm2);
    if (maxbt > s->maxbt) {
      free(s->btg);
      s->maxbt = maxbt;
      s->btg = (fo_bp *)calloc((size_t)maxbt + 2, sizeof(fo_bp));
    }
  }
  bsplrefine(s);
  if (travel(s, x, z, 1)) return 1;
  memcpy(s->ttnr, s->ttn, nn * sizeof(float));
  memcpy(s->nstsr, s->nsts, nn * sizeof(int));
  const int ogx = s->vnl, ogz = s->vnt, grdfx = s->sgdl, grdfz = s->sgdl;
  for (size_t q = 0; q < nn; ++q) s->nsts[q] = -1;
  for (int k = 1; k <= s->nnz; k += grdfz) {
    int idm1 = ogz + (k - 1) / grdfz;
    for (int l = 1; l <= s->nnx; l += grdfx) {
      int idm2 = ogx + (l - 1) / grdfx;
      AT(s, nsts, idm1, idm2) = AT(s, nstsr, k, l);
      if (AT(s, nsts, idm1, idm2) >= 0) AT(s, ttn, idm1, idm2) = AT(s, ttnr, k, l);
    }
  }
  s->nnxr = s->nnx;
  s->nnzr = s->nnz;
  s->goxr = s->gox;
  s->gozr = s->goz;
  s->dnxr = s->dnx;
  s->dnzr = s->dnz;
  s->nnx = nnxb;
  s->nnz = nnzb;
  s->dnx = dnxb;
  s->dnz = dnzb;
  s->gox = goxb;
  s->goz = gozb;
  for (int j = 1; j <= s->nnx; ++j)
    for (int k = 1; k <= s->nnz; ++k) AT(s, veln, k, j) = AT(s, velnb, k, j);
  for (int k = 1; k <= s->nnx; ++k)
    for (int l = 1; l <= s->nnz; ++l)
      if (AT(s, nsts, l, k) == 0) {
        if (l - 1 >= 1 && AT(s, nsts, l - 1, k) == -1) AT(s, nsts, l, k) = 1;
        if (l + 1 <= s->nnz && AT(s, nsts, l + 1, k) == -1) AT(s, nsts, l, k) = 1;
        if (k - 1 >= 1 && AT(s, nsts, l, k - 1) == -1) AT(s, nsts, l, k) = 1;
        if (k + 1 <= s->nnx && AT(s, nsts, l, k + 1) == -1) AT(s, nsts, l, k) = 1;
      }
  if (travel(s, x, z, 2)) return 1;
  for (int i = 0; i < nr; ++i) ttime[i] = srtimes(s, x, z, rcx[i], rcz[i]);
  return s->status;
}

/* fmst_reset (fmst.f90:334-348): the group-velocity grid */
void fo_reset(void *h, const double *velin) {
  fo_state *s = (fo_state *)h;
  gridder(s, velin, s->velnew);
}

extern float delsph_f(float c1, float l1, float c2, float l2);

/* bilinear velocity of velnew at (x, z) as fmst_raypath computes it */
static float vel_at(const fo_state *s, float x, float z) {
  int ipx = (int)((x - s->gox) / s->dnx) + 1, ipz = (int)((z - s->goz) / s->dnz) + 1;
  float drx = (x - s->gox) - (float)(ipx - 1) * s->dnx;
  float drz = (z - s->goz) - (float)(ipz - 1) * s->dnz;
  float vel = 0.0f;
  for (int l = 1; l <= 2; ++l)
    for (int m = 1; m <= 2; ++m) {
      float produ = 1.0f - fabsf(((float)(m - 1) * s->dnz - drz) / s->dnz);
      produ = produ * (1.0f - fabsf(((float)(l - 1) * s->dnx - drx) / s->dnx));
      if (ipz - 1 + m <= s->nnz && ipx - 1 + l <= s->nnx)
        vel = vel + AT(s, velnew, ipz - 1 + m, ipx - 1 + l) * produ;
    }
  return vel;
}

/* fmst_raypath (fmst.f90:350-976).  fdm: (nvx+2) x (nvz+2), x fastest (the
 * transpose the reference returns); *status 3 when the ray does not reach
 * the source within nnx*nnz steps (the reference reads an unset count). */
float fo_raypath(void *h, float scx, float scz, float surfrcx, float surfrcz, float *fdm,
                 int *status) {
  fo_state *s = (fo_state *)h;
  const int nvx = s->nvx, nvz = s->nvz;
  const int maxrp = s->nnx * s->nnz;
  float *rgx = (float *)malloc(sizeof(float) * ((size_t)maxrp + 2));
  float *rgz = (float *)malloc(sizeof(float) * ((size_t)maxrp + 2));
  float *fdm0 = (float *)calloc((size_t)(nvz + 2) * (nvx + 2), sizeof(float));
#define FDM0(iz, ix) fdm0[(size_t)(ix) * (nvz + 2) + (iz)]
#define RGX(j) rgx[(j)-1]
#define RGZ(j) rgz[(j)-1]
  *status = 0;
  int isx = (int)((scx - s->goxr) / s->dnxr) + 1;
  int isz = (int)((scz - s->gozr) / s->dnzr) + 1;
  float dpl = s->dnx * s->earth;
  float rd1 = s->dnz * s->earth * sinf(s->gox);
  if (rd1 < dpl) dpl = rd1;
  rd1 = s->dnz * s->earth * sinf(s->gox + (float)(s->nnx - 1) * s->dnx);
  if (rd1 < dpl) dpl = rd1;
  dpl = 0.5f * dpl;
  int ipx = (int)((surfrcx - s->gox) / s->dnx) + 1;
  int ipz = (int)((surfrcz - s->goz) / s->dnz) + 1;
  int nrp = -1, sw = 0;
  float tsyn = 0.0f;
  if (ipx < 1 || ipx >= s->nnx || ipz < 1 || ipz >= s->nnz) {
    *status = 1;
    goto done;
  }
  if (ipx == s->nnx) ipx -= 1;
  if (ipz == s->nnz) ipz -= 1;
  RGX(1) = surfrcx;
  RGZ(1) = surfrcz;
  {
    float e1 = (scx - RGX(1)) * s->earth, e2 = (scz - RGZ(1)) * s->earth * sinf(RGX(1));
    float sred = sqrtf(e1 * e1 + e2 * e2);
    if (sred < 2.0f * dpl) {
      RGX(2) = scx;
      RGZ(2) = scz;
      nrp = 2;
      sw = 1;
    }
  }
  int ipxr = (int)((surfrcx - s->goxr) / s->dnxr) + 1;
  int ipzr = (int)((surfrcz - s->gozr) / s->dnzr) + 1;
  int igref = 1;
  if (ipxr < 1 || ipxr >= s->nnxr) igref = 0;
  if (ipzr < 1 || ipzr >= s->nnzr) igref = 0;
  if (igref == 1) {
    if (AT(s, nstsr, ipzr, ipxr) != 0 || AT(s, nstsr, ipzr + 1, ipxr) != 0) igref = 0;
    if (AT(s, nstsr, ipzr, ipxr + 1) != 0 || AT(s, nstsr, ipzr + 1, ipxr + 1) != 0) igref = 0;
  }
  if (sw == 0 && igref == 1 && ipxr == isx && ipzr == isz) {
    RGX(2) = scx;
    RGZ(2) = scz;
    nrp = 2;
    sw = 1;
  }
  for (int j = 1; j <= maxrp; ++j) {
    if (sw == 1) break;
    float dtx, dtz;
    if (igref == 1) {
      dtx = AT(s, ttnr, ipzr, ipxr + 1) - AT(s, ttnr, ipzr, ipxr);
      dtx = dtx + AT(s, ttnr, ipzr + 1, ipxr + 1) - AT(s, ttnr, ipzr + 1, ipxr);
      dtx = dtx / (2.0f * s->earth * s->dnxr);
      dtz = AT(s, ttnr, ipzr + 1, ipxr) - AT(s, ttnr, ipzr, ipxr);
      dtz = dtz + AT(s, ttnr, ipzr + 1, ipxr + 1) - AT(s, ttnr, ipzr, ipxr + 1);
      dtz = dtz / (2.0f * s->earth * sinf(RGX(j)) * s->dnzr);
    } else {
      dtx = AT(s, ttn, ipz, ipx + 1) - AT(s, ttn, ipz, ipx);
      dtx = dtx + AT(s, ttn, ipz + 1, ipx + 1) - AT(s, ttn, ipz + 1, ipx);
      dtx = dtx / (2.0f * s->earth * s->dnx);
      dtz = AT(s, ttn, ipz + 1, ipx) - AT(s, ttn, ipz, ipx);
      dtz = dtz + AT(s, ttn, ipz + 1, ipx + 1) - AT(s, ttn, ipz, ipx + 1);
      dtz = dtz / (2.0f * s->earth * sinf(RGX(j)) * s->dnz);
    }
    rd1 = sqrtf(dtx * dtx + dtz * dtz);
    RGX(j + 1) = RGX(j) - dpl * dtx / (s->earth * rd1);
    RGZ(j + 1) = RGZ(j) - dpl * dtz / (s->earth * sinf(RGX(j)) * rd1);
    int ipxo = ipx, ipzo = ipz;
    ipxr = (int)((RGX(j + 1) - s->goxr) / s->dnxr) + 1;
    ipzr = (int)((RGZ(j + 1) - s->gozr) / s->dnzr) + 1;
    igref = 1;
    if (ipxr < 1 || ipxr >= s->nnxr) igref = 0;
    if (ipzr < 1 || ipzr >= s->nnzr) igref = 0;
    if (igref == 1) {
      if (AT(s, nstsr, ipzr, ipxr) != 0 || AT(s, nstsr, ipzr + 1, ipxr) != 0) igref = 0;
      if (AT(s, nstsr, ipzr, ipxr + 1) != 0 || AT(s, nstsr, ipzr + 1, ipxr + 1) != 0) igref = 0;
    }
    ipx = (int)((RGX(j + 1) - s->gox) / s->dnx) + 1;
    ipz = (int)((RGZ(j + 1) - s->goz) / s->dnz) + 1;
    {
      float e1 = (scx - RGX(j + 1)) * s->earth, e2 = (scz - RGZ(j + 1)) * s->earth * sinf(RGX(j + 1));
      float sred = sqrtf(e1 * e1 + e2 * e2);
      sw = 0;
      if (sred < 2.0f * dpl) {
        RGX(j + 2) = scx;
        RGZ(j + 2) = scz;
        nrp = j + 2;
        sw = 1;
      }
    }
    if (sw == 0 && igref == 1 && ipxr == isx && ipzr == isz) {
      RGX(j + 2) = scx;
      RGZ(j + 2) = scz;
      nrp = j + 2;
      sw = 1;
    }
    if (ipx < 1) {
      RGX(j + 1) = s->gox;
      ipx = 1;
      s->rbint = 1;
    }
    if (ipx >= s->nnx) {
      RGX(j + 1) = s->gox + (float)(s->nnx - 1) * s->dnx;
      ipx = s->nnx - 1;
      s->rbint = 1;
    }
    if (ipz < 1) {
      RGZ(j + 1) = s->goz;
      ipz = 1;
      s->rbint = 1;
    }
    if (ipz >= s->nnz) {
      RGZ(j + 1) = s->goz + (float)(s->nnz - 1) * s->dnz;
      ipz = s->nnz - 1;
      s->rbint = 1;
    }
    /* Frechet weights of the segment j -> j+1 (fmst.f90:700-848) */
    int ivx = (ipx - 1) / s->gdx + 1, ivz = (ipz - 1) / s->gdz + 1;
    int ivxo = (ipxo - 1) / s->gdx + 1, ivzo = (ipzo - 1) / s->gdz + 1;
    int nhp = 0, chp[4];
    float vrat[4];
    if (ivx != ivxo) {
      nhp = nhp + 1;
      float xi = (ivx > ivxo) ? s->gox + (float)(ivx - 1) * s->dvx : s->gox + (float)ivx * s->dvx;
      vrat[nhp - 1] = (xi - RGX(j)) / (RGX(j + 1) - RGX(j));
      chp[nhp - 1] = 1;
    }
    if (ivz != ivzo) {
      nhp = nhp + 1;
      float zi = (ivz > ivzo) ? s->goz + (float)(ivz - 1) * s->dvz : s->goz + (float)ivz * s->dvz;
      float r = (zi - RGZ(j)) / (RGZ(j + 1) - RGZ(j));
      if (nhp == 1) {
        vrat[nhp - 1] = r;
        chp[nhp - 1] = 2;
      } else if (r >= vrat[nhp - 2]) {
        vrat[nhp - 1] = r;
        chp[nhp - 1] = 2;
      } else {
        vrat[nhp - 1] = vrat[nhp - 2];
        chp[nhp - 1] = chp[nhp - 2];
        vrat[nhp - 2] = r;
        chp[nhp - 2] = 2;
      }
    }
    nhp = nhp + 1;
    vrat[nhp - 1] = 1.0f;
    chp[nhp - 1] = 0;
    float vel = 0.0f;
    {
      float drx = (RGX(j) - s->gox) - (float)(ipxo - 1) * s->dnx;
      float drz = (RGZ(j) - s->goz) - (float)(ipzo - 1) * s->dnz;
      for (int l = 1; l <= 2; ++l)
        for (int m = 1; m <= 2; ++m) {
          float produ = 1.0f - fabsf(((float)(m - 1) * s->dnz - drz) / s->dnz);
          produ = produ * (1.0f - fabsf(((float)(l - 1) * s->dnx - drx) / s->dnx));
          if (ipzo - 1 + m <= s->nnz && ipxo - 1 + l <= s->nnx)
            vel = vel + AT(s, velnew, ipzo - 1 + m, ipxo - 1 + l) * produ;
        }
    }
    float vi[4], wi[4], vio[4], wio[4];
    {
      float drx = (RGX(j) - s->gox) - (float)(ivxo - 1) * s->dvx;
      float drz = (RGZ(j) - s->goz) - (float)(ivzo - 1) * s->dvz;
      basis4(drx / s->dvx, vi);
      basis4(drz / s->dvz, wi);
    }
    int ivxt = ivxo, ivzt = ivzo;
    for (int k = 1; k <= nhp; ++k) {
      float velo = vel;
      memcpy(vio, vi, sizeof vi);
      memcpy(wio, wi, sizeof wi);
      if (k > 1) {
        if (chp[k - 2] == 1)
          ivxt = ivx;
        else if (chp[k - 2] == 2)
          ivzt = ivz;
      }
      float rigz = RGZ(j) + vrat[k - 1] * (RGZ(j + 1) - RGZ(j));
      float rigx = RGX(j) + vrat[k - 1] * (RGX(j + 1) - RGX(j));
      int ipxt = (int)((rigx - s->gox) / s->dnx) + 1, ipzt = (int)((rigz - s->goz) / s->dnz) + 1;
      float drx = (rigx - s->gox) - (float)(ipxt - 1) * s->dnx;
      float drz = (rigz - s->goz) - (float)(ipzt - 1) * s->dnz;
      vel = 0.0f;
      for (int m = 1; m <= 2; ++m)
        for (int n = 1; n <= 2; ++n) {
          float produ = 1.0f - fabsf(((float)(n - 1) * s->dnz - drz) / s->dnz);
          produ = produ * (1.0f - fabsf(((float)(m - 1) * s->dnx - drx) / s->dnx));
          if (ipzt - 1 + n <= s->nnz && ipxt - 1 + m <= s->nnx)
            vel = vel + AT(s, velnew, ipzt - 1 + n, ipxt - 1 + m) * produ;
        }
      drx = (rigx - s->gox) - (float)(ivxt - 1) * s->dvx;
      drz = (rigz - s->goz) - (float)(ivzt - 1) * s->dvz;
      basis4(drx / s->dvx, vi);
      basis4(drz / s->dvz, wi);
      float dinc = (k == 1) ? vrat[k - 1] * dpl : (vrat[k - 1] - vrat[k - 2]) * dpl;
      for (int l = 1; l <= 4; ++l)
        for (int m = 1; m <= 4; ++m) {
          float r1 = vi[m - 1] * wi[l - 1] / (vel * vel);
          float r2 = vio[m - 1] * wio[l - 1] / (velo * velo);
          r1 = -((r1 + r2) * dinc / 2.0f);
          r2 = FDM0(ivzt - 2 + l, ivxt - 2 + m);
          FDM0(ivzt - 2 + l, ivxt - 2 + m) = r1 + r2;
        }
    }
  }
  if (sw != 1) {
    *status = 3;
    goto done;
  }
  for (int j = 1; j <= nrp - 1; ++j) {
    float dinc = delsph_f(RGX(j), RGZ(j), RGX(j + 1), RGZ(j + 1));
    float velo = vel_at(s, RGX(j), RGZ(j));
    float vel = vel_at(s, RGX(j + 1), RGZ(j + 1));
    tsyn = tsyn + dinc * 0.5f * (1.0f / vel + 1.0f / velo);
  }
done:
  for (int ix = 0; ix <= nvx + 1; ++ix)
    for (int iz = 0; iz <= nvz + 1; ++iz) fdm[(size_t)iz * (nvx + 2) + ix] = FDM0(iz, ix);
#undef FDM0
#undef RGX
#undef RGZ
  free(rgx);
  free(rgz);
  free(fdm0);
  return tsyn;
}

/* spherical.cpp:9-26 delsph: float variables, trigonometry in double */
float delsph_f(float c1, float l1, float c2, float l2) {
  const float R = 6371.0f;
  const float dlat = c2 - c1, dlon = l2 - l1;
  const float la1 = (float)(M_PI * 0.5 - (double)c1), la2 = (float)(M_PI * 0.5 - (double)c2);
  const float a = (float)(sin((double)dlat * 0.5) * sin((double)dlat * 0.5) +
                          sin((double)dlon * 0.5) * sin((double)dlon * 0.5) * cos((double)la1) *
                              cos((double)la2));
  const float c = (float)(2.0 * atan2(sqrt((double)a), sqrt((double)(1 - a))));
  return R * c;
}

/* One station pair as SurfTime::frechet_matrix drives it
 * (SurfaceWave.cpp:612-624): fmst_init, fmst_run on the phase map, fmst_reset
 * with the output map, fmst_raypath per receiver, fmst_finalize.  ttime gets
 * the phase traveltimes of srtimes, tsyn the ray-integrated times and fdm
 * nr x nx*ny weights (may be null).  Returns 0 or the first nonzero status. */
int fo_pair(int nx, int ny, float lat0, float lon0, float dlat, float dlon, const double *vc,
            const double *vout, float srcx, float srcz, int nr, const float *rcx,
            const float *rcz, float *ttime, float *tsyn, float *fdm) {
  void *h = fo_init(nx, ny, lat0, lon0, dlat, dlon);
  int st = fo_run(h, vc, srcx, srcz, rcx, rcz, nr, ttime);
  if (st == 0) {
    fo_reset(h, vout);
    float *tmp = fdm ? NULL : (float *)malloc(sizeof(float) * (size_t)nx * ny);
    for (int i = 0; i < nr && st == 0; ++i) {
      float *f = fdm ? fdm + (size_t)i * nx * ny : tmp;
      float t = fo_raypath(h, srcx, srcz, rcx[i], rcz[i], f, &st);
      if (tsyn) tsyn[i] = t;
    }
    free(tmp);
  }
  fo_finalize(h);
  return st;
}

/* Many pairs, OpenMP over pairs (each pair has its own state). */
int fo_pairs(int nx, int ny, float lat0, float lon0, float dlat, float dlon, int kmax,
             const double *vc, const double *vout, int npair, const int *idx,
             const float *srcx, const float *srcz, const int *roff, const float *rcx,
             const float *rcz, float *ttime, float *tsyn, float *fdm, int *status,
             int nthreads) {
  const size_t nxy = (size_t)nx * ny;
  (void)kmax;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int p = 0; p < npair; ++p) {
    const int r0 = roff[p], nr = roff[p + 1] - roff[p];
    status[p] = fo_pair(nx, ny, lat0, lon0, dlat, dlon, vc + (size_t)idx[p] * nxy,
                        vout + (size_t)idx[p] * nxy, srcx[p], srcz[p], nr, rcx + r0, rcz + r0,
                        ttime + r0, tsyn + r0, fdm ? fdm + (size_t)r0 * nxy : NULL);
  }
  return 0;
}
