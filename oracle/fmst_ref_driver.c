/* Test infrastructure only: the OpenMP loop over station pairs of
 * SurfTime::frechet_matrix (src/SWD/SurfaceWave.cpp:598-672) around the
 * reference FMST compiled from source (fmst.f90: fmst_init :9, fmst_run
 * :108, fmst_reset :334, fmst_raypath :350, fmst_finalize :84).  The
 * Fortran module state is threadprivate, so each OpenMP thread holds its
 * own, as in the reference's own loop.  Linked into _ref/libfmstref.so. */
#include <stddef.h>

void fmst_init(int nx, int ny, float goxdf, float gozdf, float dvxdf, float dvzdf);
void fmst_run(const double *vel, float srcx, float srcz, const float *rcx, const float *rcz,
              int nr, float *ttime);
void fmst_reset(const double *vel);
float fmst_raypath(float scx, float scz, float rcx, float rcz, float *fdm);
void fmst_finalize(void);

/* every pair: traveltimes (srtimes), ray times and weights (fdm, nrcv x
 * nx*ny, or per-thread scratch when fdm is null: scratch holds nthreads *
 * nx*ny floats) */
int fr_pairs(int nx, int ny, float lat0, float lon0, float dlat, float dlon, const double *vc,
             const double *vout, int npair, const int *idx, const float *srcx,
             const float *srcz, const int *roff, const float *rcx, const float *rcz,
             float *ttime, float *tsyn, float *fdm, float *scratch, int nthreads) {
  const size_t nxy = (size_t)nx * ny;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int p = 0; p < npair; ++p) {
    const int r0 = roff[p], nr = roff[p + 1] - roff[p];
    fmst_init(nx, ny, lat0, lon0, dlat, dlon);
    fmst_run(vc + (size_t)idx[p] * nxy, srcx[p], srcz[p], rcx + r0, rcz + r0, nr, ttime + r0);
    fmst_reset(vout + (size_t)idx[p] * nxy);
    for (int i = 0; i < nr; ++i) {
      float *f;
      if (fdm) {
        f = fdm + (size_t)(r0 + i) * nxy;
      } else {
        int t = 0;
#ifdef _OPENMP
        extern int omp_get_thread_num(void);
        t = omp_get_thread_num();
#endif
        f = scratch + (size_t)t * nxy;
      }
      tsyn[r0 + i] = fmst_raypath(srcx[p], srcz[p], rcx[r0 + i], rcz[r0 + i], f);
    }
    fmst_finalize();
  }
  return 0;
}
