// swd_kernels.hpp -- argument blocks and host launchers of the dispersion /
// kernel path (prep, root chains, eigen stage).
#pragma once
#include "swd_common.hpp"

struct PrepOut {
  float4 *chainR, *chainL;
  float *cc, *betmx;
  int *llw, *nlay;
  double *zd, *za, *zb, *rho0, *rtpR, *rtpL, *vtp, *dtp;
};

// A long upward dc-step scan of a suspended chain (exact-root fallback):
// the chain's next points are c0 + dc, (c0 + dc) + dc, ... (the same
// additions the chain makes), the scan stops at the first sign change
// against the previous point or after the first point >= bound.
struct ScanJob {
  long long g;               // chain (per wave type)
  double c0, f0;             // the last point the chain evaluated and its value
  double omega, bound, dc;
  int col, found;            // found: a sign change at c_hit
  unsigned evals;
  double c_prev, f_prev;     // the point before c_hit and its value (found)
  double c_hit, f_hit;       // the sign-change point, or the end point (not found)
};
struct ChainSusp {
  void *saved;               // [2 * nch] chain states, wave-major
  double *cs;                // [2 * nch * (2 kmax_pad + 22)] per-chain scratch
  ScanJob *jobs[2][2];       // [wave][list] -> [nch]
  unsigned *njobs;           // [wave * 2 + list]
  unsigned *total;           // [wave] scan jobs posted (statistics)
  int list_in;               // -1: start every chain; else resume the chains of this list
  int list_out;              // list that receives new suspensions
  int allow;                 // suspend long scans in this launch
  int min_steps;             // sequential scan steps before a scan may suspend
};

struct ChainArgs {
  ColumnModel cm;
  const long long *ent_off;   // [ncol+1]
  const int *per_off;         // [ncol+1]
  const double *periods;
  const int *nmode;           // [2*ncol] (R, L)
  int nvar;                   // 1 (phase only) or 3 (T, T*1.005, T*0.995)
  int igr;                    // surfdisp96 FD group velocity (B1 ABI)
  int retry;                  // surfdisp.cpp:95-109 retry of missing roots
  long long nent;             // entries per variant
  int kmax_pad;               // scratch stride for c/cb per chain
  double *cflat;              // [nvar * nent] float-rounded flat c (or FD U)
  int *cstat;                 // [nvar * nent] JDSG_EXIT_PATH bits
  int *ierr_out;              // [nchain] or null
  double *cscr;               // [nchain_slots * 2 * kmax_pad]
  unsigned *counter;          // work counters [2] (Rayleigh, Love launches)
  unsigned long long *nevals; // period-equation evaluations (flop model)
  double tol;                 // Neville stop rule (1e-6 reference; smaller = exact roots, unrounded)
  const double *known;        // optional [nvar*nent] roots already found (>0), else solved here
  int any_wave[2];            // some column asks for Rayleigh / Love modes
  long long nchain;           // ncol * 2 * nvar
  unsigned *chain_evals;  // optional [2*ncol*nvar] evaluations per chain
  const int *corder;      // optional [ncol]: columns in the order chains are taken (deepest first)
  int susp_on;            // exact-root fallback: long scans run as parallel scan jobs
  ChainSusp susp;
};

struct EigenArgs {
  ColumnModel cm;
  const long long *ent_off;  // [ncol+1]
  const int *per_off;
  const double *periods;
  const int *nmode;          // [2*ncol]
  int wave;                  // 0 Rayleigh, 1 Love (one launch per wave)
  int group;                 // 1: sregnpu/slegnpu form
  int iflsph;
  long long nent;            // entries per variant in cflat/cstat
  const double *cflat;       // [nvar * nent]
  const int *cstat;
  // item enumeration: items [0, nitem) map to entries of this wave
  long long nitem;
  const long long *item_ent;  // [nitem] entry index, or null: grid order (see decode_item)
  int grid_K, grid_nR, grid_nL;  // grid order: item = ((mode*K)+k)*ncol + col
  // outputs
  double *c_out, *u_out;
  int *status;
  double *dc[4], *du[4];     // a, b, rho, h; null allowed
  int layout;                // 0: grid (base = col + elocal*N*ncol, lstride ncol)
                             // 1: csr  (base = kern_off[col] + elocal*N, lstride 1)
  const long long *kern_off;
  double *eig[4];            // optional eigenfunctions (ur,uz,tr,tz) / (uu,tt), csr
  double *scr;               // lane scratch, [slot * nlanes + lane]
  int nmax;
  unsigned *counter;
  const double *tv;          // optional [3*nent] explicit (T, T1, T2) per entry
};

// SurfTime output entry: which job / mode / period feeds SurfTime index idx
struct StEntry {
  int job;   // index into jobs
  int mode;  // mode within the job's plan (0-based)
  int k;     // period index within the job's list
  int kind;  // 0: Rc/Lc (vout = c, dc kernels), 1: Rg/Lg (vout = U, dU kernels)
  int love;  // Love: no Vp kernel (SurfaceWave.cpp:367,403)
};
// Outputs of one layer plan over a chunk (grid layout: value of column nl,
// entry el, layer i at (el*N + i)*nchunk + nl).
struct StJob {
  const double *c, *u;
  const int *status;
  const double *dc[3], *du[3];  // b, a, rho
  const double *cflat;          // the plan's flat roots (T variant first)
  const int *cstat;             // the plan's chain status bits
  int K, nR, nL;
};
// get_2d_map output entry: the job whose chain gives vc, the job giving U
// for Rg/Lg (-1 for Rc/Lc), mode and period index within the jobs' lists
struct StMapEntry {
  int job_c, job_u, mode, k, love;
};


struct PolishArgs {
  ColumnModel cm;
  const long long *ent_off;
  const int *per_off;
  const double *periods;
  const int *nmode;
  long long nent;
  // polish launch: items [0, nitem*nv) of one wave type, variant v0 + item/nitem
  long long nitem;            // entries of this wave type
  const long long *item_ent;  // [nitem] entry index, or null: grid order
  int grid_K, grid_nR;        // grid order: item = ((mode*K)+k)*ncol + col
  int v0, nv;
  const double *centers[3];   // [nent] bracket centres (flat c) per variant
  double hw[3];               // relative half width per variant
  // mode-order pass: one variant
  int variant;                // 0: T, 1: T(1+.005), 2: T(1-.005)
  const double *center;
  double half_width;
  double tol;                 // polish stop rule
  double *known;              // [nvar*nent] out: root or 0
  unsigned *nfail;            // counter of roots left to the plain chain
  unsigned long long *nevals;
  signed char *sg;            // [3*nvar*nent] out: sign of F at bracket low, high, floor
  double *brk;                // [2*nvar*nent] out: the bracket each root was taken from
};

void jdsg_launch_prep(int ncol, int ncol_pad, const int *layer_off, const float *thk,
                      const float *vp, const float *vs, const float *rho, int iflsph,
                      const PrepOut &o, hipStream_t st);
void jdsg_launch_chains(const ChainArgs &A, int nwaves, hipStream_t st);
// exact-root fallback with long scans suspended into parallel scan jobs for
// up to `cycles` rounds (A.susp buffers set; A.susp_on is set here)
void jdsg_launch_chains_susp(ChainArgs A, int nwaves, int cycles, bool coop, hipStream_t st);
size_t jdsg_chain_state_bytes();
// ev (optional): 2 * jdsg_eigen_pass_launches(A, nwaves) events recorded
// before and after each pass launch (the pass kernels' own time)
void jdsg_launch_eigen(const EigenArgs &A, int nwaves, hipStream_t st, hipEvent_t *ev = nullptr);
int jdsg_eigen_pass_launches(const EigenArgs &A, int nwaves);
void jdsg_launch_fill_u32(unsigned *p, long long n, unsigned v, hipStream_t st);
void jdsg_launch_water_mask(int nxy, const float *vs, int *mask, hipStream_t st);
void jdsg_launch_grid_to_layers(int col0, int nchunk, int nxy, int nz, int N, int water,
                                const int *cols, const float *vs, const float *dep, float *thk,
                                float *vp, float *vsl, float *rho, float *vp2, float *node_vp,
                                float *node_vs, hipStream_t st);
void jdsg_launch_surftime_epilogue(int col0, int nchunk, int nxy, int nz, int N, int water,
                                   int kmax, const int *cols, const void *ents, const void *jobs,
                                   const float *node_vp, const float *node_vs, double *vc,
                                   double *vout, double *kernel, int *status, int want_kernel,
                                   hipStream_t st);
void jdsg_launch_surftime_map_epilogue(int col0, int nchunk, int nxy, int kmax, const int *cols,
                                       const void *ents, const void *jobs, const double *tlist,
                                       double *vc, double *vout, int *status, hipStream_t st);
void jdsg_launch_polish(const PolishArgs &A, int wave, int nwaves, hipStream_t st);
void jdsg_launch_mlp_expand(long long e0, long long nrow, int ncol, const long long *ent_off,
                            const int *per_off, const double *periods, const int *nmode,
                            const float *acol, int lda, const float *ext, int H, uint16_t *Y,
                            hipStream_t st);
long long jdsg_mlp_fused_work_len(long long nrow);  // ints of tile workspace
void jdsg_launch_mlp_fused(long long e0, long long nrow, int ncol, const long long *ent_off,
                           const int *per_off, const double *periods, const int *nmode,
                           const float *acol, int lda, const float *ext, const uint16_t *W2,
                           const uint16_t *W3, const uint16_t *W4, const float *b2,
                           const float *b3, const float *b4, const uint16_t *w5, float b5,
                           double scale, double shift, double *pred, long long p0, int *tcol,
                           hipStream_t st);

// Fréchet rows / misfit gradient (swd_frechet.hip)
struct RowCtx {
  int nx, ny, nz, nxy, nin;   // grid; interior cells per depth (nx-2)(ny-2)
  int kmax;                   // period indices of the kernel array
  long long kstride;          // nxy * nz: kernel stride between period indices
};
void jdsg_launch_frechet_rows(const RowCtx &R, const double *kernel, int nrows, const int *row_idx,
                              const long long *ray_off, const int *ray_cell, const float *ray_w,
                              long long *row_off, int *col, float *val, long long cap,
                              hipStream_t st);
void jdsg_launch_frechet_grad(const RowCtx &R, const double *kernel, int nrows,
                              const int *row_idx, const long long *ray_off, const int *ray_cell,
                              const float *ray_w, const float *resid, const float *vs,
                              float *grad, hipStream_t st);
void jdsg_launch_mode_order(const PolishArgs &A, int kmax, hipStream_t st);
void jdsg_launch_continue(const PolishArgs &A, int wave, int v, int mmax_modes, bool coop,
                          hipStream_t st);
void jdsg_launch_count_open(const double *known, long long n, unsigned *out, hipStream_t st);
void jdsg_launch_pm_centers(const long long *ent_off, const int *per_off, const double *periods,
                            int ncol, long long nent, const double *known0, const double *pred,
                            double *cen1, double *cen2, hipStream_t st);
void jdsg_launch_column_features(int ncol, const int *layer_off, const float *thk,
                                 const float *vp, const float *vs, const float *rho, float *colf,
                                 hipStream_t st);
void jdsg_launch_entry_features(long long e0, long long nrow, int ncol, const long long *ent_off,
                                const int *per_off, const double *periods, const int *nmode,
                                const float *colf, const long long *rows, uint16_t *X,
                                hipStream_t st);
void jdsg_launch_denorm(long long n, const float *out, int ldo, double scale, double shift,
                        double *pred, long long e0, hipStream_t st);
