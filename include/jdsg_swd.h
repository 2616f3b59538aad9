/*
 * jdsg_swd.h -- C ABI of the MI355X dispersion / depth-kernel library
 * (libjdsg_swd.so).  Plain pointers and sizes only.
 *
 * Three levels, matching the reference's boundaries:
 *   1. the per-column Fortran ABI of src/SWD/swd.hpp (same symbols, same
 *      argument meaning), executed on the GPU;
 *   2. the surfdisp/groupvel wrappers of src/SWD/surfdisp.cpp as C calls that
 *      return an error code where the reference calls exit(1);
 *   3. the batched forms: a layer-stack plan (the per-column call sequence of
 *      SurfaceWave.cpp:266-412 kernel1D for a whole batch) and the SurfTime
 *      grid entry points (SurfaceWave.cpp:420-502).
 * All functions return 0 on success; jdsg_last_error() describes a failure.
 */
#ifndef JDSG_SWD_H
#define JDSG_SWD_H
#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ level 1
 * Replaces sregn96_ (swd.hpp:19-23, sregn96.f90:1631).  cp is in/out: flat
 * phase velocity in, spherical out when iflsph=1. */
void sregn96_(const float *thk, const float *vp, const float *vs, const float *rhom,
              int nlayer, const double *t, double *cp, double *cg, double *dispu,
              double *dispw, double *stressu, double *stressw, double *dc2da,
              double *dc2db, double *dc2dh, double *dc2dr, int iflsph);
/* Replaces slegn96_ (swd.hpp:36-39, slegn96.f90:672). */
void slegn96_(const float *thk, const float *vs, const float *rhom, int nlayer,
              const double *t, double *cp, double *cg, double *disp, double *stress,
              double *dc2db, double *dc2dh, double *dc2dr, int iflsph);
/* Replaces surfdisp96_ (swd.hpp:52-54, surfdisp96.f:57). mode is 1-based,
 * iwave 1 Love / 2 Rayleigh, igr>0 the finite-difference group velocity. */
void surfdisp96_(const float *thkm, const float *vpm, const float *vsm,
                 const float *rhom, int nlayer, int iflsph, int iwave, int mode,
                 int igr, int kmax, const double *t, double *cg, int *ierr);
/* Replaces slegnpu_ (swd.hpp:69-74, slegn96.f90:785). */
void slegnpu_(float *thk, float *vs, float *rhom, int nlayer, const double *t,
              double *cp, double *cg, double *disp, double *stress, double *t1,
              double *cp1, double *t2, double *cp2, double *dc2db, double *dc2dh,
              double *dc2dr, double *du2db, double *du2dh, double *du2dr, int iflsph);
/* Replaces sregnpu_ (swd.hpp:89-95, sregn96.f90:1737). */
void sregnpu_(const float *thk, const float *vp, const float *vs, const float *rhom,
              int nlayer, const double *t, double *cp, double *cg, double *dispu,
              double *dispw, double *stressu, double *stressw, double *t1,
              double *cp1, double *t2, double *cp2, double *dc2da, double *dc2db,
              double *dc2dh, double *dc2dr, double *du2da, double *du2db,
              double *du2dh, double *du2dr, int iflsph);

/* ------------------------------------------------------------------ level 2
 * surfdisp (surfdisp.cpp:64-116, swd.hpp:99-102) with the wave type coded
 * 0=Rc 1=Rg 2=Lc 3=Lg; returns 1 where the reference exit(1)s. */
int jdsg_surfdisp(const float *thk, const float *vp, const float *vs, const float *rho,
                  int nlayer, const double *t, double *cg, int kmax, int wavetype,
                  int mode, int sphere, int keep_flat);

/* ------------------------------------------------------------------ level 3
 * Layer-stack batch plan.  Host arrays (copied at creation):
 *   layer_off[ncol+1]  layer offsets (column c has layer_off[c+1]-layer_off[c]
 *                      entries, the last the half-space);
 *   per_off[ncol+1]    period offsets (ascending periods per column);
 *   nmode[2*ncol]      Rayleigh and Love mode counts per column.
 * group=1 gives the Rg/Lg form of kernel1D (three root chains, U kernels);
 * group=0 the Rc/Lc form.  iflsph=1 is the spherical Earth JDSurfG uses.
 * layout 0 ("grid", uniform batches only): kernel value of entry-local index
 *   j (= (wave*nR + mode)*K + k) at layer i of column c is at
 *   [(j*N + i)*ncol + c], the column-fastest order of SurfTime's dmat3;
 * layout 1 ("csr"): at kern_off[c] + j*N + i.
 * Entry (c, j) scalar outputs are at ent_off[c] + j. */
typedef struct jdsg_plan jdsg_plan;
int jdsg_plan_create(int ncol, const int *layer_off, const int *per_off, const int *nmode,
                     int group, int layout, int iflsph, jdsg_plan **plan);
int jdsg_plan_sizes(const jdsg_plan *plan, long long *nent, long long *nkern, int *nmax,
                    int *kmax);
int jdsg_plan_offsets(const jdsg_plan *plan, long long *ent_off, long long *kern_off);
/* Launch the solve on `stream` (a hipStream_t, may be null).  Every pointer
 * is device memory; kernel outputs may be null to skip them.  status per
 * entry: 0 ok; 1 no root at T (outputs zero); 2 no root at T(1+-0.005)
 * (c, U, dc/dm stand, dU/dm zero); 4 the reference's exit(1) path in the T
 * chain (outputs zero); 8 the exit path in a T(1+-0.005) chain (as 2).
 * Allocates nothing and does not synchronise, so it may be graph-captured. */
int jdsg_plan_execute(jdsg_plan *plan, const float *thk, const float *vp, const float *vs,
                      const float *rho, const double *periods, double *c, double *u,
                      int *status, double *dcda, double *dcdb, double *dcdr, double *dcdh,
                      double *duda, double *dudb, double *dudr, double *dudh, void *stream);
/* Allocation-free form of jdsg_plan_create + jdsg_plan_execute + destroy (no
 * counterpart in the reference; for callers that cannot allocate device
 * memory: code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as the planned solve, the
 * descriptors (layer_off, per_off, nmode) as host arrays, everything else
 * device memory.  Every device array of the plan lives in one static
 * device array of jdsg_static_scratch_bytes() (64 MiB) per GPU, held until
 * the process ends; a batch whose plan does not fit is refused with an
 * error (nonzero return, jdsg_last_error()).  Calls allocate nothing; they
 * synchronise `stream` once (descriptor upload).  Calls that overlap on one
 * GPU share that array and race: serialise them. */
int jdsg_plan_execute_static(int ncol, const int *layer_off, const int *per_off,
                             const int *nmode, int group, int layout, int iflsph,
                             const float *thk, const float *vp, const float *vs,
                             const float *rho, const double *periods, double *c, double *u,
                             int *status, double *dcda, double *dcdb, double *dcdr,
                             double *dcdh, double *duda, double *dudb, double *dudr,
                             double *dudh, void *stream);
long long jdsg_static_scratch_bytes(void);
/* Per-kernel launch tuning: waves for the chain and eigen launches (0 keeps
 * the default). */
int jdsg_plan_set_waves(jdsg_plan *plan, int chain_waves, int eigen_waves);
void jdsg_plan_destroy(jdsg_plan *plan);
/* Stop rule of the Neville polish, |c1-c2| <= tol*c1.  The default 1e-6 is
 * the reference's (surfdisp96.f:634) and returns float-rounded roots like it;
 * tol < 1e-6 is the exact-root mode (roots kept in double). */
int jdsg_plan_set_tolerance(jdsg_plan *plan, double tol);
/* How the surrogate-bracketed solve runs the chains of the roots its
 * brackets left open (all forms give the same roots):
 * 0  one lane per chain, each chain to its end;
 * 1  one lane per chain, long upward scans suspended into parallel scan jobs;
 * 2  as 1, but a chain resumed after a suspended scan runs one chain per
 *    wave (the wave forms each period equation together) -- the default. */
int jdsg_plan_set_fallback(jdsg_plan *plan, int form);
/* Number of period-equation evaluations the last execute performed
 * (synchronises; for the flop model). */
int jdsg_plan_eval_count(const jdsg_plan *plan, long long *nevals);
/* Flat roots of the last execute into a host array of nvar*nent doubles
 * (nvar = 3 with group kernels: T, T(1+.005), T(1-.005); else 1). */
int jdsg_plan_flat_roots(const jdsg_plan *plan, double *host);
/* Diagnostics of the root chains: the first call enables per-chain
 * evaluation counting (from the next solve on); later calls copy the counts
 * of the last solve and each chain's ierr into host arrays of 2*ncol*nvar
 * (chain g = variant*ncol + column of wave w at w*ncol*nvar + g; either
 * pointer may be null). */
int jdsg_plan_chain_stats(jdsg_plan *plan, unsigned *evals, int *ierr);
/* Per-stage milliseconds (prep, chains, eigen Rayleigh, eigen Love) of the
 * last execute, when JDSG_TIMING=1 was set at plan creation. */
int jdsg_plan_timings(const jdsg_plan *plan, float *ms4);
/* The eigen pass kernels' own time (ms, summed over their launches) and
 * launch count per wave type (Rayleigh, Love) in the last execute, when
 * JDSG_TIMING=1 was set at plan creation (events around each pass launch,
 * on the stream of the execute). */
int jdsg_plan_pass_timings(const jdsg_plan *plan, float *ms2, int *n2);

/* Iterated batches (no counterpart in the reference, whose inversion loop
 * re-runs get_1d_kernel each iteration, surftomo.cpp:27-138).
 * jdsg_plan_autotune times the chain and eigen launches over candidate
 * widths on the given inputs, keeps the fastest (reported through
 * chain_waves / eigen_waves, may be null) and leaves a solve in the outputs.
 * jdsg_plan_capture records one iteration -- with perturb_amp > 0 first the
 * model update Vs *= 1 + perturb_amp*u (u uniform in [-1, 1), seeded by seed
 * and the replay count; Vp, rho by Brocher's relations; buffers updated in
 * place), then the whole solve -- on a non-null stream into a graph that
 * jdsg_graph_launch replays with one launch.  Plan and buffers must outlive
 * the graph. */
typedef struct jdsg_graph jdsg_graph;
int jdsg_plan_autotune(jdsg_plan *plan, const float *thk, const float *vp, const float *vs,
                       const float *rho, const double *periods, double *c, double *u,
                       int *status, double *dcda, double *dcdb, double *dcdr, double *dcdh,
                       double *duda, double *dudb, double *dudr, double *dudh, void *stream,
                       int *chain_waves, int *eigen_waves);
int jdsg_plan_capture(jdsg_plan *plan, float *thk, float *vp, float *vs, float *rho,
                      const double *periods, double *c, double *u, int *status, double *dcda,
                      double *dcdb, double *dcdr, double *dcdh, double *duda, double *dudb,
                      double *dudr, double *dudh, float perturb_amp, unsigned seed,
                      void *stream, jdsg_graph **graph);
int jdsg_graph_launch(jdsg_graph *graph, void *stream);
void jdsg_graph_destroy(jdsg_graph *graph);

/* SurfTime grid forms (replace SurfTime::get_1d_kernel and
 * SurfTime::get_2d_map, SurfaceWave.hpp:102-103, SurfaceWave.cpp:420-502;
 * depth grid and period tables as set_model / read_swd_data give them,
 * SurfaceWave.hpp:114-115).
 *   dep[nz]      node depths (km);
 *   t{Rc,Rg,Lc,Lg}  column-major (nt x nmode) period tables padded with -1
 *                (SurfaceWave.cpp:157-169); nmode 0 or a null table = absent.
 * A column whose top node has vs == 0 has a water top, decided per column
 * at every call as grid2LayerModel / _ParamConvert do (SurfaceWave.cpp:120,
 * 71, 495); a grid may mix water and land columns.
 * Output index order is SurfaceWave.cpp:14-48 get_period_index: Rc modes,
 * Rg modes, Lc modes, Lg modes, each mode's periods contiguous. */
typedef struct jdsg_surftime jdsg_surftime;
int jdsg_surftime_create(int nx, int ny, int nz, const float *dep, const double *tRc, int ntRc,
                         int nmRc, const double *tRg, int ntRg, int nmRg, const double *tLc,
                         int ntLc, int nmLc, const double *tLg, int ntLg, int nmLg,
                         jdsg_surftime **st);
int jdsg_surftime_kmax(const jdsg_surftime *st, int *kmax);
/* the grid and period-index count of a handle (SurfTime's nx, ny, depth
 * size and kmax); any pointer may be null */
int jdsg_surftime_dims(const jdsg_surftime *st, int *nx, int *ny, int *nz, int *kmax);
/* get_1d_kernel.  Device pointers: vs (nx,ny,nz) column-major float; vc,
 * vout, status (nx*ny, kmax) and kernel (nx*ny, nz, kmax) column-major
 * (kernel may be null).  Rc/Lc entries hold the spherical phase velocity in
 * vc and vout and dc/dVs in kernel; Rg/Lg entries hold c in vc, U in vout
 * and dU/dVs in kernel (SurfaceWave.cpp:296-405, chain rule :492-500).  A
 * water column's top node is not written (:495).  status as for plans;
 * entries with a nonzero status hold zeros where the reference would
 * compute from c = 0.  Synchronises `stream` once at the start (it reads
 * the water-top mask of vs), then runs asynchronously. */
int jdsg_surftime_kernel(jdsg_surftime *st, const float *vs, double *vc, double *vout,
                         double *kernel, int *status, void *stream);
/* get_2d_map (disp1D, SurfaceWave.cpp:177-256): vc is the spherical phase
 * velocity from one root chain per (wave, mode) and _flat2sphere
 * (surfdisp.cpp:17-51); vout is vc for Rc/Lc and the analytic group
 * velocity of groupvel_r / groupvel_l for Rg/Lg (surfdisp.cpp:126-174; the
 * Love one on the chain groupvel_l runs with Vp = 1.732 Vs).  No kernels.
 * Device pointers and synchronisation as jdsg_surftime_kernel. */
int jdsg_surftime_map(jdsg_surftime *st, const float *vs, double *vc, double *vout, int *status,
                      void *stream);
void jdsg_surftime_destroy(jdsg_surftime *st);
/* Host-array convenience forms of the above (allocate, copy, solve). */
int jdsg_get_1d_kernel(int nx, int ny, int nz, const float *vs, const float *dep,
                       const double *tRc, int ntRc, int nmRc, const double *tRg, int ntRg,
                       int nmRg, const double *tLc, int ntLc, int nmLc, const double *tLg,
                       int ntLg, int nmLg, double *vc, double *vout, double *kernel,
                       int *status);
int jdsg_get_2d_map(int nx, int ny, int nz, const float *vs, const float *dep,
                    const double *tRc, int ntRc, int nmRc, const double *tRg, int ntRg,
                    int nmRg, const double *tLc, int ntLc, int nmLc, const double *tLg,
                    int ntLg, int nmLg, double *vc, double *vout, int *status);

/* Fréchet rows and misfit gradient of the travel-time data (replace the row
 * loops of SurfTime::frechet_matrix, SurfaceWave.cpp:584-672, and of
 * SurfTime::compute_grad, :695-781, after the eikonal ray tracing).  Device
 * pointers except where noted.
 *   kernel     (nx*ny, nz, kmax) column-major, as jdsg_surftime_kernel fills it;
 *   row_idx    [nrows] period index of each data row (get_period_index order);
 *   ray_off    [nrows+1], ray_cell, ray_w: the row's ray sensitivity fdm
 *              (fmst_raypath) in sparse form: the grid cells n = ix + iy*nx
 *              where it is nonzero, increasing, and their float weights.
 * Row r holds, in increasing column c = iz*(nx-2)*(ny-2) + (iy-1)*(nx-2) +
 * (ix-1) over the interior cells (0 <= iz < nz-1), the nonzero values of
 * float(fdm(ix,iy) * kernel(n, iz, idx)) (:613-641).
 * jdsg_frechet_rows writes the CSR offsets row_off [nrows+1], col and val
 * (capacity cap) and the total in *nnz (host); it synchronises `stream` and
 * fails with the total set when it exceeds cap, or when a row's period index
 * or ray cells are invalid.
 * jdsg_frechet_grad: grad[(nx-2)*(ny-2)*(nz-1)] = sum over rows of the row
 * times resid[r] (t - t_obs), then times vs at the cell (vs (nx,ny,nz)
 * column-major) (:754-773); asynchronous.
 * jdsg_write_frechet (host arrays): the reference's frechet.bin, header
 * (m, n) then per row (row id, nnz, int[nnz] columns, float[nnz] values), as
 * merge_csr_files (csr_matrix.cpp:205-242) leaves it. */
int jdsg_frechet_rows(int nx, int ny, int nz, int kmax, const double *kernel, int nrows,
                      const int *row_idx, const long long *ray_off, const int *ray_cell,
                      const float *ray_w, long long *row_off, int *col, float *val,
                      long long cap, long long *nnz, void *stream);
int jdsg_frechet_grad(int nx, int ny, int nz, int kmax, const double *kernel, int nrows,
                      const int *row_idx, const long long *ray_off, const int *ray_cell,
                      const float *ray_w, const float *resid, const float *vs, float *grad,
                      void *stream);
int jdsg_write_frechet(const char *path, int m, int n, int nrows, const int *row_id,
                       const long long *row_off, const int *col, const float *val);

/* Learned-bracket root finding (no counterpart in the reference: an optional
 * accelerated form of the root search of surfdisp96.f:213-339 that returns
 * the same roots in exact-root mode).  An MLP predicts every flat phase
 * velocity; each prediction is bracketed, sign-checked and polished in FP64,
 * and the roots that fail (no sign change, mode order, an odd number of
 * roots skipped below the bracket) are solved by the plain chain.
 *
 * Flat roots of the plain solver (nvar*nent doubles, 0 where none; status
 * as for plans, may be null) -- the training targets. */
int jdsg_plan_roots(jdsg_plan *plan, const float *thk, const float *vp, const float *vs,
                    const float *rho, const double *periods, double *c_out, int *stat_out,
                    void *stream);
/* MLP input rows, bf16 [nrow][192], for entries e0..e0+nrow-1 or, with rows
 * non-null, for the entry indices rows[0..nrow); colf is a device scratch
 * of ncol*148 floats rebuilt when column_pass is 1. */
int jdsg_plan_features(jdsg_plan *plan, const float *thk, const float *vp, const float *vs,
                       const float *rho, const double *periods, long long e0, long long nrow,
                       const long long *rows, void *X, float *colf, int column_pass,
                       void *stream);
/* First hidden layer split by what varies: only ln T and the wave and mode
 * one-hots of an entry's 192 inputs differ between entries of one column.
 * jdsg_mlp_column_rows writes per-column input rows (bf16 [ncol][192]: the
 * column features, entry inputs zero) from colf; one GEMM of them with the
 * first weights gives acol (fp32 [ncol][lda]).  jdsg_plan_mlp_expand then
 * writes Y = gelu(acol + bf16(lnT) ext[0] + ext[1+wave] + ext[3+mode] +
 * ext[7]) (bf16 [nrow][H]) for entries e0..e0+nrow-1, ext (fp32 [8][H])
 * holding the weight columns 144..150 and the bias; H a multiple of 8 up to
 * 1024, lda >= H a multiple of 4, acol, ext and Y 16-byte aligned. */
int jdsg_mlp_column_rows(int ncol, const float *colf, void *X, void *stream);
int jdsg_plan_mlp_expand(jdsg_plan *plan, const double *periods, long long e0, long long nrow,
                         const float *acol, int lda, const float *ext, int H, void *Y,
                         void *stream);
/* The whole network of hidden width 256 at inference in one kernel: first
 * layer as jdsg_plan_mlp_expand, hidden layers 2-4 (bf16 weights W2..W4,
 * 256x256 row-major out x in, fp32 biases b2..b4) with GELU, and the head as
 * jdsg_mlp_head (w5 the bf16 output row, b5 its bias): pred[p0 + r] for
 * entries e0 + r, r < nrow, the same bits as those kernels in sequence.
 * Operands 16-byte aligned, lda >= 256 and a multiple of 4; work is a
 * device workspace of work_len ints, at least jdsg_mlp_fused_work(nrow). */
long long jdsg_mlp_fused_work(long long nrow);
int jdsg_plan_mlp_fused(jdsg_plan *plan, const double *periods, long long e0, long long nrow,
                        const float *acol, int lda, const float *ext, const void *W2,
                        const void *W3, const void *W4, const float *b2, const float *b3,
                        const float *b4, const void *w5, float b5, double scale, double shift,
                        double *pred, long long p0, int *work, long long work_len,
                        void *stream);
/* pred[e0 + r] = exp(out[r*ldo]*scale + shift), r < n (the network predicts
 * a standardised ln c). */
int jdsg_denorm(long long n, const float *out, int ldo, double scale, double shift,
                double *pred, long long e0, void *stream);
/* Allocates the surrogate buffers of a plan (once). */
int jdsg_plan_enable_surrogate(jdsg_plan *plan);
/* The solve with bracket centres pred[nent] (device); the plan must be in
 * exact-root mode (tolerance < 1e-6).  Outputs as jdsg_plan_execute. */
int jdsg_plan_execute_surrogate(jdsg_plan *plan, const float *thk, const float *vp,
                                const float *vs, const float *rho, const double *periods,
                                const double *pred, double half_width, double *c, double *u,
                                int *status, double *dcda, double *dcdb, double *dcdr,
                                double *dcdh, double *duda, double *dudb, double *dudr,
                                double *dudh, void *stream);
/* The whole surrogate iteration as device work only.  jdsg_plan_set_network
 * binds a network to the plan (W[0..4] bf16 row-major out x in: hid x 192,
 * three hid x hid, 64 x hid; b[0..4] fp32; the caller keeps them; the
 * prediction is exp(out * scale + shift)) and allocates the inference
 * buffers; jdsg_plan_execute_surrogate_net then predicts every root
 * (column features, first-layer GEMM per column, hidden layers on MFMA, the
 * head) and runs jdsg_plan_execute_surrogate on the predictions, with no
 * host synchronisation; jdsg_plan_capture_surrogate records the optional
 * model update (as jdsg_plan_capture) and that iteration into a graph that
 * jdsg_graph_launch replays.  jdsg_plan_network_pred copies the last
 * predictions (nent doubles) to the host. */
int jdsg_plan_set_network(jdsg_plan *plan, int hid, const void *const W[5],
                          const float *const b[5], double scale, double shift);
int jdsg_plan_execute_surrogate_net(jdsg_plan *plan, const float *thk, const float *vp,
                                    const float *vs, const float *rho, const double *periods,
                                    double half_width, double *c, double *u, int *status,
                                    double *dcda, double *dcdb, double *dcdr, double *dcdh,
                                    double *duda, double *dudb, double *dudr, double *dudh,
                                    void *stream);
int jdsg_plan_capture_surrogate(jdsg_plan *plan, float *thk, float *vp, float *vs, float *rho,
                                const double *periods, double half_width, double *c, double *u,
                                int *status, double *dcda, double *dcdb, double *dcdr,
                                double *dcdh, double *duda, double *dudb, double *dudr,
                                double *dudh, float perturb_amp, unsigned seed, void *stream,
                                jdsg_graph **graph);
int jdsg_plan_network_pred(const jdsg_plan *plan, double *host);
/* The polished roots the identity checks kept after the last surrogate
 * execute (nvar*nent doubles, 0 where the chain took over) and the signs of
 * the period equation at bracket low, high and the chain floor (3*nvar*nent
 * bytes); either pointer may be null. */
int jdsg_plan_surrogate_roots(const jdsg_plan *plan, double *known, signed char *signs);
/* Roots the last surrogate execute left to the chain, and the period-
 * equation evaluations of its bracket/polish and chain stages. */
int jdsg_plan_surrogate_stats(const jdsg_plan *plan, long long *nfail, long long *polish_evals,
                              long long *chain_evals);
/* Long chain scans the last surrogate execute ran as parallel scan jobs: in
 * its fallback a chain whose upward dc-step scan (surfdisp96.f getsol) has
 * gone two points without a sign change is suspended, its remaining scan
 * points (the same c + dc additions) are evaluated 64 at a time by one
 * wavefront, and the chain resumes at the first sign change or at the end
 * of the scan, as the sequential scan would. */
int jdsg_plan_scan_jobs(const jdsg_plan *plan, long long *jobs);

/* BF16 MFMA kernels of the surrogate network (training and inference).
 * C = op(A) op(B), A(i,k) = ta ? A[k*lda+i] : A[i*lda+k], B(k,j) = tb ?
 * B[j*ldb+k] : B[k*ldb+j]; M, N multiples of 64, K of 32; bf16 operands,
 * fp32 accumulation.  epi 0: C; 1: C = Z = acc+bias, Y = bf16(gelu(Z));
 * 2: C = acc+bias; 3: C += acc; 4: Y = bf16(gelu(acc+bias)) only. */
int jdsg_gemm_bf16(int ta, int tb, int M, int N, int K, const void *A, int lda, const void *B,
                   int ldb, float *C, int ldc, int epi, const float *bias, void *Y, int ldy,
                   void *stream);
/* dZ = dY * gelu'(Z): bf16 dZ and (if non-null) fp32 dZf. */
int jdsg_gelu_bwd(long long n, const float *dY, const float *Z, void *dZ, float *dZf,
                  void *stream);
/* out[j] = sum_i X[i*N + j] (bias gradients). */
int jdsg_colsum(int M, int N, const float *X, float *out, void *stream);
/* Weighted loss on column 0 of out (ld ldo): squared error, or Huber with
 * delta > 0; dY = dloss/dout / M, per-256-row partial losses. */
int jdsg_mse(int M, int ldo, const float *out, const float *target, const float *wt, float delta,
             float *dY, float *loss_part, void *stream);
/* Adam (bias-corrected; wd adds wd*w to the gradient) on fp32 master
 * weights; wbf, if non-null, receives the bf16 copy. */
int jdsg_adam(long long n, float *w, float *m, float *v, const float *g, float lr, float b1,
              float b2, float eps, int step, float wd, void *wbf, void *stream);
int jdsg_f32_to_bf16(long long n, const float *x, void *y, void *stream);
/* Inference output head: pred[e0 + i] = exp((Y_i . w + b)*scale + shift) for
 * the n rows of the last hidden layer Y (bf16, row stride H, H a multiple of
 * 64, 16-byte aligned) and the first row w of the output weights (bf16);
 * fp32 accumulation in a fixed order.  Nonzero on bad sizes or alignment. */
int jdsg_mlp_head(long long n, int H, const void *Y, const void *w, float b, double scale,
                  double shift, double *pred, long long e0, void *stream);

const char *jdsg_last_error(void);

/* ---- least-squares solve of the inversion step ------------------------ */
/* LSMR on a CSR matrix (m rows, n columns; val, indices, indptr, b and x in
 * device memory, 0-based, nnz = indptr[m] < 2^31), replacing
 * LSMR_csr (src/shared/lsmrModule_csr.f90:516, bind(C)) as called by
 * csr_matrix::lsmr_solver (src/shared/csr_matrix.cpp:114-128): the same
 * arguments in the same order and meaning, float vectors and scalars, x set
 * to the solution, istop/itn and the norm estimates returned; verbose and
 * num_threads give way to a HIP stream.  Nonzero on a bad argument or a HIP
 * failure (message in jdsg_lsmr_last_error). */
int jdsg_lsmr_csr(int m, int n, const float *val, const int *indices, const int *indptr,
                  const float *b, float damp, float atol, float btol, float conlim, int itnlim,
                  int localSize, float *x, int *istop, int *itn, float *normA, float *condA,
                  float *normr, float *normAr, float *normx, void *stream);
/* The same solve for a matrix whose columns are numbered layer-major,
 * c = layer * (n / layers) + cell, as SurfTime::frechet_matrix numbers them
 * (src/SWD/SurfaceWave.cpp:630-637: depth outermost, then y, then x): the
 * iterations run in the cell-major numbering (cell * layers + layer), where
 * the layers of one cell share a cache line of x, and x comes back in the
 * caller's numbering.  No counterpart in the reference (LSMR's iterates do
 * not depend on the column order; only the rounding order changes).  layers
 * must divide n. */
int jdsg_lsmr_csr_layers(int m, int n, const float *val, const int *indices, const int *indptr,
                         const float *b, float damp, float atol, float btol, float conlim,
                         int itnlim, int localSize, float *x, int *istop, int *itn, float *normA,
                         float *condA, float *normr, float *normAr, float *normx, int layers,
                         void *stream);
const char *jdsg_lsmr_last_error(void);

/* ---- model smoothing and regularisation rows (device arrays, column-major,
 * x fastest; 0 on success, else nonzero with jdsg_smooth_last_error) ---- */

/* smooth_cart_pde (src/shared/smoothing.cpp:177-249): Gaussian smoothing by
 * explicit diffusion steps on a grid of unit spacing, in place; the
 * finite-difference coefficients are the exact central-difference rationals
 * (the reference solves for them with a float LU). */
int jdsg_smooth_cart_pde(float *grad, int nx, int ny, int nz, float sigma_h, float sigma_v,
                         void *stream);
/* smooth_sph_pde (smoothing.cpp:256-362): the same in spherical coordinates,
 * nodes (colatitude from lat0, longitude from lon0, depth from z0) spaced
 * dx, dy degrees and dz km, in place. */
int jdsg_smooth_sph_pde(float *grad, int nx, int ny, int nz, float dx, float dy, float dz,
                        float lat0, float lon0, float z0, float sigma_h, float sigma_v,
                        void *stream);
/* interp_irregular_z (smoothing.cpp:122-170): fwd != 0 fills gradr (nx, ny,
 * nz1) from grad (nx, ny, nz) on the increasing host depths dep[nz]; else
 * fills grad from gradr. */
int jdsg_interp_irregular_z(float *grad, float *gradr, int nx, int ny, int nz, int nz1,
                            const float *dep, int fwd, void *stream);
/* DSurfTomo::smoothing (src/surftomo/surftomo.cpp:147-178): grad of shape
 * (nx-2, ny-2, nz-1) for a model of nx x ny x nz nodes with host axes lat[nx],
 * lon[ny], dep[nz]: resampled to a regular depth grid, smoothed (spherical
 * when smooth_in_km, else Cartesian), resampled back, in place. */
int jdsg_smooth_gradient(float *grad, int nx, int ny, int nz, const float *dep, const float *lat,
                         const float *lon, int smooth_in_km, float sigma_h, float sigma_v,
                         void *stream);
/* add_regularization (src/shared/csr_matrix.cpp:268-328): appends the
 * nx*ny*nz Laplacian rows (2w on a face node; 6w and six -w inside) after the
 * m rows of a CSR matrix in device memory, entries from ptr[m] on; ptr holds
 * m + nx*ny*nz + 1 ints and col / val cap entries.  *nnz = the new entry
 * count; fails, as the reference exits, when cap is too small. */
int jdsg_add_regularization(int m, int *ptr, int *col, float *val, long long cap, float weight,
                            int nx, int ny, int nz, long long *nnz, void *stream);
const char *jdsg_smooth_last_error(void);

/* ---- surface-wave data file (host code) ---- */

/* SurfTime::read_swd_data (src/SWD/io.cpp:29-238, read_receiver :6-27): the
 * period tables and the station pairs of surfdataSC.dat.  Wave type codes:
 * 0 Rc, 1 Rg, 2 Lc, 3 Lg; periods 0-based; positions as colatitude and
 * longitude in radians; distances (km, haversine) and observed times
 * (distance / velocity) as float.  0 on success, else nonzero with
 * jdsg_io_last_error (the reference prints and exits). */
typedef struct jdsg_swd_data jdsg_swd_data;
int jdsg_read_swd_data(const char *path, jdsg_swd_data **out);
void jdsg_swd_data_sizes(const jdsg_swd_data *d, int *nmode, int *nt, long long *npairs,
                         long long *nobs);
/* nmode and nt hold 4 ints (Rc, Rg, Lc, Lg) */
int jdsg_swd_data_table(const jdsg_swd_data *d, int wave, float *t, int *kmax);
int jdsg_swd_data_arrays(const jdsg_swd_data *d, int *pair_off, int *wave, int *mode, int *period,
                         float *src_colat, float *src_lon, float *rcv_colat, float *rcv_lon,
                         float *dist, float *obst);
void jdsg_swd_data_free(jdsg_swd_data *d);
const char *jdsg_io_last_error(void);

/* ---- the model file (host code) ---- */

/* DSurfTomo::read_model (src/surftomo/io.cpp:27-114; JSurfGTomo::read_model,
 * src/JSurfGTomo/JSurfGtomo.cpp:57-) split in two so the caller sizes the
 * buffers: the header gives nx (nlat), ny (nlon), nz, the origin goxd (lat),
 * gozd (lon) and the spacing dvxd, dvzd; then dep[nz] and the column-major
 * vs[nx*ny*nz] (i fastest), the layout jdsg_surftime_* take.  0 on success,
 * else nonzero with jdsg_io_last_error (the reference prints and exits; a
 * file with fewer values than the grid is an error here). */
int jdsg_read_model_header(const char *path, int *nx, int *ny, int *nz, float *goxd, float *gozd,
                           float *dvxd, float *dvzd);
int jdsg_read_model(const char *path, int nx, int ny, int nz, float *dep, float *vs);
/* The per-iteration model snapshot "lon lat dep vs" (%g), k slowest then j
 * then i, with lat[i] = goxd - i*dvxd and lon[j] = gozd + j*dvzd
 * (src/surftomo/main.cpp:53-59, :113-120). */
int jdsg_write_model(const char *path, int nx, int ny, int nz, float goxd, float gozd, float dvxd,
                     float dvzd, const float *dep, const float *vs);

/* ---- the nonlinear model update (device-resident NonlinOPT) ---- */

/* NonlinOPT<T> (include/nonlinear_opt.hpp:15-390): CG (Polak-Ribiere) or
 * L-BFGS (5 pairs) directions in log-model space, the trial step capped at
 * max_step relative change, x1 = x * exp(step * direc), and the bisecting
 * Wolfe line search.  The history the reference keeps in grad/direc/logmd/
 * syn .bin files is a ring in device memory.  The model object's calls are
 * callbacks on device buffers, issued on the handle's stream:
 *   grad   -- T::compute_grad(x, dsyn, grad): n floats in, m + n out
 *   misfit -- T::compute_misfit(dsyn)
 *   smooth -- T::smoothing(direc), in place (may be null: no smoothing).
 * A callback returns nonzero to abort the update.  jdsg_nlopt_update is
 * NonlinOPT::update (:271-389): x (n device floats) is replaced by the
 * accepted model, dsyn (m device floats) holds the synthetics of the model
 * on entry; method 0 "CG", 1 "LBFGS".  0 on success; 2 when 100 trial
 * steps meet no Wolfe condition (the reference loops on); else nonzero,
 * with jdsg_nlopt_last_error. */
typedef struct jdsg_nlopt jdsg_nlopt;
typedef int (*jdsg_nlopt_grad_fn)(void *user, const float *x, float *dsyn, float *grad, void *stream);
typedef float (*jdsg_nlopt_misfit_fn)(void *user, const float *dsyn, void *stream);
typedef int (*jdsg_nlopt_smooth_fn)(void *user, float *direc, void *stream);
/* NonlinOPT(outdir, max_step, iter_start, iter_current) (:36-48).  work: null
 * (the library allocates one block) or jdsg_nlopt_work_floats(n, m) device
 * floats, 8-byte aligned, kept by the caller until jdsg_nlopt_destroy; every
 * buffer a callback sees lies in it (or is the caller's x / dsyn). */
long long jdsg_nlopt_work_floats(int n, int m);
int jdsg_nlopt_create(int n, int m, float max_step, int iter_start, int iter_current, float *work,
                      void *stream, jdsg_nlopt **out);
int jdsg_nlopt_update(jdsg_nlopt *h, float *x, float *dsyn, int method, jdsg_nlopt_grad_fn grad,
                      jdsg_nlopt_misfit_fn misfit, jdsg_nlopt_smooth_fn smooth, void *user);
/* iteration, ITER_START, trial steps of the last update, STEP_FAC, and the
 * misfits at the start and at the last trial of the last update */
void jdsg_nlopt_state(const jdsg_nlopt *h, int *iter, int *iter_start, int *trials,
                      float *step_fac, float *fcost, float *fcost1);
void jdsg_nlopt_destroy(jdsg_nlopt *h);
const char *jdsg_nlopt_last_error(void);

/* FMST: first-arrival traveltimes and ray Frechet weights of the station
 * pairs on the GPU.  Replaces the per-pair sequence fmst_init(nx, ny, lat0,
 * lon0, dlat, dlon) / fmst_run(&vc(0,idx), srcx, srcz, rcx, rcz, nr, ttime) /
 * fmst_reset(&vout(0,idx)) / fmst_raypath(srcx, srcz, rcx[i], rcz[i], fdm) /
 * fmst_finalize() (fmst.f90:9, :108, :334, :350, :84) that
 * SurfTime::travel_time, frechet_matrix and compute_grad run per pair
 * (SurfaceWave.cpp:524-536, 612-624, 740-752).
 * jdsg_fmst_create takes host arrays: the grid of fmst_init (degrees); per
 * pair its period index (get_period_index), source colatitude/longitude
 * (rad), receivers rcv_off[p] .. rcv_off[p+1]-1 of rcv_colat/rcv_lon (rad),
 * and ray_time[p] != 0 where the receivers' times are the ray-integrated
 * group times (Rg, Lg; may be null).
 * jdsg_fmst_run: vc, vout the device maps (nx*ny, kmax) column-major as
 * jdsg_surftime_map / _kernel fill them; mode 0 traces rays only for
 * ray-time pairs (travel_time), mode 1 for every receiver and keeps the
 * weights (frechet_matrix, compute_grad).  Device outputs: ttime[nrcv] the
 * receiver times the reference reports, tsyn[nrcv] the ray-integrated times
 * (0 where no ray was traced), rstatus[nrcv], pstatus[npair] (0 ok; 1 source
 * or receiver outside the model, where the reference stops; 2 the band
 * outgrew the reference's heap of NINT(0.5*nnx*nnz) entries; 3 the ray did
 * not reach the source).  Synchronises the stream.
 * jdsg_fmst_rays: the last mode-1 run's weights fdm(ix, iy) per receiver in
 * the sparse form jdsg_frechet_rows takes: device ray_off[nrcv+1], ray_cell
 * (n = ix + iy*nx increasing), ray_w; fails with *nnz set to the size needed
 * when that exceeds cap.  Errors: jdsg_fmst_last_error. */
typedef struct jdsg_fmst jdsg_fmst;
int jdsg_fmst_create(int nx, int ny, float lat0, float lon0, float dlat, float dlon, int npair,
                     const int *idx, const float *src_colat, const float *src_lon,
                     const int *rcv_off, const float *rcv_colat, const float *rcv_lon,
                     const int *ray_time, jdsg_fmst **h);
int jdsg_fmst_sizes(const jdsg_fmst *h, int *npair, int *nrcv, int *nplane);
int jdsg_fmst_run(jdsg_fmst *h, const double *vc, const double *vout, int kmax, int mode,
                  float *ttime, float *tsyn, int *rstatus, int *pstatus, void *stream);
int jdsg_fmst_rays(jdsg_fmst *h, long long *ray_off, int *ray_cell, float *ray_w, long long cap,
                   long long *nnz, void *stream);
void jdsg_fmst_destroy(jdsg_fmst *h);
const char *jdsg_fmst_last_error(void);
/* The SurfTime methods that run FMST, one call each on the device model vs
 * (nx,ny,nz) column-major float, with the grid/tables handle and the pairs'
 * FMST handle (same nx, ny); data[nrcv], rstatus[nrcv] device arrays in the
 * order of the pairs' receivers (the data vector's order); errors through
 * jdsg_fmst_last_error.
 * jdsg_surftime_travel_time   SurfTime::travel_time (SurfaceWave.cpp:509-541):
 *   get_2d_map, then the receiver times (ray times for ray-time pairs).
 * jdsg_surftime_frechet       SurfTime::frechet_matrix (:551-676): get_1d_kernel,
 *   every receiver's ray, data, and the rows in CSR as jdsg_frechet_rows
 *   writes them (row_off[nrcv+1], col, val, capacity cap, *nnz on the host);
 *   jdsg_write_frechet then gives the reference's file.
 * jdsg_surftime_compute_grad  SurfTime::compute_grad (:695-781): as frechet,
 *   then grad[(nx-2)(ny-2)(nz-1)] = sum of rows times (data - obst), times
 *   vs at the cell. */
int jdsg_surftime_travel_time(jdsg_surftime *st, jdsg_fmst *fm, const float *vs, float *data,
                              int *rstatus, void *stream);
int jdsg_surftime_frechet(jdsg_surftime *st, jdsg_fmst *fm, const float *vs, float *data,
                          int *rstatus, long long *row_off, int *col, float *val, long long cap,
                          long long *nnz, void *stream);
int jdsg_surftime_compute_grad(jdsg_surftime *st, jdsg_fmst *fm, const float *vs,
                               const float *obst, float *data, float *grad, int *rstatus,
                               void *stream);

#ifdef __cplusplus
}
#endif
#endif
