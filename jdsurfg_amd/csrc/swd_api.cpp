// swd_api.cpp -- C ABI (include/jdsg_swd.h): batch plans and the reference's
// per-column entry points, all executed by the HIP kernels of this library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jdsg_swd.h"
#include "swd_kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(const std::string &m) {
  g_err = m;
  return 1;
}

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) return fail(std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

#define HIPCHK_P(x)                                                            \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) return fail(std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

int default_waves() {
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) ncu = p.multiProcessorCount;
  }
  return ncu * 16;
}

}  // namespace

struct jdsg_plan {
  int ncol = 0, ncol_pad = 0, nmax = 0, kmax = 0, kpad = 0;
  int group = 0, layout = 0, iflsph = 1, nvar = 1;
  long long nent = 0, nkern = 0, nchain = 0, nitem[2] = {0, 0};
  bool uniform = false;
  int uK = 0, unR = 0, unL = 0;
  std::vector<long long> ent_off, kern_off;
  std::vector<void *> allocs;
  // device
  int *layer_off_d = nullptr, *per_off_d = nullptr, *nmode_d = nullptr;
  long long nlay_total = 0;  // layer entries of all columns
  long long *ent_off_d = nullptr, *kern_off_d = nullptr, *items_d[2] = {nullptr, nullptr};
  int *corder_d = nullptr;  // uneven batches: columns deepest first, the order chains are taken
  PrepOut prep{};
  ColumnModel cm{};
  double *cflat = nullptr, *cscr = nullptr, *escr = nullptr;
  int *cstat = nullptr, *ierr = nullptr;
  unsigned *counters = nullptr;
  unsigned long long *nevals = nullptr;
  int chain_waves = 0, eigen_waves = 0;
  double tol = 1.0e-6;
  int fallback = 2;  // chains of the open roots: see jdsg_plan_set_fallback
  // surrogate path buffers (allocated by jdsg_plan_enable_surrogate)
  double *known = nullptr, *cen = nullptr;
  signed char *sg = nullptr;
  double *brk = nullptr;
  unsigned *chain_ev = nullptr;  // per-chain evaluation counts (diagnostics)
  unsigned *sfail = nullptr;
  // suspended long scans of the exact-root fallback
  char *susp_saved = nullptr;
  double *susp_cs = nullptr;
  ScanJob *susp_jobs = nullptr;
  unsigned *susp_njobs = nullptr;
  const double *known_in = nullptr;
  // the surrogate network bound by jdsg_plan_set_network (weights are the
  // caller's; the buffers are the plan's)
  struct Net {
    int hid = 0;
    const uint16_t *W[5] = {};
    const float *b[5] = {};
    float b5 = 0.0f;
    double scale = 0.0, shift = 0.0;
    float *ext = nullptr, *colf = nullptr, *acol = nullptr;
    uint16_t *Xc = nullptr, *Y[2] = {nullptr, nullptr};
    int *work = nullptr;
    long long work_len = 0, chunk = 0;
    double *pred = nullptr;
  };
  std::unique_ptr<Net> net;
  // explicit per-entry periods for the eigen stage (per-column ABI only)
  double *tv = nullptr;
  double *eig[4] = {nullptr, nullptr, nullptr, nullptr};
  // timing (event records are skipped while a graph is being captured)
  bool timing = false;
  bool capturing = false;
  hipEvent_t ev[5] = {};
  float last_ms[4] = {0, 0, 0, 0};
  // the eigen pass kernels' own time per wave type (timing on)
  std::vector<hipEvent_t> pev[2];
  int pass_n[2] = {0, 0};
  float pass_ms[2] = {0, 0};
  int pass_events(int w, int n) {
    while ((int)pev[w].size() < 2 * n) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return fail("hipEventCreate");
      pev[w].push_back(e);
    }
    return 0;
  }
  int finish_timing() {
    HIPCHK_P(hipEventSynchronize(ev[4]));
    for (int i = 0; i < 4; ++i) HIPCHK_P(hipEventElapsedTime(&last_ms[i], ev[i], ev[i + 1]));
    for (int w = 0; w < 2; ++w) {
      pass_ms[w] = 0.0f;
      for (int i = 0; i < pass_n[w]; ++i) {
        float t = 0.0f;
        HIPCHK_P(hipEventElapsedTime(&t, pev[w][2 * i], pev[w][2 * i + 1]));
        pass_ms[w] += t;
      }
    }
    return 0;
  }

  template <class T>
  int dalloc(T **p, size_t n) {
    if (n == 0) n = 1;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return fail(std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(q);
    *p = (T *)q;
    return 0;
  }
  ~jdsg_plan() {
    for (void *p : allocs) (void)hipFree(p);
    if (timing)
      for (auto &e : ev) (void)hipEventDestroy(e);
    for (auto &v : pev)
      for (auto &e : v) (void)hipEventDestroy(e);
  }
  int size_scratch() {
    if (cscr) {
      (void)hipFree(cscr);
      allocs.erase(std::find(allocs.begin(), allocs.end(), (void *)cscr));
      cscr = nullptr;
    }
    if (escr) {
      (void)hipFree(escr);
      allocs.erase(std::find(allocs.begin(), allocs.end(), (void *)escr));
      escr = nullptr;
    }
    if (dalloc(&cscr, (size_t)chain_waves * 64 * (2 * kpad + 22))) return 1;
    if (dalloc(&escr, (size_t)eigen_waves * 64 * (18 * (size_t)nmax + 9))) return 1;
    return 0;
  }
};

extern "C" const char *jdsg_last_error(void) { return g_err.c_str(); }

namespace {

// Host part of a plan: entry and kernel offsets, sizes, launch widths and
// (for uneven batches) the per-wave item lists.  No device work.
int plan_host(jdsg_plan *P, int ncol, const int *layer_off, const int *per_off,
              const int *nmode, int group, int layout, int iflsph,
              std::vector<long long> items[2]) {
  if (ncol <= 0) return fail("ncol must be positive");
  if (layer_off[0] != 0 || per_off[0] != 0) return fail("offsets must start at 0");
  P->ncol = ncol;
  P->ncol_pad = (ncol + 63) / 64 * 64;
  P->group = group ? 1 : 0;
  P->nvar = group ? 3 : 1;
  P->layout = layout;
  P->iflsph = iflsph;
  P->ent_off.resize(ncol + 1);
  P->kern_off.resize(ncol + 1);
  long long e = 0, q = 0;
  int N0 = layer_off[1] - layer_off[0], K0 = per_off[1] - per_off[0];
  bool uni = true;
  for (int c = 0; c < ncol; ++c) {
    int N = layer_off[c + 1] - layer_off[c];
    int K = per_off[c + 1] - per_off[c];
    int nR = nmode[2 * c], nL = nmode[2 * c + 1];
    if (N < 2) return fail("every column needs at least one layer and a half-space");
    if (K < 0 || nR < 0 || nL < 0) return fail("negative period or mode count");
    if (nR > JDSG_MAXMODE || nL > JDSG_MAXMODE) return fail("at most 8 modes per wave");
    P->nmax = std::max(P->nmax, N);
    P->kmax = std::max(P->kmax, K);
    P->ent_off[c] = e;
    P->kern_off[c] = q;
    long long ne = (long long)(nR + nL) * K;
    e += ne;
    q += ne * N;
    P->nitem[0] += (long long)nR * K;
    P->nitem[1] += (long long)nL * K;
    if (N != N0 || K != K0 || nR != nmode[0] || nL != nmode[1]) uni = false;
  }
  P->ent_off[ncol] = e;
  P->kern_off[ncol] = q;
  P->nent = e;
  P->nkern = q;
  P->uniform = uni;
  P->uK = K0;
  P->unR = nmode[0];
  P->unL = nmode[1];
  if (layout == 0 && !uni) return fail("grid layout needs a uniform batch");
  P->kpad = std::max(1, P->kmax);
  P->nchain = (long long)ncol * 2 * P->nvar;
  int dw = default_waves();
  P->chain_waves = (int)std::min<long long>(dw, ((long long)ncol * P->nvar + 63) / 64);
  // eigen chunks of 64 waves per CU: each pass then runs eight rounds of two
  // waves per SIMD, so the last round's tail is a small part of it (measured
  // on 256x256 columns x 60 periods: 32 waves per CU 201 ms, 64 194 ms, 128
  // 194 ms, 256 197 ms for the Rayleigh stage; lane scratch 9.9 GB at 64)
  P->eigen_waves =
      (int)std::min<long long>(4LL * dw, (std::max(P->nitem[0], P->nitem[1]) + 63) / 64);
  P->chain_waves = std::max(1, P->chain_waves);
  P->eigen_waves = std::max(1, P->eigen_waves);
  P->nlay_total = layer_off[ncol];
  if (!uni) {
    // columns deepest first (stable): a wave's lanes then work on columns of
    // about the same depth, whose layer loops end together, and the longest
    // work starts first; every kernel that walks the list treats each entry
    // on its own, so the order changes no result
    std::vector<int> corder(ncol);
    for (int c = 0; c < ncol; ++c) corder[c] = c;
    std::stable_sort(corder.begin(), corder.end(), [&](int a, int b) {
      return layer_off[a + 1] - layer_off[a] > layer_off[b + 1] - layer_off[b];
    });
    for (int w = 0; w < 2; ++w) {
      items[w].clear();
      items[w].reserve(P->nitem[w]);
      for (int c : corder) {
        int K = per_off[c + 1] - per_off[c];
        int nR = nmode[2 * c], nL = nmode[2 * c + 1];
        int M = w == 0 ? nR : nL, off = w == 0 ? 0 : nR;
        for (int m = 0; m < M; ++m)
          for (int k = 0; k < K; ++k)
            items[w].push_back(P->ent_off[c] + (long long)(off + m) * K + k);
      }
    }
  }
  return 0;
}

// Device arrays of a plan placed by `take(pointer, count)` -- hipMalloc in
// jdsg_plan_create, a carve of the static scratch in the allocation-free
// form.  Returns nonzero when an array cannot be placed.
template <class Take>
int plan_place(jdsg_plan *P, bool uneven, Take take) {
  size_t cp = (size_t)P->ncol_pad * P->nmax;
  int ncol = P->ncol;
  if (take(&P->layer_off_d, ncol + 1) || take(&P->per_off_d, ncol + 1) ||
      take(&P->nmode_d, 2 * ncol) || take(&P->ent_off_d, ncol + 1) ||
      take(&P->kern_off_d, ncol + 1) || take(&P->prep.chainR, cp) ||
      take(&P->prep.chainL, cp) || take(&P->prep.cc, ncol) || take(&P->prep.betmx, ncol) ||
      take(&P->prep.llw, ncol) || take(&P->prep.nlay, ncol) || take(&P->prep.zd, cp) ||
      take(&P->prep.za, cp) || take(&P->prep.zb, cp) || take(&P->prep.rho0, cp) ||
      take(&P->prep.rtpR, cp) || take(&P->prep.rtpL, cp) || take(&P->prep.vtp, cp) ||
      take(&P->prep.dtp, cp) || take(&P->cflat, (size_t)P->nvar * P->nent) ||
      take(&P->cstat, (size_t)P->nvar * P->nent) || take(&P->ierr, P->nchain) ||
      take(&P->counters, 4) || take(&P->nevals, 4))
    return 1;
  if (uneven)
    for (int w = 0; w < 2; ++w)
      if (take(&P->items_d[w], P->nitem[w])) return 1;
  return 0;
}

// Copies the descriptors to the device arrays and binds the column model.
int plan_upload(jdsg_plan *P, const int *layer_off, const int *per_off, const int *nmode,
                const std::vector<long long> items[2], hipStream_t st) {
  int ncol = P->ncol;
  HIPCHK(hipMemcpyAsync(P->layer_off_d, layer_off, sizeof(int) * (ncol + 1),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P->per_off_d, per_off, sizeof(int) * (ncol + 1), hipMemcpyHostToDevice,
                        st));
  HIPCHK(hipMemcpyAsync(P->nmode_d, nmode, sizeof(int) * 2 * ncol, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P->ent_off_d, P->ent_off.data(), sizeof(long long) * (ncol + 1),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P->kern_off_d, P->kern_off.data(), sizeof(long long) * (ncol + 1),
                        hipMemcpyHostToDevice, st));
  if (!P->uniform)
    for (int w = 0; w < 2; ++w)
      if (!items[w].empty())
        HIPCHK(hipMemcpyAsync(P->items_d[w], items[w].data(), sizeof(long long) * items[w].size(),
                              hipMemcpyHostToDevice, st));
  // the host arrays are the caller's and this function's: wait for the copies
  HIPCHK(hipStreamSynchronize(st));
  P->cm.ncol = ncol;
  P->cm.ncol_pad = P->ncol_pad;
  P->cm.nmax = P->nmax;
  P->cm.nlay = P->prep.nlay;
  P->cm.chainR = P->prep.chainR;
  P->cm.chainL = P->prep.chainL;
  P->cm.cc = P->prep.cc;
  P->cm.betmx = P->prep.betmx;
  P->cm.llw = P->prep.llw;
  P->cm.zd = P->prep.zd;
  P->cm.za = P->prep.za;
  P->cm.zb = P->prep.zb;
  P->cm.rho0 = P->prep.rho0;
  P->cm.rtpR = P->prep.rtpR;
  P->cm.rtpL = P->prep.rtpL;
  P->cm.vtp = P->prep.vtp;
  P->cm.dtp = P->prep.dtp;
  return 0;
}

}  // namespace

extern "C" int jdsg_plan_create(int ncol, const int *layer_off, const int *per_off,
                                const int *nmode, int group, int layout, int iflsph,
                                jdsg_plan **out) {
  *out = nullptr;
  auto P = std::make_unique<jdsg_plan>();
  std::vector<long long> items[2];
  if (plan_host(P.get(), ncol, layer_off, per_off, nmode, group, layout, iflsph, items)) return 1;
  jdsg_plan *Q = P.get();
  if (plan_place(Q, !Q->uniform, [Q](auto **p, size_t n) { return Q->dalloc(p, n); }) ||
      Q->size_scratch())
    return 1;
  if (plan_upload(Q, layer_off, per_off, nmode, items, nullptr)) return 1;
  if (!Q->uniform) {
    std::vector<int> corder(ncol);
    for (int c = 0; c < ncol; ++c) corder[c] = c;
    std::stable_sort(corder.begin(), corder.end(), [&](int a, int b) {
      return layer_off[a + 1] - layer_off[a] > layer_off[b + 1] - layer_off[b];
    });
    if (Q->dalloc(&Q->corder_d, (size_t)ncol)) return 1;
    HIPCHK(hipMemcpy(Q->corder_d, corder.data(), sizeof(int) * ncol, hipMemcpyHostToDevice));
  }
  const char *tenv = getenv("JDSG_TIMING");
  if (tenv && tenv[0] == '1') {
    P->timing = true;
    for (auto &ev : P->ev) (void)hipEventCreate(&ev);
  }
  *out = P.release();
  return 0;
}

extern "C" int jdsg_plan_sizes(const jdsg_plan *P, long long *nent, long long *nkern,
                               int *nmax, int *kmax) {
  if (nent) *nent = P->nent;
  if (nkern) *nkern = P->nkern;
  if (nmax) *nmax = P->nmax;
  if (kmax) *kmax = P->kmax;
  return 0;
}

extern "C" int jdsg_plan_offsets(const jdsg_plan *P, long long *ent_off, long long *kern_off) {
  if (ent_off) memcpy(ent_off, P->ent_off.data(), sizeof(long long) * (P->ncol + 1));
  if (kern_off) memcpy(kern_off, P->kern_off.data(), sizeof(long long) * (P->ncol + 1));
  return 0;
}

extern "C" int jdsg_plan_set_waves(jdsg_plan *P, int cw, int ew) {
  if (cw > 0) P->chain_waves = cw;
  if (ew > 0) P->eigen_waves = ew;
  return P->size_scratch();
}

extern "C" void jdsg_plan_destroy(jdsg_plan *P) { delete P; }

extern "C" int jdsg_plan_set_tolerance(jdsg_plan *P, double tol) {
  if (!(tol > 0.0) || tol > 1.0e-2) return fail("tolerance must be in (0, 1e-2]");
  P->tol = tol;
  return 0;
}

extern "C" int jdsg_plan_set_fallback(jdsg_plan *P, int form) {
  if (form < 0 || form > 2) return fail("fallback form must be 0, 1 or 2");
  P->fallback = form;
  return 0;
}

extern "C" int jdsg_plan_eval_count(const jdsg_plan *P, long long *out) {
  unsigned long long v = 0;
  HIPCHK(hipMemcpy(&v, P->nevals, sizeof(v), hipMemcpyDeviceToHost));
  *out = (long long)v;
  return 0;
}

// Flat roots of the last execute ([nvar*nent]: T, then T(1+.005), T(1-.005);
// float-rounded in the reference mode), device to host.  Synchronises.
// Diagnostics: per-chain evaluation counts and the reference's ierr of
// each root chain (chain g of wave w at [w*ncol*nvar + g], g = variant*ncol
// + column) of the last solve, into host arrays.  Counting starts with the
// first call and costs one store per chain.
extern "C" int jdsg_plan_chain_stats(jdsg_plan *P, unsigned *evals, int *ierr) {
  size_t n = (size_t)2 * P->ncol * P->nvar;
  if (!P->chain_ev) {
    if (P->dalloc(&P->chain_ev, n)) return 1;
    HIPCHK(hipMemset(P->chain_ev, 0, sizeof(unsigned) * n));
    return 0;
  }
  HIPCHK(hipDeviceSynchronize());
  if (evals) HIPCHK(hipMemcpy(evals, P->chain_ev, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
  if (ierr) HIPCHK(hipMemcpy(ierr, P->ierr, sizeof(int) * n, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int jdsg_plan_flat_roots(const jdsg_plan *P, double *host) {
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(host, P->cflat, sizeof(double) * P->nvar * P->nent, hipMemcpyDeviceToHost));
  return 0;
}

// Prepared root-search model of the plan's columns (after an execute):
// float4 {d, a, b, rho} per layer for Rayleigh and Love ([layer*ncol_pad+col]),
// starting velocity and betmx per column.  For tests.
extern "C" int jdsg_plan_chain_model(const jdsg_plan *P, float *R, float *L, float *cc,
                                     float *betmx, int *ncol_pad) {
  size_t n = (size_t)P->ncol_pad * P->nmax;
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(R, P->prep.chainR, sizeof(float4) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(L, P->prep.chainL, sizeof(float4) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(cc, P->prep.cc, sizeof(float) * P->ncol, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(betmx, P->prep.betmx, sizeof(float) * P->ncol, hipMemcpyDeviceToHost));
  *ncol_pad = P->ncol_pad;
  return 0;
}

extern "C" int jdsg_plan_pass_timings(const jdsg_plan *P, float *ms2, int *n2) {
  for (int w = 0; w < 2; ++w) {
    if (ms2) ms2[w] = P->pass_ms[w];
    if (n2) n2[w] = P->pass_n[w];
  }
  return P->timing ? 0 : fail("timing disabled (set JDSG_TIMING=1 before plan creation)");
}

extern "C" int jdsg_plan_timings(const jdsg_plan *P, float *ms4) {
  for (int i = 0; i < 4; ++i) ms4[i] = P->last_ms[i];
  return P->timing ? 0 : fail("timing disabled (set JDSG_TIMING=1 before plan creation)");
}

namespace {

// Prepare + chains.  igr/retry as in surfdisp96 / surfdisp.
int run_front(jdsg_plan *P, const float *thk, const float *vp, const float *vs,
              const float *rho, const double *periods, int igr, int retry, hipStream_t st) {
  if (P->timing && !P->capturing) HIPCHK(hipEventRecord(P->ev[0], st));
  jdsg_launch_prep(P->ncol, P->ncol_pad, P->layer_off_d, thk, vp, vs, rho, P->iflsph, P->prep,
                   st);
  jdsg_launch_fill_u32(P->counters, 4, 0u, st);
  jdsg_launch_fill_u32((unsigned *)P->nevals, 2, 0u, st);
  jdsg_launch_fill_u32((unsigned *)P->cstat, (long long)P->nvar * P->nent, 0u, st);
  if (P->timing && !P->capturing) HIPCHK(hipEventRecord(P->ev[1], st));
  ChainArgs A{};
  A.corder = P->corder_d;
  A.cm = P->cm;
  A.ent_off = P->ent_off_d;
  A.per_off = P->per_off_d;
  A.periods = periods;
  A.nmode = P->nmode_d;
  A.nvar = P->nvar;
  A.igr = igr;
  A.retry = retry;
  A.nent = P->nent;
  A.kmax_pad = P->kpad;
  A.cflat = P->cflat;
  A.cstat = P->cstat;
  A.ierr_out = P->ierr;
  A.cscr = P->cscr;
  A.counter = P->counters + 0;
  A.nchain = P->nchain;
  A.nevals = P->nevals;
  A.tol = P->tol;
  A.known = P->known_in;
  A.chain_evals = P->chain_ev;
  A.any_wave[0] = P->nitem[0] > 0;
  A.any_wave[1] = P->nitem[1] > 0;
  jdsg_launch_chains(A, P->chain_waves, st);
  HIPCHK(hipGetLastError());
  if (P->timing && !P->capturing) HIPCHK(hipEventRecord(P->ev[2], st));
  return 0;
}

int run_eigen(jdsg_plan *P, const double *periods, double *c, double *u, int *status,
              double *const dc[4], double *const du[4], hipStream_t st) {
  for (int w = 0; w < 2; ++w) {
    if (P->nitem[w] == 0) {
      if (P->timing && !P->capturing) HIPCHK(hipEventRecord(P->ev[3 + w], st));
      continue;
    }
    EigenArgs E{};
    E.cm = P->cm;
    E.ent_off = P->ent_off_d;
    E.per_off = P->per_off_d;
    E.periods = periods;
    E.nmode = P->nmode_d;
    E.wave = w;
    E.group = P->group;
    E.iflsph = P->iflsph;
    E.nent = P->nent;
    E.cflat = P->cflat;
    E.cstat = P->cstat;
    E.nitem = P->nitem[w];
    E.item_ent = P->uniform ? nullptr : P->items_d[w];
    E.grid_K = P->uK;
    E.grid_nR = P->unR;
    E.grid_nL = P->unL;
    E.c_out = c;
    E.u_out = u;
    E.status = status;
    for (int a = 0; a < 4; ++a) {
      E.dc[a] = dc[a];
      E.du[a] = du[a];
      E.eig[a] = P->eig[a];
    }
    E.layout = P->layout;
    E.kern_off = P->kern_off_d;
    E.scr = P->escr;
    E.nmax = P->nmax;
    E.counter = P->counters + 2 + w;
    E.tv = P->tv;
    int waves = (int)std::min<long long>(P->eigen_waves, (P->nitem[w] + 63) / 64);
    hipEvent_t *pev = nullptr;
    P->pass_n[w] = 0;
    if (P->timing && !P->capturing) {
      int np = jdsg_eigen_pass_launches(E, waves);
      if (P->pass_events(w, np)) return 1;
      pev = P->pev[w].data();
      P->pass_n[w] = np;
    }
    jdsg_launch_eigen(E, waves, st, pev);
    HIPCHK(hipGetLastError());
    if (P->timing && !P->capturing) HIPCHK(hipEventRecord(P->ev[3 + w], st));
  }
  return 0;
}

}  // namespace

extern "C" int jdsg_plan_execute(jdsg_plan *P, const float *thk, const float *vp,
                                 const float *vs, const float *rho, const double *periods,
                                 double *c, double *u, int *status, double *dcda, double *dcdb,
                                 double *dcdr, double *dcdh, double *duda, double *dudb,
                                 double *dudr, double *dudh, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (!thk || !vp || !vs || !rho || !periods || !c || !u || !status)
    return fail("null input or scalar output pointer");
  if (run_front(P, thk, vp, vs, rho, periods, 0, 1, st)) return 1;
  double *dc[4] = {dcda, dcdb, dcdr, dcdh};
  double *du[4] = {duda, dudb, dudr, dudh};
  if (run_eigen(P, periods, c, u, status, dc, du, st)) return 1;
  if (P->timing && !P->capturing) {
    if (P->finish_timing()) return 1;
  }
  return 0;
}

// ============================================================================
// Allocation-free batch solve: jdsg_plan_create + jdsg_plan_execute with
// every device array of the plan carved from one file-scope device array.
namespace {
constexpr size_t kStaticScratchBytes = 64ull << 20;
}  // namespace

static __device__ __align__(256) char scratch[kStaticScratchBytes];

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
int static_scratch(void **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("static scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(std::string("static scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" long long jdsg_static_scratch_bytes(void) { return (long long)kStaticScratchBytes; }

extern "C" int jdsg_plan_execute_static(int ncol, const int *layer_off, const int *per_off,
                                        const int *nmode, int group, int layout, int iflsph,
                                        const float *thk, const float *vp, const float *vs,
                                        const float *rho, const double *periods, double *c,
                                        double *u, int *status, double *dcda, double *dcdb,
                                        double *dcdr, double *dcdh, double *duda, double *dudb,
                                        double *dudr, double *dudh, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (!thk || !vp || !vs || !rho || !periods || !c || !u || !status)
    return fail("null input or scalar output pointer");
  jdsg_plan P;  // host descriptor; its device arrays live in the static scratch
  std::vector<long long> items[2];
  if (plan_host(&P, ncol, layer_off, per_off, nmode, group, layout, iflsph, items)) return 1;
  void *base = nullptr;
  if (static_scratch(&base)) return 1;
  size_t off = 0;
  auto take = [&](auto **p, size_t n) -> int {
    using T = std::remove_reference_t<decltype(**p)>;
    size_t b = (std::max<size_t>(n, 1) * sizeof(T) + 255) / 256 * 256;
    if (off + b > kStaticScratchBytes) return 1;
    *p = (T *)((char *)base + off);
    off += b;
    return 0;
  };
  auto too_big = [&]() {
    return fail("batch too large for the allocation-free solve: its plan needs more than the " +
                std::to_string(kStaticScratchBytes >> 20) +
                " MiB static scratch (use jdsg_plan_create / jdsg_plan_execute)");
  };
  if (plan_place(&P, !P.uniform, take)) return too_big();
  // lane scratch of the chain and eigen launches from what is left, at most
  // the plan's default widths (results do not depend on the widths: chains
  // and eigen entries are taken one lane each from work counters)
  size_t cwb = (size_t)64 * (2 * P.kpad + 22) * sizeof(double) + 256;
  size_t ewb = (size_t)64 * (18 * (size_t)P.nmax + 9) * sizeof(double) + 256;
  size_t left = kStaticScratchBytes - off;
  long long cw = std::min<long long>(P.chain_waves, (long long)(left / 2 / cwb));
  long long ew = cw > 0 ? std::min<long long>(P.eigen_waves, (long long)((left - cw * cwb) / ewb))
                        : 0;
  if (cw < 1 || ew < 1) return too_big();
  P.chain_waves = (int)cw;
  P.eigen_waves = (int)ew;
  if (take(&P.cscr, (size_t)cw * 64 * (2 * P.kpad + 22)) ||
      take(&P.escr, (size_t)ew * 64 * (18 * (size_t)P.nmax + 9)))
    return too_big();
  if (plan_upload(&P, layer_off, per_off, nmode, items, st)) return 1;
  if (run_front(&P, thk, vp, vs, rho, periods, 0, 1, st)) return 1;
  double *dc[4] = {dcda, dcdb, dcdr, dcdh};
  double *du[4] = {duda, dudb, dudr, dudh};
  return run_eigen(&P, periods, c, u, status, dc, du, st);
}

// ============================================================================
// Per-column reference ABI (host arrays), each call one small plan.
namespace {

struct DevBuf {
  std::vector<void *> p;
  ~DevBuf() {
    for (void *q : p) (void)hipFree(q);
  }
  template <class T>
  T *up(const T *h, size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
    p.push_back(q);
    if (h && n) (void)hipMemcpy(q, h, sizeof(T) * n, hipMemcpyHostToDevice);
    return (T *)q;
  }
};

void report(const char *where) {
  fprintf(stderr, "libjdsg_swd: %s failed: %s\n", where, g_err.c_str());
}

// Per-thread cache for the per-column entry points: the reference calls
// them from its OpenMP column loop (SurfaceWave.cpp:469), so each host
// thread keeps its own small plans (keyed by the column's sizes) and one
// growing device arena, and a call allocates nothing once its sizes have
// been seen.  The cache lives until the process ends.
struct ColumnCache {
  struct Entry {
    long long key;
    jdsg_plan *plan;
    unsigned long long used;
  };
  std::vector<Entry> plans;
  unsigned long long tick = 0;
  void *arena = nullptr;
  size_t cap = 0;
  jdsg_plan *get(int N, int K, int nR, int nL, int group, int iflsph) {
    long long key = ((((long long)N * 4096 + K) * 16 + nR) * 16 + nL) * 4 + group * 2 + iflsph;
    ++tick;
    for (auto &e : plans)
      if (e.key == key) {
        e.used = tick;
        return e.plan;
      }
    int lo[2] = {0, N}, po[2] = {0, K}, nm[2] = {nR, nL};
    jdsg_plan *P = nullptr;
    if (jdsg_plan_create(1, lo, po, nm, group, 1, iflsph, &P)) return nullptr;
    if (plans.size() >= 16) {  // drop the least recently used plan
      auto it = std::min_element(plans.begin(), plans.end(),
                                 [](const Entry &a, const Entry &b) { return a.used < b.used; });
      jdsg_plan_destroy(it->plan);
      plans.erase(it);
    }
    plans.push_back({key, P, tick});
    return P;
  }
  // device arena of at least n bytes (contents not kept across growth)
  char *space(size_t n) {
    if (n > cap) {
      if (arena) (void)hipFree(arena);
      arena = nullptr;
      cap = 0;
      size_t want = std::max<size_t>(n, 1 << 16);
      if (hipMalloc(&arena, want) != hipSuccess) return nullptr;
      cap = want;
    }
    return (char *)arena;
  }
};

ColumnCache &column_cache() {
  thread_local ColumnCache *c = new ColumnCache();
  return *c;
}

// Places arrays one after another (256-byte aligned) in an arena.
struct Carve {
  char *base;
  size_t off = 0;
  template <class T>
  T *take(size_t n) {
    T *p = (T *)(base + off);
    off += (n * sizeof(T) + 255) / 256 * 256;
    return p;
  }
};
size_t carve_size(std::initializer_list<size_t> bytes) {
  size_t s = 0;
  for (size_t b : bytes) s += (b + 255) / 256 * 256;
  return s;
}

// One column, chains only: returns flat roots for modes 1..mode (mode-major)
int column_chains(const float *thk, const float *vp, const float *vs, const float *rho, int N,
                  int iflsph, int iwave, int mode, int igr, int retry, int kmax,
                  const double *t, std::vector<double> &cg_all, std::vector<int> &st_all,
                  int *ierr) {
  ColumnCache &CC = column_cache();
  jdsg_plan *P = CC.get(N, kmax, iwave == 2 ? mode : 0, iwave == 1 ? mode : 0, 0, iflsph);
  if (!P) return 1;
  char *a = CC.space(carve_size({4 * sizeof(float) * N, sizeof(double) * kmax}));
  if (!a) return fail("device allocation");
  Carve cv{a};
  float *d[4] = {cv.take<float>(N), cv.take<float>(N), cv.take<float>(N), cv.take<float>(N)};
  double *dt = cv.take<double>(kmax);
  const float *h[4] = {thk, vp, vs, rho};
  for (int i = 0; i < 4; ++i) HIPCHK(hipMemcpy(d[i], h[i], sizeof(float) * N, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dt, t, sizeof(double) * kmax, hipMemcpyHostToDevice));
  if (run_front(P, d[0], d[1], d[2], d[3], dt, igr, retry, nullptr)) return 1;
  cg_all.assign((size_t)mode * kmax, 0.0);
  st_all.assign((size_t)mode * kmax, 0);
  HIPCHK(hipMemcpy(cg_all.data(), P->cflat, sizeof(double) * mode * kmax, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(st_all.data(), P->cstat, sizeof(int) * mode * kmax, hipMemcpyDeviceToHost));
  // the chain kernel of wave W writes chain g's ierr at W * ncol * nvar + g
  int ie = 0;
  HIPCHK(hipMemcpy(&ie, P->ierr + (iwave == 1 ? P->nchain / 2 : 0), sizeof(int),
                   hipMemcpyDeviceToHost));
  if (ierr) *ierr = ie;
  return 0;
}

// One column, eigen stage from given flat roots.  nv = 1 (sregn96/slegn96)
// or 3 (sregnpu/slegnpu, explicit periods).
int column_eigen(int wave, const float *thk, const float *vp, const float *vs, const float *rho,
                 int N, int iflsph, const double *tv, const double *cv, int nv, double *cp,
                 double *cg, double *eig[4], double *dc[4], double *du[4]) {
  ColumnCache &CC = column_cache();
  jdsg_plan *P = CC.get(N, 1, wave == 0 ? 1 : 0, wave == 1 ? 1 : 0, nv == 3, iflsph);
  if (!P) return 1;
  size_t nout = 2 + 12 * (size_t)N;
  char *a = CC.space(carve_size({4 * sizeof(float) * N, 3 * sizeof(double),
                                 sizeof(double) * nout, sizeof(int)}));
  if (!a) return fail("device allocation");
  Carve c{a};
  float *d[4] = {c.take<float>(N), c.take<float>(N), c.take<float>(N), c.take<float>(N)};
  double *dtv = c.take<double>(3);
  double *dout = c.take<double>(nout);
  int *dst = c.take<int>(1);
  const float *h[4] = {thk, vp ? vp : vs, vs, rho};
  for (int i = 0; i < 4; ++i) HIPCHK(hipMemcpy(d[i], h[i], sizeof(float) * N, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dtv, tv, sizeof(double) * 3, hipMemcpyHostToDevice));
  for (int q = 0; q < 4; ++q) P->eig[q] = dout + 2 + (size_t)(8 + q) * N;
  P->tv = (nv == 3) ? dtv : nullptr;
  jdsg_launch_prep(1, P->ncol_pad, P->layer_off_d, d[0], d[1], d[2], d[3], iflsph, P->prep,
                   nullptr);
  jdsg_launch_fill_u32(P->counters, 4, 0u, nullptr);
  jdsg_launch_fill_u32((unsigned *)P->cstat, (long long)P->nvar * P->nent, 0u, nullptr);
  HIPCHK(hipMemcpy(P->cflat, cv, sizeof(double) * nv, hipMemcpyHostToDevice));
  double *ddc[4], *ddu[4];
  for (int q = 0; q < 4; ++q) {
    ddc[q] = dout + 2 + (size_t)q * N;
    ddu[q] = dout + 2 + (size_t)(4 + q) * N;
  }
  int rc = run_eigen(P, dtv, dout, dout + 1, dst, ddc, ddu, nullptr);
  P->tv = nullptr;
  for (int q = 0; q < 4; ++q) P->eig[q] = nullptr;
  if (rc) return rc;
  std::vector<double> hb(nout);
  HIPCHK(hipMemcpy(hb.data(), dout, sizeof(double) * nout, hipMemcpyDeviceToHost));
  *cp = hb[0];
  *cg = hb[1];
  for (int q = 0; q < 4; ++q) {
    if (dc[q]) memcpy(dc[q], &hb[2 + q * N], sizeof(double) * N);
    if (du[q]) memcpy(du[q], &hb[2 + (4 + q) * N], sizeof(double) * N);
    if (eig[q]) memcpy(eig[q], &hb[2 + (8 + q) * N], sizeof(double) * N);
  }
  return 0;
}

}  // namespace

extern "C" void surfdisp96_(const float *thkm, const float *vpm, const float *vsm,
                            const float *rhom, int nlayer, int iflsph, int iwave, int mode,
                            int igr, int kmax, const double *t, double *cg, int *ierr) {
  std::vector<double> all;
  std::vector<int> st;
  if (kmax <= 0 || mode <= 0) {
    *ierr = 0;
    return;
  }
  if (column_chains(thkm, vpm, vsm, rhom, nlayer, iflsph, iwave, mode, igr, 0, kmax, t, all, st,
                    ierr)) {
    report("surfdisp96_");
    for (int i = 0; i < kmax; ++i) cg[i] = 0.0;
    *ierr = 1;
    return;
  }
  memcpy(cg, all.data() + (size_t)(mode - 1) * kmax, sizeof(double) * kmax);
}

extern "C" int jdsg_surfdisp(const float *thk, const float *vp, const float *vs,
                             const float *rho, int nlayer, const double *t, double *cg,
                             int kmax, int wavetype, int mode, int sphere, int keep_flat) {
  if (wavetype < 0 || wavetype > 3) return fail("wavetype must be 0..3 (Rc,Rg,Lc,Lg)");
  int iwave = wavetype >= 2 ? 1 : 2;
  int igr = (wavetype == 1 || wavetype == 3) ? 1 : 0;
  std::vector<double> all;
  std::vector<int> st;
  int ierr = 0;
  if (kmax <= 0) return 0;
  if (column_chains(thk, vp, vs, rho, nlayer, sphere ? 1 : 0, iwave, mode + 1, igr, 1, kmax, t,
                    all, st, &ierr))
    return 1;
  memcpy(cg, all.data() + (size_t)mode * kmax, sizeof(double) * kmax);
  for (int i = 0; i < kmax; ++i)
    if (st[(size_t)mode * kmax + i] & JDSG_EXIT_PATH)
      return fail("cannot find dispersion (the reference exits here)");
  if (sphere && !keep_flat) {
    const double ar = 6371.0;
    for (int i = 0; i < kmax; ++i) {
      double omega = 2.0 * M_PI / t[i];
      double tm = (wavetype >= 2) ? 1. + pow(1.5 * cg[i] / (ar * omega), 2)
                                  : 1. + pow(0.5 * cg[i] / (ar * omega), 2);
      tm = sqrt(tm);
      cg[i] = (wavetype == 0 || wavetype == 2) ? cg[i] / tm : cg[i] * tm;
    }
  }
  return 0;
}

extern "C" void sregn96_(const float *thk, const float *vp, const float *vs, const float *rhom,
                         int nlayer, const double *t, double *cp, double *cg, double *dispu,
                         double *dispw, double *stressu, double *stressw, double *dc2da,
                         double *dc2db, double *dc2dh, double *dc2dr, int iflsph) {
  double *eig[4] = {dispu, dispw, stressu, stressw};
  double *dc[4] = {dc2da, dc2db, dc2dr, dc2dh};
  double *du[4] = {nullptr, nullptr, nullptr, nullptr};
  double tv[3] = {*t, 0, 0}, cv[3] = {*cp, 0, 0};
  if (column_eigen(0, thk, vp, vs, rhom, nlayer, iflsph, tv, cv, 1, cp, cg, eig, dc, du))
    report("sregn96_");
}

extern "C" void slegn96_(const float *thk, const float *vs, const float *rhom, int nlayer,
                         const double *t, double *cp, double *cg, double *disp, double *stress,
                         double *dc2db, double *dc2dh, double *dc2dr, int iflsph) {
  double *eig[4] = {disp, nullptr, nullptr, nullptr};
  double *dc[4] = {nullptr, dc2db, dc2dr, dc2dh};
  double *du[4] = {nullptr, nullptr, nullptr, nullptr};
  double tv[3] = {*t, 0, 0}, cv[3] = {*cp, 0, 0};
  if (column_eigen(1, thk, nullptr, vs, rhom, nlayer, iflsph, tv, cv, 1, cp, cg, eig, dc, du))
    report("slegn96_");
  // slegn96.f90:770 copies the displacement into `stress`
  if (stress && disp) memcpy(stress, disp, sizeof(double) * nlayer);
}

extern "C" void sregnpu_(const float *thk, const float *vp, const float *vs, const float *rhom,
                         int nlayer, const double *t, double *cp, double *cg, double *dispu,
                         double *dispw, double *stressu, double *stressw, double *t1,
                         double *cp1, double *t2, double *cp2, double *dc2da, double *dc2db,
                         double *dc2dh, double *dc2dr, double *du2da, double *du2db,
                         double *du2dh, double *du2dr, int iflsph) {
  double *eig[4] = {dispu, dispw, stressu, stressw};
  double *dc[4] = {dc2da, dc2db, dc2dr, dc2dh};
  double *du[4] = {du2da, du2db, du2dr, du2dh};
  double tv[3] = {*t, *t1, *t2}, cv[3] = {*cp, *cp1, *cp2};
  if (column_eigen(0, thk, vp, vs, rhom, nlayer, iflsph, tv, cv, 3, cp, cg, eig, dc, du))
    report("sregnpu_");
}

extern "C" void slegnpu_(float *thk, float *vs, float *rhom, int nlayer, const double *t,
                         double *cp, double *cg, double *disp, double *stress, double *t1,
                         double *cp1, double *t2, double *cp2, double *dc2db, double *dc2dh,
                         double *dc2dr, double *du2db, double *du2dh, double *du2dr,
                         int iflsph) {
  double *eig[4] = {disp, stress, nullptr, nullptr};
  double *dc[4] = {nullptr, dc2db, dc2dr, dc2dh};
  double *du[4] = {nullptr, du2db, du2dr, du2dh};
  double tv[3] = {*t, *t1, *t2}, cv[3] = {*cp, *cp1, *cp2};
  if (column_eigen(1, thk, nullptr, vs, rhom, nlayer, iflsph, tv, cv, 3, cp, cg, eig, dc, du))
    report("slegnpu_");
}

// ============================================================================
// Surrogate (learned bracket) path.
extern "C" int jdsg_plan_roots(jdsg_plan *P, const float *thk, const float *vp, const float *vs,
                               const float *rho, const double *periods, double *c_out,
                               int *stat_out, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (run_front(P, thk, vp, vs, rho, periods, 0, 1, st)) return 1;
  size_t n = (size_t)P->nvar * P->nent;
  HIPCHK(hipMemcpyAsync(c_out, P->cflat, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  if (stat_out)
    HIPCHK(hipMemcpyAsync(stat_out, P->cstat, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
  return 0;
}

// MLP input rows (bf16, 192 wide) for entries e0..e0+nrow-1, or for the
// entry indices rows[0..nrow) when rows is not null; column_pass=1 first
// rebuilds the per-column profile features colf[ncol*148].
extern "C" int jdsg_plan_features(jdsg_plan *P, const float *thk, const float *vp,
                                  const float *vs, const float *rho, const double *periods,
                                  long long e0, long long nrow, const long long *rows, void *X,
                                  float *colf, int column_pass, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (!rows && (e0 < 0 || e0 + nrow > P->nent)) return fail("feature rows out of range");
  if (column_pass)
    jdsg_launch_column_features(P->ncol, P->layer_off_d, thk, vp, vs, rho, colf, st);
  jdsg_launch_entry_features(e0, nrow, P->ncol, P->ent_off_d, P->per_off_d, periods, P->nmode_d,
                             colf, rows, (uint16_t *)X, st);
  HIPCHK(hipGetLastError());
  return 0;
}

// First hidden layer of the network for entries e0..e0+nrow from the
// per-column pre-activations acol (ncol x lda fp32, one GEMM over the column
// feature rows) and the weight columns of the per-entry inputs (ext, 8 x H):
// Y (nrow x H, bf16) = gelu(acol + entry terms).
extern "C" int jdsg_plan_mlp_expand(jdsg_plan *P, const double *periods, long long e0,
                                    long long nrow, const float *acol, int lda, const float *ext,
                                    int H, void *Y, void *stream) {
  if (e0 < 0 || nrow < 0 || e0 + nrow > P->nent) return fail("expand rows out of range");
  if (H <= 0 || H % 8 || H > 1024 || lda < H || lda % 4)
    return fail("expand: H must be a positive multiple of 8, at most 1024 and <= lda, lda a "
                "multiple of 4");
  if (((uintptr_t)acol | (uintptr_t)ext | (uintptr_t)Y) & 15)
    return fail("expand: acol, ext and Y must be 16-byte aligned");
  jdsg_launch_mlp_expand(e0, nrow, P->ncol, P->ent_off_d, P->per_off_d, periods, P->nmode_d, acol,
                         lda, ext, H, (uint16_t *)Y, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return 0;
}

// The whole width-256 network at inference for entries e0..e0+nrow:
// pred[p0 + r] = exp(head(hidden layers(first layer)) * scale + shift).
extern "C" int jdsg_plan_mlp_fused(jdsg_plan *P, const double *periods, long long e0,
                                   long long nrow, const float *acol, int lda, const float *ext,
                                   const void *W2, const void *W3, const void *W4,
                                   const float *b2, const float *b3, const float *b4,
                                   const void *w5, float b5, double scale, double shift,
                                   double *pred, long long p0, int *work, long long work_len,
                                   void *stream) {
  if (e0 < 0 || nrow < 0 || e0 + nrow > P->nent) return fail("fused network rows out of range");
  if (nrow > 0 && (!work || work_len < jdsg_mlp_fused_work_len(nrow)))
    return fail("fused network: workspace missing or shorter than jdsg_mlp_fused_work(nrow)");
  if (lda < 256 || lda % 4) return fail("fused network: lda must be >= 256 and a multiple of 4");
  const void *ptrs[] = {acol, ext, W2, W3, W4, b2, b3, b4, w5};
  for (const void *q : ptrs)
    if (!q || (reinterpret_cast<uintptr_t>(q) & 15)) return fail("fused network: null or unaligned operand");
  jdsg_launch_mlp_fused(e0, nrow, P->ncol, P->ent_off_d, P->per_off_d, periods, P->nmode_d, acol,
                        lda, ext, (const uint16_t *)W2, (const uint16_t *)W3,
                        (const uint16_t *)W4, b2, b3, b4, (const uint16_t *)w5, b5, scale, shift,
                        pred, p0, work, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" long long jdsg_mlp_fused_work(long long nrow) { return jdsg_mlp_fused_work_len(nrow); }

// pred[e0 + r] = exp(out[r*ldo]*scale + shift) for r < n (MLP output, a
// standardised ln c, to flat c)
extern "C" int jdsg_denorm(long long n, const float *out, int ldo, double scale, double shift,
                           double *pred, long long e0, void *stream) {
  jdsg_launch_denorm(n, out, ldo, scale, shift, pred, e0, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int jdsg_plan_enable_surrogate(jdsg_plan *P) {
  if (P->known) return 0;
  if (P->dalloc(&P->known, (size_t)P->nvar * P->nent) || P->dalloc(&P->cen, 2 * (size_t)P->nent) ||
      P->dalloc(&P->sg, 3 * (size_t)P->nvar * P->nent) ||
      P->dalloc(&P->brk, 2 * (size_t)P->nvar * P->nent) || P->dalloc(&P->sfail, 2))
    return 1;
  // per chain of each wave type: its saved state, its own scratch, a job slot
  // in each of two lists (256x256 columns x 60 periods: about 0.7 GB)
  const size_t nch = (size_t)P->ncol * P->nvar;
  if (P->dalloc(&P->susp_saved, 2 * nch * jdsg_chain_state_bytes()) ||
      P->dalloc(&P->susp_cs, 2 * nch * (2 * (size_t)P->kpad + 22)) ||
      P->dalloc(&P->susp_jobs, 4 * nch) || P->dalloc(&P->susp_njobs, 6))
    return 1;
  return 0;
}

extern "C" int jdsg_plan_execute_surrogate(jdsg_plan *P, const float *thk, const float *vp,
                                           const float *vs, const float *rho,
                                           const double *periods, const double *pred,
                                           double half_width, double *c, double *u, int *status,
                                           double *dcda, double *dcdb, double *dcdr, double *dcdh,
                                           double *duda, double *dudb, double *dudr, double *dudh,
                                           void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (!P->known) return fail("call jdsg_plan_enable_surrogate first");
  if (P->tol >= 1.0e-6) return fail("the surrogate path polishes to exact roots: set tol < 1e-6");
  if (P->timing && !P->capturing) HIPCHK(hipEventRecord(P->ev[0], st));
  jdsg_launch_prep(P->ncol, P->ncol_pad, P->layer_off_d, thk, vp, vs, rho, P->iflsph, P->prep,
                   st);
  jdsg_launch_fill_u32(P->sfail, 2, 0u, st);
  jdsg_launch_fill_u32((unsigned *)(P->nevals + 1), 2, 0u, st);
  PolishArgs Q{};
  Q.cm = P->cm;
  Q.ent_off = P->ent_off_d;
  Q.per_off = P->per_off_d;
  Q.periods = periods;
  Q.nmode = P->nmode_d;
  Q.nent = P->nent;
  Q.tol = P->tol;
  Q.known = P->known;
  Q.sg = P->sg;
  Q.brk = P->brk;
  Q.nfail = P->sfail + 1;  // failure events; sfail[0] counts the roots left open at the end
  Q.nevals = P->nevals + 1;
  Q.grid_K = P->uK;
  Q.grid_nR = P->unR;
  Q.centers[0] = pred;
  Q.hw[0] = half_width;
  Q.centers[1] = P->cen;
  Q.centers[2] = P->cen + P->nent;
  // T(1 +- 0.005) centres come from the polished T roots moved along the
  // slope of the neighbouring periods' roots (about 1e-5 off), but the
  // bracket stays wider than the chain's dc step (0.005 km/s): a pair of
  // roots closer than one step is skipped by the chain, and only a bracket
  // that holds both (no sign change) sends such an entry back to it
  Q.hw[1] = Q.hw[2] = 0.5 * half_width;
  const int pw = default_waves();
  auto polish = [&](int v0, int nv) {
    Q.v0 = v0;
    Q.nv = nv;
    for (int w = 0; w < 2; ++w) {
      if (P->nitem[w] == 0) continue;
      Q.nitem = P->nitem[w];
      Q.item_ent = P->uniform ? nullptr : P->items_d[w];
      jdsg_launch_polish(Q, w, pw, st);
    }
  };
  auto order = [&](int v) {
    Q.variant = v;
    Q.center = Q.centers[v];
    Q.half_width = Q.hw[v];
    jdsg_launch_mode_order(Q, P->kmax, st);
  };
  // open T entries next to a kept root of the same mode, continued from it
  // (the T(1 +- 0.005) brackets then start from those roots through
  // pm_centers; continuing those variants as well changed statuses where a
  // T(1 +- 0.005) chain misses a root near cutoff)
  auto cont = [&](int v) {
    for (int w = 0; w < 2; ++w)
      if (P->nitem[w] > 0) jdsg_launch_continue(Q, w, v, JDSG_MAXMODE, P->fallback == 2, st);
  };
  polish(0, 1);
  order(0);
  cont(0);
  order(0);
  if (P->group) {
    jdsg_launch_pm_centers(P->ent_off_d, P->per_off_d, periods, P->ncol, P->nent, P->known, pred,
                           P->cen, P->cen + P->nent, st);
    polish(1, 2);
    order(1);
    order(2);
  }
  jdsg_launch_fill_u32(P->sfail, 1, 0u, st);
  jdsg_launch_count_open(P->known, (long long)P->nvar * P->nent, P->sfail, st);
  HIPCHK(hipGetLastError());
  if (P->timing && !P->capturing) HIPCHK(hipEventRecord(P->ev[1], st));
  // the chains walk in reference order and solve only the roots left open
  P->known_in = P->known;
  int rc = 0;
  {
    jdsg_launch_fill_u32(P->counters, 4, 0u, st);
    jdsg_launch_fill_u32((unsigned *)P->nevals, 2, 0u, st);
    jdsg_launch_fill_u32((unsigned *)P->cstat, (long long)P->nvar * P->nent, 0u, st);
    ChainArgs A{};
    A.corder = P->corder_d;
    A.cm = P->cm;
    A.ent_off = P->ent_off_d;
    A.per_off = P->per_off_d;
    A.periods = periods;
    A.nmode = P->nmode_d;
    A.nvar = P->nvar;
    A.igr = 0;
    A.retry = 1;
    A.nent = P->nent;
    A.kmax_pad = P->kpad;
    A.cflat = P->cflat;
    A.cstat = P->cstat;
    A.ierr_out = P->ierr;
    A.cscr = P->cscr;
    A.counter = P->counters + 0;
    A.nchain = P->nchain;
    A.nevals = P->nevals;
    A.tol = P->tol;
    A.known = P->known;
    A.chain_evals = P->chain_ev;
    A.any_wave[0] = P->nitem[0] > 0;
    A.any_wave[1] = P->nitem[1] > 0;
    // long upward scans (above all the walk to betmx that establishes a
    // mode's cutoff) go to parallel scan jobs after two sequential steps, up
    // to 64 suspensions per chain (measured on 128x128 columns x 40 periods:
    // fallback 375 ms without, 243 ms with 8 after 8 steps, 210 ms with 64
    // after 2; a cycle without jobs costs two near-empty launches)
    const size_t nch = (size_t)P->ncol * P->nvar;
    A.susp.saved = P->susp_saved;
    A.susp.cs = P->susp_cs;
    for (int w = 0; w < 2; ++w)
      for (int l = 0; l < 2; ++l) A.susp.jobs[w][l] = P->susp_jobs + (size_t)(2 * w + l) * nch;
    A.susp.njobs = P->susp_njobs;
    A.susp.total = P->susp_njobs + 4;
    jdsg_launch_fill_u32(A.susp.total, 2, 0u, st);
    A.susp.min_steps = 2;
    jdsg_launch_chains_susp(A, P->chain_waves, P->fallback == 0 ? 0 : 64, P->fallback == 2, st);
    rc = hipGetLastError() != hipSuccess;
  }
  P->known_in = nullptr;
  if (rc) return fail("chain launch failed");
  if (P->timing && !P->capturing) HIPCHK(hipEventRecord(P->ev[2], st));
  double *dc[4] = {dcda, dcdb, dcdr, dcdh};
  double *du[4] = {duda, dudb, dudr, dudh};
  if (run_eigen(P, periods, c, u, status, dc, du, st)) return 1;
  if (P->timing && !P->capturing) {
    if (P->finish_timing()) return 1;
  }
  return 0;
}

// ---- the whole surrogate iteration in device work only (network + solve)
extern "C" int jdsg_plan_set_network(jdsg_plan *P, int hid, const void *const W[5],
                                     const float *const b[5], double scale, double shift) {
  if (hid <= 0 || hid % 128) return fail("network width must be a positive multiple of 128");
  for (int i = 0; i < 5; ++i)
    if (!W[i] || !b[i] || (reinterpret_cast<uintptr_t>(W[i]) & 15) ||
        (reinterpret_cast<uintptr_t>(b[i]) & 15))
      return fail("network: null or unaligned weight or bias");
  if (jdsg_plan_enable_surrogate(P)) return 1;
  auto N = std::make_unique<jdsg_plan::Net>();
  N->hid = hid;
  for (int i = 0; i < 5; ++i) {
    N->W[i] = (const uint16_t *)W[i];
    N->b[i] = b[i];
  }
  N->scale = scale;
  N->shift = shift;
  // the first layer's weight columns of the per-entry inputs (ln T, two wave
  // and four mode one-hots) and its bias, as jdsg_plan_mlp_expand takes them
  std::vector<uint16_t> w0((size_t)hid * 192);
  std::vector<float> b0(hid), ext((size_t)8 * hid);
  HIPCHK(hipMemcpy(w0.data(), W[0], sizeof(uint16_t) * w0.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(b0.data(), b[0], sizeof(float) * hid, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&N->b5, b[4], sizeof(float), hipMemcpyDeviceToHost));
  for (int r = 0; r < 7; ++r)
    for (int h = 0; h < hid; ++h) {
      uint32_t bits = (uint32_t)w0[(size_t)h * 192 + 144 + r] << 16;
      float f;
      memcpy(&f, &bits, sizeof(f));
      ext[(size_t)r * hid + h] = f;
    }
  for (int h = 0; h < hid; ++h) ext[(size_t)7 * hid + h] = b0[h];
  size_t ncp = (size_t)(P->ncol + 127) / 128 * 128;
  if (P->dalloc(&N->ext, ext.size()) || P->dalloc(&N->colf, (size_t)P->ncol * 148) ||
      P->dalloc(&N->Xc, ncp * 192) || P->dalloc(&N->acol, ncp * hid) ||
      P->dalloc(&N->pred, P->nent))
    return 1;
  HIPCHK(hipMemcpy(N->ext, ext.data(), sizeof(float) * ext.size(), hipMemcpyHostToDevice));
  if (hid == 256) {
    N->work_len = jdsg_mlp_fused_work_len(P->nent);
    if (P->dalloc(&N->work, std::max<long long>(N->work_len, 1))) return 1;
  } else {
    N->chunk = std::min<long long>(1 << 20, (P->nent + 63) / 64 * 64);
    for (int q = 0; q < 2; ++q)
      if (P->dalloc(&N->Y[q], (size_t)std::max<long long>(N->chunk, 64) * hid)) return 1;
  }
  P->net = std::move(N);
  return 0;
}

namespace {
// predicted flat c of every entry into net->pred
int run_network(jdsg_plan *P, const float *thk, const float *vp, const float *vs,
                const float *rho, const double *periods, hipStream_t st) {
  jdsg_plan::Net &N = *P->net;
  const int H = N.hid;
  const int ncp = (P->ncol + 127) / 128 * 128;
  jdsg_launch_column_features(P->ncol, P->layer_off_d, thk, vp, vs, rho, N.colf, st);
  if (jdsg_mlp_column_rows(P->ncol, N.colf, N.Xc, st)) return 1;
  if (jdsg_gemm_bf16(0, 1, ncp, H, 192, N.Xc, 192, N.W[0], 192, N.acol, H, 0, nullptr, nullptr, 0,
                     st))
    return 1;
  if (H == 256)
    return jdsg_plan_mlp_fused(P, periods, 0, P->nent, N.acol, H, N.ext, N.W[1], N.W[2], N.W[3],
                               N.b[1], N.b[2], N.b[3], N.W[4], N.b5, N.scale, N.shift, N.pred, 0,
                               N.work, N.work_len, st);
  for (long long s0 = 0; s0 < P->nent; s0 += N.chunk) {
    long long m = std::min(N.chunk, P->nent - s0);
    int mp = (int)((m + 63) / 64 * 64);
    if (jdsg_plan_mlp_expand(P, periods, s0, m, N.acol, H, N.ext, H, N.Y[0], st)) return 1;
    int a = 0;
    for (int li = 1; li < 4; ++li, a ^= 1)
      if (jdsg_gemm_bf16(0, 1, mp, H, H, N.Y[a], H, N.W[li], H, nullptr, H, 4, N.b[li],
                         N.Y[a ^ 1], H, st))
        return 1;
    if (jdsg_mlp_head(m, H, N.Y[a], N.W[4], N.b5, N.scale, N.shift, N.pred, s0, st)) return 1;
  }
  return 0;
}
}  // namespace

extern "C" int jdsg_plan_execute_surrogate_net(jdsg_plan *P, const float *thk, const float *vp,
                                               const float *vs, const float *rho,
                                               const double *periods, double half_width,
                                               double *c, double *u, int *status, double *dcda,
                                               double *dcdb, double *dcdr, double *dcdh,
                                               double *duda, double *dudb, double *dudr,
                                               double *dudh, void *stream) {
  if (!P->net) return fail("call jdsg_plan_set_network first");
  hipStream_t st = (hipStream_t)stream;
  if (run_network(P, thk, vp, vs, rho, periods, st)) return 1;
  return jdsg_plan_execute_surrogate(P, thk, vp, vs, rho, periods, P->net->pred, half_width, c,
                                     u, status, dcda, dcdb, dcdr, dcdh, duda, dudb, dudr, dudh,
                                     stream);
}

extern "C" int jdsg_plan_network_pred(const jdsg_plan *P, double *host) {
  if (!P->net) return fail("no network bound");
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(host, P->net->pred, sizeof(double) * P->nent, hipMemcpyDeviceToHost));
  return 0;
}

// Diagnostics of the last surrogate execute: the polished roots that the
// checks kept (0 where the chain took over), nvar*nent doubles, and the signs
// of the period equation at bracket low, high and the chain floor, 3*nvar*nent
// bytes (either pointer may be null).
extern "C" int jdsg_plan_surrogate_roots(const jdsg_plan *P, double *known, signed char *signs) {
  if (!P->known) return fail("surrogate path not enabled");
  HIPCHK(hipDeviceSynchronize());
  size_t n = (size_t)P->nvar * P->nent;
  if (known) HIPCHK(hipMemcpy(known, P->known, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (signs) HIPCHK(hipMemcpy(signs, P->sg, 3 * n, hipMemcpyDeviceToHost));
  return 0;
}

// Roots left to the plain chain by the last surrogate execute, and the
// period-equation evaluations of the bracket/polish and chain stages.
extern "C" int jdsg_plan_surrogate_stats(const jdsg_plan *P, long long *nfail,
                                         long long *polish_evals, long long *chain_evals) {
  unsigned f[2] = {0, 0};
  unsigned long long ev[2] = {0, 0};
  if (P->sfail) HIPCHK(hipMemcpy(f, P->sfail, sizeof(f), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ev, P->nevals, sizeof(ev), hipMemcpyDeviceToHost));
  if (nfail) *nfail = f[0];
  if (chain_evals) *chain_evals = (long long)ev[0];
  if (polish_evals) *polish_evals = (long long)ev[1];
  return 0;
}

extern "C" int jdsg_plan_scan_jobs(const jdsg_plan *P, long long *jobs) {
  unsigned t[2] = {0, 0};
  if (P->susp_njobs) HIPCHK(hipMemcpy(t, P->susp_njobs + 4, sizeof(t), hipMemcpyDeviceToHost));
  if (jobs) *jobs = (long long)t[0] + t[1];
  return 0;
}

// ============================================================================
// Iterated batch: launch widths picked by measurement, and the per-iteration
// model update + solve captured once in a hipGraph and replayed.
void jdsg_launch_perturb(long long n, float *vs, float *vp, float *rho, const int *iter,
                         float amp, uint32_t seed, hipStream_t st);
void jdsg_launch_bump(int *iter, hipStream_t st);

namespace {
// milliseconds of fn() on stream st between two events
template <class Fn>
int time_ms(hipStream_t st, Fn fn, float *ms) {
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, st));
  int rc = fn();
  HIPCHK(hipEventRecord(b, st));
  HIPCHK(hipEventSynchronize(b));
  HIPCHK(hipEventElapsedTime(ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}
}  // namespace

// Autotuner: times the chain stage over candidate launch widths (workgroups
// of one wave; multiples of the CU count up to the chains available), keeps
// the fastest, then does the same for the eigen stage on the roots just
// found, and sizes the scratch for the winners.  The outputs are written as
// by jdsg_plan_execute (the last candidate's solve).  The chosen widths are
// returned through chain_waves / eigen_waves (may be null).
extern "C" int jdsg_plan_autotune(jdsg_plan *P, const float *thk, const float *vp,
                                  const float *vs, const float *rho, const double *periods,
                                  double *c, double *u, int *status, double *dcda, double *dcdb,
                                  double *dcdr, double *dcdh, double *duda, double *dudb,
                                  double *dudr, double *dudh, void *stream, int *chain_waves,
                                  int *eigen_waves) {
  hipStream_t st = (hipStream_t)stream;
  double *dc[4] = {dcda, dcdb, dcdr, dcdh};
  double *du[4] = {duda, dudb, dudr, dudh};
  const int ncu = std::max(1, default_waves() / 16);
  const long long max_cw = std::max(1LL, ((long long)P->ncol * P->nvar + 63) / 64);
  const long long max_ew =
      std::max(1LL, (std::max(P->nitem[0], P->nitem[1]) + 63) / 64);
  bool was_timing = P->timing;
  P->timing = false;
  auto run = [&]() {
    if (run_front(P, thk, vp, vs, rho, periods, 0, 1, st)) return 1;
    return run_eigen(P, periods, c, u, status, dc, du, st);
  };
  // warm-up at the current widths
  float ms = 0.f;
  int rc = time_ms(st, run, &ms);
  int best_cw = P->chain_waves, best_ew = P->eigen_waves;
  float best = 3.0e38f;
  const int cmul[] = {4, 8, 12, 16, 24};
  for (int m : cmul) {
    if (rc) break;
    int cw = (int)std::min<long long>((long long)m * ncu, max_cw);
    P->chain_waves = cw;
    if (P->size_scratch()) { rc = 1; break; }
    rc = time_ms(st, [&]() { return run_front(P, thk, vp, vs, rho, periods, 0, 1, st); }, &ms);
    if (!rc && ms < best) { best = ms; best_cw = cw; }
    if (cw == max_cw) break;
  }
  best = 3.0e38f;
  const int emul[] = {2, 4, 8, 16, 32};
  for (int m : emul) {
    if (rc) break;
    int ew = (int)std::min<long long>((long long)m * ncu, max_ew);
    P->eigen_waves = ew;
    if (P->size_scratch()) { rc = 1; break; }
    rc = time_ms(st, [&]() { return run_eigen(P, periods, c, u, status, dc, du, st); }, &ms);
    if (!rc && ms < best) { best = ms; best_ew = ew; }
    if (ew == max_ew) break;
  }
  P->chain_waves = best_cw;
  P->eigen_waves = best_ew;
  P->timing = was_timing;
  if (P->size_scratch()) return 1;
  if (chain_waves) *chain_waves = best_cw;
  if (eigen_waves) *eigen_waves = best_ew;
  return rc;
}

struct jdsg_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int *iter = nullptr;
  ~jdsg_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (iter) (void)hipFree(iter);
  }
};

// Captures one iteration -- optionally the model update (Vs *= 1 + amp*u,
// u uniform in [-1, 1) seeded by `seed` and the replay count; Vp and rho by
// Brocher's relations) followed by the whole solve of jdsg_plan_execute --
// on `stream` (which must not be the null stream) into a graph.  Each
// jdsg_graph_launch then runs one iteration with a single launch; the model
// buffers are updated in place.  The plan and the buffers must outlive the
// graph.
extern "C" int jdsg_plan_capture(jdsg_plan *P, float *thk, float *vp, float *vs, float *rho,
                                 const double *periods, double *c, double *u, int *status,
                                 double *dcda, double *dcdb, double *dcdr, double *dcdh,
                                 double *duda, double *dudb, double *dudr, double *dudh,
                                 float perturb_amp, unsigned seed, void *stream,
                                 jdsg_graph **out) {
  hipStream_t st = (hipStream_t)stream;
  if (!st) return fail("graph capture needs a non-null stream");
  if (!thk || !vp || !vs || !rho || !periods || !c || !u || !status)
    return fail("null input or scalar output pointer");
  auto *G = new jdsg_graph();
  if (hipMalloc(&G->iter, sizeof(int)) != hipSuccess ||
      hipMemset(G->iter, 0, sizeof(int)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    delete G;
    return fail("graph counter allocation failed");
  }
  double *dc[4] = {dcda, dcdb, dcdr, dcdh};
  double *du[4] = {duda, dudb, dudr, dudh};
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    delete G;
    return fail(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  }
  P->capturing = true;
  int rc = 0;
  long long nlay = P->nlay_total;
  if (perturb_amp > 0.f) jdsg_launch_perturb(nlay, vs, vp, rho, G->iter, perturb_amp, seed, st);
  rc = run_front(P, thk, vp, vs, rho, periods, 0, 1, st);
  if (!rc) rc = run_eigen(P, periods, c, u, status, dc, du, st);
  if (perturb_amp > 0.f) jdsg_launch_bump(G->iter, st);
  P->capturing = false;
  e = hipStreamEndCapture(st, &G->graph);
  if (rc || e != hipSuccess) {
    delete G;
    return rc ? 1 : fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    delete G;
    return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  *out = G;
  return 0;
}

// The surrogate iteration (model update, network, bracket + polish, fallback
// chains, eigen stage) captured like jdsg_plan_capture; the plan needs a
// bound network (jdsg_plan_set_network) and exact-root mode.
extern "C" int jdsg_plan_capture_surrogate(jdsg_plan *P, float *thk, float *vp, float *vs,
                                           float *rho, const double *periods, double half_width,
                                           double *c, double *u, int *status, double *dcda,
                                           double *dcdb, double *dcdr, double *dcdh,
                                           double *duda, double *dudb, double *dudr,
                                           double *dudh, float perturb_amp, unsigned seed,
                                           void *stream, jdsg_graph **out) {
  hipStream_t st = (hipStream_t)stream;
  if (!st) return fail("graph capture needs a non-null stream");
  if (!P->net) return fail("call jdsg_plan_set_network first");
  if (P->tol >= 1.0e-6) return fail("the surrogate path polishes to exact roots: set tol < 1e-6");
  auto G = std::make_unique<jdsg_graph>();
  if (hipMalloc(&G->iter, sizeof(int)) != hipSuccess ||
      hipMemset(G->iter, 0, sizeof(int)) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return fail("graph counter allocation failed");
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fail(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  P->capturing = true;
  long long nlay = P->nlay_total;
  if (perturb_amp > 0.f) jdsg_launch_perturb(nlay, vs, vp, rho, G->iter, perturb_amp, seed, st);
  int rc = jdsg_plan_execute_surrogate_net(P, thk, vp, vs, rho, periods, half_width, c, u, status,
                                           dcda, dcdb, dcdr, dcdh, duda, dudb, dudr, dudh, st);
  if (perturb_amp > 0.f) jdsg_launch_bump(G->iter, st);
  P->capturing = false;
  e = hipStreamEndCapture(st, &G->graph);
  if (rc) return 1;
  if (e != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  *out = G.release();
  return 0;
}

extern "C" int jdsg_graph_launch(jdsg_graph *G, void *stream) {
  HIPCHK(hipGraphLaunch(G->exec, (hipStream_t)stream));
  return 0;
}

extern "C" void jdsg_graph_destroy(jdsg_graph *G) { delete G; }

// ============================================================================
// SurfTime grid entry points: get_1d_kernel (SurfaceWave.cpp:458-502 over
// kernel1D :266-412) and get_2d_map (:420-450 over disp1D :177-256).
//
// Columns are split by their water top (vs(i,j,0) == 0, decided per column
// as grid2LayerModel and _ParamConvert do, :120, :71, :495) into two
// classes with 4nz-3 and 4nz-6 layers.  Each class runs in chunks of
// columns through uniform layer plans, one plan per job (wave type, period
// list, mode count, form); an epilogue writes SurfTime's column-major
// outputs through the class's column list.
struct jdsg_surftime {
  int nx = 0, ny = 0, nz = 0, nxy = 0, kmax = 0;
  std::vector<double> tab[4];
  int ntm[4] = {0, 0, 0, 0}, nmd[4] = {0, 0, 0, 0};
  struct Job {
    int wave;   // 0 Rayleigh, 1 Love
    int group;  // kernel form: Rg/Lg (three chains); map form: eigen stage for U
    int vp2;    // map form: the chain runs on Vp = 1.732 Vs (groupvel_l)
    int M;
    std::vector<double> list;
  };
  struct Set {
    int cls = 0, ncol = 0, N = 0;
    std::vector<jdsg_plan *> plans;
    std::vector<void *> allocs;
    float *thk = nullptr, *vp = nullptr, *vs = nullptr, *rho = nullptr, *vp2 = nullptr;
    std::vector<double *> periods;
    std::vector<StJob> hjobs;
    std::vector<double *> out_c, out_u;
    std::vector<int *> out_st;
    std::vector<std::array<double *, 6>> out_k;
    StJob *djobs = nullptr;
    ~Set() {
      for (auto *p : plans) jdsg_plan_destroy(p);
      for (void *q : allocs) (void)hipFree(q);
    }
    template <class T>
    int dalloc(T **p, size_t n) {
      void *q = nullptr;
      if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
        return fail("hipMalloc (surftime)");
      allocs.push_back(q);
      *p = (T *)q;
      return 0;
    }
  };
  struct Mode {
    int map = 0;
    std::vector<Job> jobs;
    std::vector<StEntry> ents;
    std::vector<StMapEntry> ments;
    void *ents_d = nullptr;
    std::vector<std::unique_ptr<Set>> sets;
    bool built = false;
  };
  Mode modes[2];  // 0: get_1d_kernel, 1: get_2d_map
  std::vector<void *> allocs;
  float *dep_d = nullptr, *node_vp = nullptr, *node_vs = nullptr;
  int *mask_d = nullptr, *cols_d = nullptr;
  double *tlist_d = nullptr;
  std::vector<int> mask_h;
  int ncls[2] = {0, 0};
  bool have_mask = false;
  double budget = 0.0;  // device bytes for the chunks' intermediate results
  ~jdsg_surftime() {
    modes[0].sets.clear();
    modes[1].sets.clear();
    for (void *q : allocs) (void)hipFree(q);
  }
  template <class T>
  int dalloc(T **p, size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
      return fail("hipMalloc (surftime)");
    allocs.push_back(q);
    *p = (T *)q;
    return 0;
  }
  int layers(int cls) const { return cls ? 4 * nz - 6 : 4 * nz - 3; }
  int ntrow(int w, int m) const {  // SurfaceWave.cpp:157-168
    const double *col = tab[w].data() + (size_t)m * ntm[w];
    int n = 0;
    while (n < ntm[w] && col[n] != -1) ++n;
    return n;
  }
  std::vector<double> list(int w, int m) const {
    const double *col = tab[w].data() + (size_t)m * ntm[w];
    return std::vector<double>(col, col + ntrow(w, m));
  }
};

namespace {

// Jobs and entries of one form.  Output index order is get_period_index's
// (SurfaceWave.cpp:14-48): Rc modes, Rg modes, Lc modes, Lg modes.
int build_mode(jdsg_surftime *S, jdsg_surftime::Mode &Md) {
  auto &jobs = Md.jobs;
  // a job serves a request when wave, list and Vp form agree and it runs at
  // least the stages needed
  auto find_job = [&](int wave, int group, int vp2, const std::vector<double> &L) {
    for (size_t j = 0; j < jobs.size(); ++j)
      if (jobs[j].wave == wave && jobs[j].list == L && jobs[j].vp2 == vp2 &&
          (jobs[j].group || !group))
        return (int)j;
    jobs.push_back({wave, group, vp2, 0, L});
    return (int)jobs.size() - 1;
  };
  // job per (table, mode, part): part 0 gives the value / root, part 1 U
  std::vector<std::vector<std::array<int, 2>>> jobof(4);
  for (int w : {1, 3, 0, 2}) {  // group tables first so phase tables reuse them
    int wave = w >= 2 ? 1 : 0, grp = (w == 1 || w == 3) ? 1 : 0;
    jobof[w].assign(std::max(S->nmd[w], 0), {-1, -1});
    for (int m = 0; m < S->nmd[w]; ++m) {
      std::vector<double> L = S->list(w, m);
      if (L.empty()) continue;
      std::array<int, 2> jj{-1, -1};
      if (!Md.map) {
        jj[0] = find_job(wave, grp, 0, L);
      } else if (grp && wave == 0) {  // Rg: one chain gives vc and, with sregn96, U
        jj[1] = find_job(0, 1, 0, L);
        jj[0] = jj[1];
      } else if (grp) {  // Lg: vc from the chain on the model's Vp, U from Vp = 1.732 Vs
        jj[1] = find_job(1, 1, 1, L);
        jj[0] = find_job(1, 0, 0, L);
      } else {
        jj[0] = find_job(wave, 0, 0, L);
      }
      for (int p : jj)
        if (p >= 0) {
          jobs[p].M = std::max(jobs[p].M, m + 1);
          if (jobs[p].M > JDSG_MAXMODE) return fail("at most 8 modes per table");
        }
      jobof[w][m] = jj;
    }
  }
  std::vector<double> tl;
  for (int w = 0; w < 4; ++w)
    for (int m = 0; m < S->nmd[w]; ++m) {
      int nt = S->ntrow(w, m);
      int grp = (w == 1 || w == 3) ? 1 : 0, love = w >= 2 ? 1 : 0;
      for (int k = 0; k < nt; ++k) {
        if (!Md.map)
          Md.ents.push_back({jobof[w][m][0], m, k, grp, love});
        else
          Md.ments.push_back({jobof[w][m][0], jobof[w][m][1], m, k, love});
        tl.push_back(S->tab[w][(size_t)m * S->ntm[w] + k]);
      }
    }
  size_t ne = Md.map ? Md.ments.size() : Md.ents.size();
  if ((int)ne != S->kmax) return fail("surftime: inconsistent period index");
  if (Md.map) {
    StMapEntry *d = nullptr;
    if (S->dalloc(&d, ne)) return 1;
    if (ne && hipMemcpy(d, Md.ments.data(), sizeof(StMapEntry) * ne, hipMemcpyHostToDevice) !=
                  hipSuccess)
      return fail("surftime upload");
    Md.ents_d = d;
    if (!S->tlist_d) {
      if (S->dalloc(&S->tlist_d, ne)) return 1;
      if (ne && hipMemcpy(S->tlist_d, tl.data(), sizeof(double) * ne, hipMemcpyHostToDevice) !=
                    hipSuccess)
        return fail("surftime upload");
    }
  } else {
    StEntry *d = nullptr;
    if (S->dalloc(&d, ne)) return 1;
    if (ne && hipMemcpy(d, Md.ents.data(), sizeof(StEntry) * ne, hipMemcpyHostToDevice) !=
                  hipSuccess)
      return fail("surftime upload");
    Md.ents_d = d;
  }
  Md.built = true;
  return 0;
}

// Device bytes of one column's intermediate results in a form.
double column_bytes(const jdsg_surftime *S, const jdsg_surftime::Mode &Md, int cls) {
  double b = 16.0 * S->layers(cls) + (Md.map ? 4.0 * S->layers(cls) : 0.0);
  for (auto &J : Md.jobs) {
    double ne = (double)J.M * J.list.size();
    b += ne * 20.0 + (J.group ? 3.0 : 1.0) * ne * 12.0;
    if (!Md.map) b += ne * S->layers(cls) * 8.0 * (J.group ? 6 : 3);
  }
  return b;
}

int build_set(jdsg_surftime *S, jdsg_surftime::Mode &Md, jdsg_surftime::Set &T, int cls,
              int ncol) {
  T.cls = cls;
  T.ncol = ncol;
  T.N = S->layers(cls);
  const int N = T.N;
  size_t nl = (size_t)ncol * N;
  if (T.dalloc(&T.thk, nl) || T.dalloc(&T.vp, nl) || T.dalloc(&T.vs, nl) || T.dalloc(&T.rho, nl))
    return 1;
  bool need_vp2 = false;
  for (auto &J : Md.jobs) need_vp2 |= J.vp2 != 0;
  if (need_vp2 && T.dalloc(&T.vp2, nl)) return 1;
  std::vector<int> lo(ncol + 1);
  for (int c = 0; c <= ncol; ++c) lo[c] = c * N;
  for (auto &J : Md.jobs) {
    int K = (int)J.list.size();
    std::vector<int> po(ncol + 1), nm(2 * (size_t)ncol);
    for (int c = 0; c <= ncol; ++c) po[c] = c * K;
    for (int c = 0; c < ncol; ++c) {
      nm[2 * c] = J.wave == 0 ? J.M : 0;
      nm[2 * c + 1] = J.wave == 1 ? J.M : 0;
    }
    jdsg_plan *P = nullptr;
    int pg = Md.map ? 0 : J.group;
    if (jdsg_plan_create(ncol, lo.data(), po.data(), nm.data(), pg, 0, 1, &P)) return 1;
    T.plans.push_back(P);
    std::vector<double> per((size_t)ncol * K);
    for (int c = 0; c < ncol; ++c)
      std::copy(J.list.begin(), J.list.end(), per.begin() + (size_t)c * K);
    double *pd = nullptr;
    if (T.dalloc(&pd, per.size())) return 1;
    if (hipMemcpy(pd, per.data(), sizeof(double) * per.size(), hipMemcpyHostToDevice) !=
        hipSuccess)
      return fail("surftime period upload");
    T.periods.push_back(pd);
    double *c = nullptr, *u = nullptr;
    int *st = nullptr;
    std::array<double *, 6> k{};
    bool eig = !Md.map || J.group;
    if (eig && (T.dalloc(&c, P->nent) || T.dalloc(&u, P->nent) || T.dalloc(&st, P->nent)))
      return 1;
    int nk = Md.map ? 0 : (J.group ? 6 : 3);
    for (int a = 0; a < nk; ++a)
      if (T.dalloc(&k[a], P->nkern)) return 1;
    T.out_c.push_back(c);
    T.out_u.push_back(u);
    T.out_st.push_back(st);
    T.out_k.push_back(k);
    StJob h{};
    h.c = c;
    h.u = u;
    h.status = st;
    for (int a = 0; a < 3; ++a) {
      h.dc[a] = k[a];
      h.du[a] = (!Md.map && J.group) ? k[3 + a] : nullptr;
    }
    h.cflat = P->cflat;
    h.cstat = P->cstat;
    h.K = K;
    h.nR = J.wave == 0 ? J.M : 0;
    h.nL = J.wave == 1 ? J.M : 0;
    T.hjobs.push_back(h);
  }
  if (T.dalloc(&T.djobs, std::max<size_t>(T.hjobs.size(), 1))) return 1;
  if (!T.hjobs.empty() &&
      hipMemcpy(T.djobs, T.hjobs.data(), sizeof(StJob) * T.hjobs.size(),
                hipMemcpyHostToDevice) != hipSuccess)
    return fail("surftime job upload");
  return 0;
}

jdsg_surftime::Set *get_set(jdsg_surftime *S, jdsg_surftime::Mode &Md, int cls, int ncol) {
  for (auto &p : Md.sets)
    if (p->cls == cls && p->ncol == ncol) return p.get();
  auto T = std::make_unique<jdsg_surftime::Set>();
  if (build_set(S, Md, *T, cls, ncol)) return nullptr;
  Md.sets.push_back(std::move(T));
  return Md.sets.back().get();
}

// Reads the water-top mask of vs (synchronises st) and, when it changed,
// rebuilds the two column lists (land columns first, then water columns).
int update_mask(jdsg_surftime *S, const float *vs, hipStream_t st) {
  jdsg_launch_water_mask(S->nxy, vs, S->mask_d, st);
  HIPCHK(hipGetLastError());
  std::vector<int> m(S->nxy);
  HIPCHK(hipMemcpyAsync(m.data(), S->mask_d, sizeof(int) * S->nxy, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (S->have_mask && m == S->mask_h) return 0;
  std::vector<int> cols;
  cols.reserve(S->nxy);
  for (int n = 0; n < S->nxy; ++n)
    if (!m[n]) cols.push_back(n);
  S->ncls[0] = (int)cols.size();
  for (int n = 0; n < S->nxy; ++n)
    if (m[n]) cols.push_back(n);
  S->ncls[1] = S->nxy - S->ncls[0];
  if (S->ncls[1] && S->nz < 3) return fail("a water top needs nz >= 3");
  HIPCHK(hipMemcpy(S->cols_d, cols.data(), sizeof(int) * S->nxy, hipMemcpyHostToDevice));
  S->mask_h.swap(m);
  S->have_mask = true;
  for (auto &Md : S->modes) Md.sets.clear();  // chunk sizes follow the class sizes
  return 0;
}

int surftime_run(jdsg_surftime *S, int mapform, const float *vs, double *vc, double *vout,
                 double *kernel, int *status, hipStream_t st) {
  if (S->kmax == 0) return 0;
  auto &Md = S->modes[mapform];
  if (!Md.built && build_mode(S, Md)) return 1;
  if (update_mask(S, vs, st)) return 1;
  if (S->budget <= 0.0) {
    // intermediate layer kernels of all chunks together: most of the free
    // device memory (a whole 256x256-column grid fits one chunk on an MI355X), or
    // JDSG_SURFTIME_BUDGET_GB
    size_t fr = 0, tot = 0;
    S->budget = (hipMemGetInfo(&fr, &tot) == hipSuccess) ? 0.6 * (double)fr : 12e9;
    if (const char *b = getenv("JDSG_SURFTIME_BUDGET_GB")) S->budget = atof(b) * 1e9;
  }
  const double budget = S->budget;
  for (int cls = 0; cls < 2; ++cls) {
    int nc = S->ncls[cls];
    if (nc == 0) continue;
    const int *cols = S->cols_d + (cls ? S->ncls[0] : 0);
    double share = budget * nc / S->nxy;
    int chunk = (int)std::max(1.0, std::min<double>(nc, share / column_bytes(S, Md, cls)));
    for (int col0 = 0; col0 < nc; col0 += chunk) {
      int nch = std::min(chunk, nc - col0);
      jdsg_surftime::Set *T = get_set(S, Md, cls, nch);
      if (!T) return 1;
      jdsg_launch_grid_to_layers(col0, nch, S->nxy, S->nz, T->N, cls, cols, vs, S->dep_d, T->thk,
                                 T->vp, T->vs, T->rho, T->vp2, S->node_vp, S->node_vs, st);
      for (size_t j = 0; j < Md.jobs.size(); ++j) {
        auto &J = Md.jobs[j];
        jdsg_plan *P = T->plans[j];
        const float *vpj = J.vp2 ? T->vp2 : T->vp;
        if (!mapform) {
          auto &k = T->out_k[j];
          bool g = J.group;
          if (jdsg_plan_execute(P, T->thk, vpj, T->vs, T->rho, T->periods[j], T->out_c[j],
                                T->out_u[j], T->out_st[j], k[1], k[0], k[2], nullptr,
                                g ? k[4] : nullptr, g ? k[3] : nullptr, g ? k[5] : nullptr,
                                nullptr, st))
            return 1;
        } else if (J.group) {
          if (jdsg_plan_execute(P, T->thk, vpj, T->vs, T->rho, T->periods[j], T->out_c[j],
                                T->out_u[j], T->out_st[j], nullptr, nullptr, nullptr, nullptr,
                                nullptr, nullptr, nullptr, nullptr, st))
            return 1;
        } else if (run_front(P, T->thk, vpj, T->vs, T->rho, T->periods[j], 0, 1, st)) {
          return 1;
        }
      }
      if (!mapform)
        jdsg_launch_surftime_epilogue(col0, nch, S->nxy, S->nz, T->N, cls, S->kmax, cols,
                                      Md.ents_d, T->djobs, S->node_vp, S->node_vs, vc, vout,
                                      kernel, status, kernel != nullptr, st);
      else
        jdsg_launch_surftime_map_epilogue(col0, nch, S->nxy, S->kmax, cols, Md.ents_d, T->djobs,
                                          S->tlist_d, vc, vout, status, st);
      HIPCHK(hipGetLastError());
    }
  }
  return 0;
}

}  // namespace

extern "C" int jdsg_surftime_create(int nx, int ny, int nz, const float *dep, const double *tRc,
                                    int ntRc, int nmRc, const double *tRg, int ntRg, int nmRg,
                                    const double *tLc, int ntLc, int nmLc, const double *tLg,
                                    int ntLg, int nmLg, jdsg_surftime **out) {
  *out = nullptr;
  if (nx <= 0 || ny <= 0 || nz < 2) return fail("bad grid dimensions");
  auto S = std::make_unique<jdsg_surftime>();
  S->nx = nx;
  S->ny = ny;
  S->nz = nz;
  S->nxy = nx * ny;
  const double *tab[4] = {tRc, tRg, tLc, tLg};
  int ntm[4] = {ntRc, ntRg, ntLc, ntLg}, nmd[4] = {nmRc, nmRg, nmLc, nmLg};
  for (int w = 0; w < 4; ++w) {
    if (!tab[w] || nmd[w] <= 0 || ntm[w] <= 0) continue;
    S->ntm[w] = ntm[w];
    S->nmd[w] = nmd[w];
    S->tab[w].assign(tab[w], tab[w] + (size_t)ntm[w] * nmd[w]);
  }
  S->kmax = 0;
  for (int w = 0; w < 4; ++w)
    for (int m = 0; m < S->nmd[w]; ++m) S->kmax += S->ntrow(w, m);
  S->modes[1].map = 1;
  if (S->dalloc(&S->dep_d, nz) || S->dalloc(&S->node_vp, (size_t)nz * S->nxy) ||
      S->dalloc(&S->node_vs, (size_t)nz * S->nxy) || S->dalloc(&S->mask_d, S->nxy) ||
      S->dalloc(&S->cols_d, S->nxy))
    return 1;
  if (hipMemcpy(S->dep_d, dep, sizeof(float) * nz, hipMemcpyHostToDevice) != hipSuccess)
    return fail("surftime upload");
  if (build_mode(S.get(), S->modes[0])) return 1;
  *out = S.release();
  return 0;
}

extern "C" int jdsg_surftime_kmax(const jdsg_surftime *S, int *kmax) {
  *kmax = S->kmax;
  return 0;
}

extern "C" int jdsg_surftime_dims(const jdsg_surftime *S, int *nx, int *ny, int *nz, int *kmax) {
  if (!S) return fail("surftime_dims: null handle");
  if (nx) *nx = S->nx;
  if (ny) *ny = S->ny;
  if (nz) *nz = S->nz;
  if (kmax) *kmax = S->kmax;
  return 0;
}

extern "C" void jdsg_surftime_destroy(jdsg_surftime *S) { delete S; }

extern "C" int jdsg_surftime_kernel(jdsg_surftime *S, const float *vs, double *vc, double *vout,
                                    double *kernel, int *status, void *stream) {
  return surftime_run(S, 0, vs, vc, vout, kernel, status, (hipStream_t)stream);
}

extern "C" int jdsg_surftime_map(jdsg_surftime *S, const float *vs, double *vc, double *vout,
                                 int *status, void *stream) {
  return surftime_run(S, 1, vs, vc, vout, nullptr, status, (hipStream_t)stream);
}

namespace {
int host_grid_call(int mapform, int nx, int ny, int nz, const float *vs, const float *dep,
                   const double *tRc, int ntRc, int nmRc, const double *tRg, int ntRg, int nmRg,
                   const double *tLc, int ntLc, int nmLc, const double *tLg, int ntLg, int nmLg,
                   double *vc, double *vout, double *kernel, int *status) {
  int nxy = nx * ny;
  jdsg_surftime *S = nullptr;
  if (jdsg_surftime_create(nx, ny, nz, dep, tRc, ntRc, nmRc, tRg, ntRg, nmRg, tLc, ntLc, nmLc,
                           tLg, ntLg, nmLg, &S))
    return 1;
  std::unique_ptr<jdsg_surftime> own(S);
  DevBuf B;
  size_t nk = (size_t)nxy * S->kmax;
  float *dvs = B.up(vs, (size_t)nxy * nz);
  double *dvc = B.up<double>(nullptr, nk), *dvo = B.up<double>(nullptr, nk);
  double *dker = (kernel && !mapform) ? B.up<double>(nullptr, nk * nz) : nullptr;
  int *dst = B.up<int>(nullptr, nk);
  if (!dvs || !dvc || !dvo || !dst || (kernel && !mapform && !dker))
    return fail("device allocation");
  if (dker) HIPCHK(hipMemset(dker, 0, sizeof(double) * nk * nz));
  if (surftime_run(S, mapform, dvs, dvc, dvo, dker, dst, nullptr)) return 1;
  HIPCHK(hipMemcpy(vc, dvc, sizeof(double) * nk, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(vout, dvo, sizeof(double) * nk, hipMemcpyDeviceToHost));
  if (dker) HIPCHK(hipMemcpy(kernel, dker, sizeof(double) * nk * nz, hipMemcpyDeviceToHost));
  if (status) HIPCHK(hipMemcpy(status, dst, sizeof(int) * nk, hipMemcpyDeviceToHost));
  return 0;
}
}  // namespace

extern "C" int jdsg_get_1d_kernel(int nx, int ny, int nz, const float *vs, const float *dep,
                                  const double *tRc, int ntRc, int nmRc, const double *tRg,
                                  int ntRg, int nmRg, const double *tLc, int ntLc, int nmLc,
                                  const double *tLg, int ntLg, int nmLg, double *vc,
                                  double *vout, double *kernel, int *status) {
  return host_grid_call(0, nx, ny, nz, vs, dep, tRc, ntRc, nmRc, tRg, ntRg, nmRg, tLc, ntLc,
                        nmLc, tLg, ntLg, nmLg, vc, vout, kernel, status);
}

extern "C" int jdsg_get_2d_map(int nx, int ny, int nz, const float *vs, const float *dep,
                               const double *tRc, int ntRc, int nmRc, const double *tRg, int ntRg,
                               int nmRg, const double *tLc, int ntLc, int nmLc, const double *tLg,
                               int ntLg, int nmLg, double *vc, double *vout, int *status) {
  return host_grid_call(1, nx, ny, nz, vs, dep, tRc, ntRc, nmRc, tRg, ntRg, nmRg, tLc, ntLc,
                        nmLc, tLg, ntLg, nmLg, vc, vout, nullptr, status);
}

// ---------------------------------------------------------------- Fréchet rows
// SurfTime::frechet_matrix / compute_grad (SurfaceWave.cpp:554-676, 695-781)
// over the kernels of jdsg_surftime_kernel; device pointers, see jdsg_swd.h.
namespace {
int frechet_ctx(int nx, int ny, int nz, int kmax, int nrows, RowCtx *R) {
  if (nx < 3 || ny < 3 || nz < 2 || kmax < 1 || nrows < 0)
    return fail("frechet: need nx, ny >= 3, nz >= 2, kmax >= 1 and nrows >= 0");
  R->nx = nx;
  R->ny = ny;
  R->nz = nz;
  R->nxy = nx * ny;
  R->nin = (nx - 2) * (ny - 2);
  R->kmax = kmax;
  R->kstride = (long long)R->nxy * nz;
  return 0;
}
}  // namespace

extern "C" int jdsg_frechet_rows(int nx, int ny, int nz, int kmax, const double *kernel,
                                 int nrows, const int *row_idx, const long long *ray_off,
                                 const int *ray_cell, const float *ray_w, long long *row_off,
                                 int *col, float *val, long long cap, long long *nnz,
                                 void *stream) {
  RowCtx R;
  if (frechet_ctx(nx, ny, nz, kmax, nrows, &R)) return 1;
  hipStream_t st = (hipStream_t)stream;
  jdsg_launch_frechet_rows(R, kernel, nrows, row_idx, ray_off, ray_cell, ray_w, row_off, col, val,
                           cap, st);
  if (hipGetLastError() != hipSuccess) return fail("frechet: kernel launch failed");
  long long head[2] = {0, 0};
  if (nrows > 0) {
    HIPCHK(hipMemcpyAsync(&head[0], row_off, sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&head[1], row_off + nrows, sizeof(long long), hipMemcpyDeviceToHost,
                          st));
    HIPCHK(hipStreamSynchronize(st));
  }
  if (head[0] < 0)
    return fail("frechet: a row has a period index outside [0, kmax) or ray cells that are not "
                "increasing indices in [0, nx*ny)");
  if (nnz) *nnz = head[1];
  if (head[1] > cap) return fail("frechet: more nonzero entries than the output capacity");
  return 0;
}

extern "C" int jdsg_frechet_grad(int nx, int ny, int nz, int kmax, const double *kernel,
                                 int nrows, const int *row_idx, const long long *ray_off,
                                 const int *ray_cell, const float *ray_w, const float *resid,
                                 const float *vs, float *grad, void *stream) {
  RowCtx R;
  if (frechet_ctx(nx, ny, nz, kmax, nrows, &R)) return 1;
  jdsg_launch_frechet_grad(R, kernel, nrows, row_idx, ray_off, ray_cell, ray_w, resid, vs, grad,
                           (hipStream_t)stream);
  if (hipGetLastError() != hipSuccess) return fail("frechet: kernel launch failed");
  return 0;
}

extern "C" int jdsg_write_frechet(const char *path, int m, int n, int nrows, const int *row_id,
                                  const long long *row_off, const int *col, const float *val) {
  FILE *f = fopen(path, "wb");
  if (!f) return fail(std::string("frechet: cannot open ") + path);
  bool ok = fwrite(&m, sizeof(int), 1, f) == 1 && fwrite(&n, sizeof(int), 1, f) == 1;
  for (int r = 0; ok && r < nrows; ++r) {
    const long long a = row_off[r], b = row_off[r + 1];
    const int nar = (int)(b - a);
    ok = fwrite(&row_id[r], sizeof(int), 1, f) == 1 && fwrite(&nar, sizeof(int), 1, f) == 1 &&
         fwrite(col + a, sizeof(int), (size_t)nar, f) == (size_t)nar &&
         fwrite(val + a, sizeof(float), (size_t)nar, f) == (size_t)nar;
  }
  if (fclose(f) != 0) ok = false;
  if (!ok) return fail(std::string("frechet: write to ") + path + " failed");
  return 0;
}
