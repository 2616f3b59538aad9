// swd_nlopt.hip -- the nonlinear model update of an inversion iteration,
// NonlinOPT (include/nonlinear_opt.hpp:15-390), with its state in device
// memory: conjugate-gradient (Polak-Ribiere) or L-BFGS search directions in
// the log-model space, a trial step capped by the largest relative change,
// the multiplicative update x1 = x * exp(step * direc) and a bisecting
// line search on the Wolfe conditions.
//
// The reference keeps its history in files (grad<it>.bin, direc<it>.bin,
// logmd<it>.bin, syn<it>.bin under the result directory) and re-reads them
// for every direction.  Here the history is a ring of MSTORE + 1 = 6
// gradient and log-model vectors plus two search directions, resident in
// HBM, so an iteration reads each vector once per use and nothing crosses
// PCIe but the scalars of the line search.  The model object's three calls
// (compute_grad, compute_misfit, smoothing) are callbacks on device buffers.
//
// Reductions: each inner product sums float differences (computed in float,
// as the reference forms grad1 - grad0 and x1 - x0) as exact double
// products, 256 blocks in a fixed order and a fixed tree per block, then
// one block over the partials: the same bits every run.  The reference sums
// in float in Eigen's packet order; the double sums are closer to the exact
// value, and the coefficient computed from them is rounded to float as the
// reference's is.  Elementwise updates round each operation to float
// without contraction, in the reference's order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

namespace {

constexpr int kTB = 256;   // threads per block
constexpr int kRB = 256;   // reduction blocks
constexpr int kRing = 6;   // MSTORE + 1 iterations of history
constexpr int kMaxTrials = 100;

// sum over i of (a[i] - a0[i]) * (b[i] - b0[i]); a0 / b0 may be null
__global__ void __launch_bounds__(kTB) nl_dot_part(int n, const float *a, const float *a0,
                                                   const float *b, const float *b0, double *part) {
  __shared__ double s[kTB];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < n; i += (long long)kRB * kTB) {
    float x = a0 ? __fsub_rn(a[i], a0[i]) : a[i];
    float y = b0 ? __fsub_rn(b[i], b0[i]) : b[i];
    acc += (double)x * (double)y;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kRB) nl_dot_final(const double *part, double *out) {
  __shared__ double s[kRB];
  s[threadIdx.x] = part[threadIdx.x];
  __syncthreads();
  for (int o = kRB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// max |a[i]| (order-free)
__global__ void __launch_bounds__(kTB) nl_maxabs_part(int n, const float *a, float *part) {
  __shared__ float s[kTB];
  float m = 0.0f;
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < n; i += (long long)kRB * kTB)
    m = fmaxf(m, fabsf(a[i]));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kRB) nl_maxabs_final(const float *part, float *out) {
  __shared__ float s[kRB];
  s[threadIdx.x] = part[threadIdx.x];
  __syncthreads();
  for (int o = kRB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

#define NL_FOR_EACH(n) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (long long)gridDim.x * blockDim.x)

// out = -g + beta * d  (nonlinear_opt.hpp:84)
__global__ void nl_cg_dir(int n, const float *g, const float *d, float beta, float *out) {
  NL_FOR_EACH(n) out[i] = __fadd_rn(-g[i], __fmul_rn(beta, d[i]));
}
// out = s * a
__global__ void nl_scale(int n, float s, const float *a, float *out) {
  NL_FOR_EACH(n) out[i] = __fmul_rn(s, a[i]);
}
// q = q - a * (g1 - g0)  (nonlinear_opt.hpp:145)
__global__ void nl_q_update(int n, float a, const float *g1, const float *g0, float *q) {
  NL_FOR_EACH(n) q[i] = __fsub_rn(q[i], __fmul_rn(a, __fsub_rn(g1[i], g0[i])));
}
// r = r + (x1 - x0) * c  (nonlinear_opt.hpp:183)
__global__ void nl_r_update(int n, float c, const float *x1, const float *x0, float *r) {
  NL_FOR_EACH(n) r[i] = __fadd_rn(r[i], __fmul_rn(__fsub_rn(x1[i], x0[i]), c));
}
// x1 = x * exp(step * d)  (nonlinear_opt.hpp:259)
__global__ void nl_expstep(int n, const float *x, float step, const float *d, float *x1) {
  NL_FOR_EACH(n) x1[i] = __fmul_rn(x[i], expf(__fmul_rn(step, d[i])));
}
__global__ void nl_log(int n, const float *x, float *out) {
  NL_FOR_EACH(n) out[i] = logf(x[i]);
}

thread_local std::string g_nl_err;

}  // namespace

typedef int (*jdsg_nlopt_grad_fn)(void *user, const float *x, float *dsyn, float *grad, void *stream);
typedef float (*jdsg_nlopt_misfit_fn)(void *user, const float *dsyn, void *stream);
typedef int (*jdsg_nlopt_smooth_fn)(void *user, float *direc, void *stream);

struct jdsg_nlopt {
  int n = 0, m = 0;
  float max_step = 0.0f;
  int iter_start = 0, iter = 0;
  int flag = 0;  // 0 INIT, 1 GRAD, 2 LS
  int iter_ls = 0;
  float step_l = 0.0f, step_r = 0.0f, step_fac = -1.0f;
  hipStream_t st = nullptr;
  float *grad_ring[kRing] = {}, *logmd_ring[kRing] = {};
  float *direc_hist[2] = {};  // the direction saved per iteration, before the angle check
  float *grad = nullptr, *direc = nullptr, *x1 = nullptr, *grad_next = nullptr;
  float *dsyn_next = nullptr, *dsyn_saved = nullptr, *q = nullptr, *r = nullptr;
  double *dpart = nullptr, *dout = nullptr;
  float *fpart = nullptr, *fout = nullptr;
  float *owned = nullptr;  // the block the library allocated, if any
  // statistics of the last update
  int trials = 0;
  float fcost = 0.0f, fcost1 = 0.0f;
  // callbacks of the running update
  jdsg_nlopt_grad_fn gradf = nullptr;
  jdsg_nlopt_misfit_fn misfitf = nullptr;
  jdsg_nlopt_smooth_fn smoothf = nullptr;
  void *user = nullptr;
};

namespace {

bool ck(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    g_nl_err = std::string("nlopt: ") + what + ": " + hipGetErrorString(e);
    return false;
  }
  return true;
}

int grid_for(int n) {
  long long b = ((long long)n + kTB - 1) / kTB;
  return (int)(b < 2048 ? (b > 0 ? b : 1) : 2048);
}

// inner product, read back (synchronises the stream)
bool dot(jdsg_nlopt *h, const float *a, const float *a0, const float *b, const float *b0,
         double *res) {
  hipLaunchKernelGGL(nl_dot_part, dim3(kRB), dim3(kTB), 0, h->st, h->n, a, a0, b, b0, h->dpart);
  hipLaunchKernelGGL(nl_dot_final, dim3(1), dim3(kRB), 0, h->st, h->dpart, h->dout);
  if (!ck(hipGetLastError(), "dot launch")) return false;
  if (!ck(hipMemcpyAsync(res, h->dout, sizeof(double), hipMemcpyDeviceToHost, h->st), "dot read"))
    return false;
  return ck(hipStreamSynchronize(h->st), "dot sync");
}

bool maxabs(jdsg_nlopt *h, const float *a, float *res) {
  hipLaunchKernelGGL(nl_maxabs_part, dim3(kRB), dim3(kTB), 0, h->st, h->n, a, h->fpart);
  hipLaunchKernelGGL(nl_maxabs_final, dim3(1), dim3(kRB), 0, h->st, h->fpart, h->fout);
  if (!ck(hipGetLastError(), "maxabs launch")) return false;
  if (!ck(hipMemcpyAsync(res, h->fout, sizeof(float), hipMemcpyDeviceToHost, h->st), "maxabs read"))
    return false;
  return ck(hipStreamSynchronize(h->st), "maxabs sync");
}

bool copy(jdsg_nlopt *h, float *dst, const float *src, int n) {
  return ck(hipMemcpyAsync(dst, src, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, h->st),
            "copy");
}

float *ring(float *const *r, int it) { return r[((it % kRing) + kRing) % kRing]; }

// get_cg_direction (nonlinear_opt.hpp:65-90)
bool cg_direction(jdsg_nlopt *h, int it) {
  const int n = h->n, g = grid_for(n);
  if (it == h->iter_start) {
    hipLaunchKernelGGL(nl_scale, dim3(g), dim3(kTB), 0, h->st, n, -1.0f, h->grad, h->direc);
  } else {
    const float *gp = ring(h->grad_ring, it - 1), *dp = h->direc_hist[(it - 1) & 1];
    double num, den;
    if (!dot(h, h->grad, nullptr, h->grad, gp, &num) || !dot(h, gp, nullptr, gp, nullptr, &den))
      return false;
    const float beta = (float)num / (float)den;
    hipLaunchKernelGGL(nl_cg_dir, dim3(g), dim3(kTB), 0, h->st, n, h->grad, dp, beta, h->direc);
  }
  return copy(h, h->direc_hist[it & 1], h->direc, n);
}

// get_lbfgs_direction (nonlinear_opt.hpp:92-188): the two-loop recursion
// over the pairs (logmd, grad) of iterations iter_store .. it
bool lbfgs_direction(jdsg_nlopt *h, int it) {
  const int n = h->n, g = grid_for(n);
  if (it == h->iter_start) {
    hipLaunchKernelGGL(nl_scale, dim3(g), dim3(kTB), 0, h->st, n, -1.0f, ring(h->grad_ring, it),
                       h->direc);
    return ck(hipGetLastError(), "lbfgs launch");
  }
  int iter_store = it - 5;
  if (iter_store <= h->iter_start) iter_store = h->iter_start;
  float a[kRing] = {}, p[kRing] = {};
  if (!copy(h, h->q, ring(h->grad_ring, it), n)) return false;
  for (int is = it - 1; is >= iter_store; --is) {
    const float *g1 = ring(h->grad_ring, is + 1), *g0 = ring(h->grad_ring, is);
    const float *x1 = ring(h->logmd_ring, is + 1), *x0 = ring(h->logmd_ring, is);
    double pt, at;
    if (!dot(h, g1, g0, x1, x0, &pt) || !dot(h, x1, x0, h->q, nullptr, &at)) return false;
    const int k = is - iter_store;
    p[k] = (float)(1. / (float)pt);
    a[k] = p[k] * (float)at;
    hipLaunchKernelGGL(nl_q_update, dim3(g), dim3(kTB), 0, h->st, n, a[k], g1, g0, h->q);
  }
  {
    const float *g1 = ring(h->grad_ring, it), *g0 = ring(h->grad_ring, it - 1);
    const float *x1 = ring(h->logmd_ring, it), *x0 = ring(h->logmd_ring, it - 1);
    double up, down;
    if (!dot(h, g1, g0, x1, x0, &up) || !dot(h, g1, g0, g1, g0, &down)) return false;
    const float pk = (float)up / (float)down;
    hipLaunchKernelGGL(nl_scale, dim3(g), dim3(kTB), 0, h->st, n, pk, h->q, h->r);
  }
  for (int is = iter_store; is <= it - 1; ++is) {
    const float *g1 = ring(h->grad_ring, is + 1), *g0 = ring(h->grad_ring, is);
    const float *x1 = ring(h->logmd_ring, is + 1), *x0 = ring(h->logmd_ring, is);
    double bt;
    if (!dot(h, g1, g0, h->r, nullptr, &bt)) return false;
    const int k = is - iter_store;
    const float b = p[k] * (float)bt;
    hipLaunchKernelGGL(nl_r_update, dim3(g), dim3(kTB), 0, h->st, n, a[k] - b, x1, x0, h->r);
  }
  hipLaunchKernelGGL(nl_scale, dim3(g), dim3(kTB), 0, h->st, n, -1.0f, h->r, h->direc);
  return ck(hipGetLastError(), "lbfgs launch");
}

// get_next_model (nonlinear_opt.hpp:201-268), writing h->x1 and h->direc
bool next_model(jdsg_nlopt *h, const float *x, float chi, int method) {
  const int n = h->n, g = grid_for(n), it = h->iter;
  if (!(method == 0 ? cg_direction(h, it) : lbfgs_direction(h, it))) return false;
  if (it != h->iter_start) {
    // get_angle (:190-199) between direc and -grad
    double dd, gg, dg;
    if (!dot(h, h->direc, nullptr, h->direc, nullptr, &dd) ||
        !dot(h, h->grad, nullptr, h->grad, nullptr, &gg) ||
        !dot(h, h->direc, nullptr, h->grad, nullptr, &dg))
      return false;
    const float l1 = std::sqrt((float)dd), l2 = std::sqrt((float)gg);
    const float ca = (float)(-dg / ((double)l1 * (double)l2));
    const float theta = (float)(std::acos(ca) * 180. / M_PI);
    if (theta > 92) {
      hipLaunchKernelGGL(nl_scale, dim3(g), dim3(kTB), 0, h->st, n, -1.0f, h->grad, h->direc);
      h->iter_start = it;
    }
  }
  if (h->smoothf && h->smoothf(h->user, h->direc, (void *)h->st) != 0) {
    g_nl_err = "nlopt: the smoothing callback failed";
    return false;
  }
  float dmax;
  if (!maxabs(h, h->direc, &dmax)) return false;
  float step_fac = h->step_fac;
  if (it == h->iter_start && h->iter_ls == 0) step_fac = -1.;
  if (it == h->iter_start + 1 && h->iter_ls == 0) {
    if (method == 0) {
      double gd;
      if (!dot(h, h->grad, nullptr, h->direc, nullptr, &gd)) return false;
      const float g1 = (float)gd;
      step_fac = -2. / g1 * chi;
    } else {
      step_fac = 1.;
    }
  }
  if (step_fac < 0 || step_fac * dmax > h->max_step) step_fac = h->max_step / dmax;
  h->step_fac = step_fac;
  hipLaunchKernelGGL(nl_expstep, dim3(g), dim3(kTB), 0, h->st, n, x, h->step_fac, h->direc, h->x1);
  if (h->flag != 2) {
    h->flag = 2;
    h->iter_ls = 0;
  }
  return ck(hipGetLastError(), "next-model launch");
}

}  // namespace

extern "C" const char *jdsg_nlopt_last_error(void) { return g_nl_err.c_str(); }

// one block holds every buffer: partial sums first (8-byte aligned), then
// the vectors, each padded to 64 floats
extern "C" long long jdsg_nlopt_work_floats(int n, int m) {
  const long long vn = ((long long)n + 63) / 64 * 64, vm = ((long long)(m > 0 ? m : 1) + 63) / 64 * 64;
  return 2 * kRB + 64 + kRB + 64 + (2 * kRing + 8) * vn + 2 * vm;
}

extern "C" void jdsg_nlopt_destroy(jdsg_nlopt *h) {
  if (!h) return;
  if (h->owned) (void)hipFree(h->owned);
  delete h;
}

// NonlinOPT(outdir, max_step, iter_start, iter_current) (nonlinear_opt.hpp:36-48).
// work: null (the library allocates) or jdsg_nlopt_work_floats(n, m) device
// floats, 8-byte aligned, that the caller keeps until jdsg_nlopt_destroy.
extern "C" int jdsg_nlopt_create(int n, int m, float max_step, int iter_start, int iter_current,
                                 float *work, void *stream, jdsg_nlopt **out) {
  if (!out || n <= 0 || m < 0 || ((size_t)work & 7)) {
    g_nl_err = "nlopt_create: bad argument";
    return 1;
  }
  // the history of earlier iterations (the reference's grad/direc/logmd/syn
  // files) is a device ring that starts empty: a resumed run with
  // iter_current past iter_start would read an unset history
  if (iter_current != iter_start) {
    g_nl_err = "nlopt_create: iter_current must equal iter_start (the history starts empty; "
               "restart the optimiser at the current iteration)";
    return 1;
  }
  const long long nw = jdsg_nlopt_work_floats(n, m);
  float *owned = nullptr;
  if (!work) {
    if (!ck(hipMalloc(&owned, (size_t)nw * sizeof(float)), "alloc")) return 1;
    work = owned;
  }
  jdsg_nlopt *h = new jdsg_nlopt;
  h->owned = owned;
  h->n = n;
  h->m = m;
  h->max_step = max_step;
  h->iter_start = iter_start;
  h->iter = iter_current;
  h->st = (hipStream_t)stream;
  const long long vn = ((long long)n + 63) / 64 * 64, vm = ((long long)(m > 0 ? m : 1) + 63) / 64 * 64;
  float *p = work;
  h->dpart = (double *)p;
  p += 2 * kRB;
  h->dout = (double *)p;
  p += 64;
  h->fpart = p;
  p += kRB;
  h->fout = p;
  p += 64;
  for (int i = 0; i < kRing; ++i) {
    h->grad_ring[i] = p;
    p += vn;
    h->logmd_ring[i] = p;
    p += vn;
  }
  float **vecs[] = {&h->grad, &h->direc, &h->x1, &h->grad_next, &h->q, &h->r, &h->direc_hist[0],
                    &h->direc_hist[1]};
  for (float **v : vecs) {
    *v = p;
    p += vn;
  }
  h->dsyn_next = p;
  p += vm;
  h->dsyn_saved = p;
  *out = h;
  return 0;
}

// NonlinOPT::update (nonlinear_opt.hpp:271-389).  x: n device floats, the
// model, replaced by the accepted one; dsyn: m device floats, the
// synthetics of the model as it was on entry.  method 0 = "CG", 1 = "LBFGS".
extern "C" int jdsg_nlopt_update(jdsg_nlopt *h, float *x, float *dsyn, int method,
                                 jdsg_nlopt_grad_fn gradf, jdsg_nlopt_misfit_fn misfitf,
                                 jdsg_nlopt_smooth_fn smoothf, void *user) {
  if (!h || !x || !dsyn || !gradf || !misfitf || (method != 0 && method != 1)) {
    g_nl_err = "nlopt_update: bad argument";
    return 1;
  }
  h->gradf = gradf;
  h->misfitf = misfitf;
  h->smoothf = smoothf;
  h->user = user;
  h->trials = 0;
  const int n = h->n, g = grid_for(n);
  void *st = (void *)h->st;
  if (h->flag == 0) {
    if (gradf(user, x, dsyn, h->grad, st) != 0) {
      g_nl_err = "nlopt: the gradient callback failed";
      return 1;
    }
    if (!copy(h, ring(h->grad_ring, h->iter), h->grad, n)) return 1;
    hipLaunchKernelGGL(nl_log, dim3(g), dim3(kTB), 0, h->st, n, x, ring(h->logmd_ring, h->iter));
  } else if (h->flag == 1) {
    if (!copy(h, h->grad, ring(h->grad_ring, h->iter), n)) return 1;
    if (h->m > 0 &&
        !ck(hipMemcpyAsync(dsyn, h->dsyn_saved, (size_t)h->m * sizeof(float),
                           hipMemcpyDeviceToDevice, h->st), "copy"))
      return 1;
  }
  const float fcost = misfitf(user, dsyn, st);
  h->fcost = fcost;
  if (!next_model(h, x, fcost, method)) return 1;
  const float m1 = 1.0E-4f, m2 = method == 0 ? 0.1f : 0.9f;
  while (h->flag == 2) {
    if (++h->trials > kMaxTrials) {
      g_nl_err = "nlopt: no step met the Wolfe conditions";
      return 2;
    }
    if (gradf(user, h->x1, h->dsyn_next, h->grad_next, st) != 0) {
      g_nl_err = "nlopt: the gradient callback failed";
      return 1;
    }
    if (h->iter_ls == 0) {
      h->step_l = 0.;
      h->step_r = 0.;
    }
    const float fcost1 = misfitf(user, h->dsyn_next, st);
    h->fcost1 = fcost1;
    double qd, q1d;
    if (!dot(h, h->grad, nullptr, h->direc, nullptr, &qd) ||
        !dot(h, h->grad_next, nullptr, h->direc, nullptr, &q1d))
      return 1;
    const float q = (float)qd, q1 = (float)q1d;
    const bool cond1 = fcost1 <= (fcost + m1 * h->step_fac * q);
    const bool cond2 = q1 >= m2 * q;
    if (cond1 && cond2) {
      h->flag = 1;
      h->iter += 1;
      h->iter_ls = 0;
      if (!copy(h, ring(h->grad_ring, h->iter), h->grad_next, n)) return 1;
      if (h->m > 0 &&
          !ck(hipMemcpyAsync(h->dsyn_saved, h->dsyn_next, (size_t)h->m * sizeof(float),
                             hipMemcpyDeviceToDevice, h->st), "copy"))
        return 1;
      if (!copy(h, x, h->x1, n)) return 1;
      hipLaunchKernelGGL(nl_log, dim3(g), dim3(kTB), 0, h->st, n, x, ring(h->logmd_ring, h->iter));
      if (!ck(hipGetLastError(), "log launch")) return 1;
      break;
    } else if (!cond1) {
      h->step_r = h->step_fac;
      h->step_fac = 0.5 * (h->step_r + h->step_l);
      h->iter_ls += 1;
    } else {
      h->step_l = h->step_fac;
      const float tried = h->step_fac;
      const bool grow = h->step_r == 0.;
      if (!grow)
        h->step_fac = 0.5 * (h->step_l + h->step_r);
      else
        h->step_fac *= 10.0f;
      h->iter_ls += 1;
      if (!next_model(h, x, fcost, method)) return 1;
      // the tenfold step capped back to the step just tried: every further
      // trial would repeat it (the reference loops on), so stop here
      if (grow && h->step_fac == tried) {
        g_nl_err = "nlopt: no step met the Wolfe conditions (the capped trial step repeats)";
        return 2;
      }
      continue;
    }
    if (!next_model(h, x, fcost, method)) return 1;
  }
  return ck(hipStreamSynchronize(h->st), "update sync") ? 0 : 1;
}

// iter, iter_start, the line-search trials of the last update, the trial
// step, and the misfits before and after the last update
extern "C" void jdsg_nlopt_state(const jdsg_nlopt *h, int *iter, int *iter_start, int *trials,
                                 float *step_fac, float *fcost, float *fcost1) {
  if (!h) return;
  if (iter) *iter = h->iter;
  if (iter_start) *iter_start = h->iter_start;
  if (trials) *trials = h->trials;
  if (step_fac) *step_fac = h->step_fac;
  if (fcost) *fcost = h->fcost;
  if (fcost1) *fcost1 = h->fcost1;
}
