// swd_lsmr.hip -- LSMR on a CSR matrix on the GPU: the least-squares solve
// that follows the Frechet rows in an inversion step (src/shared/
// lsmrModule_csr.f90 LSMR of Fong and Saunders with local
// reorthogonalisation, called through csr_matrix::lsmr_solver,
// src/shared/csr_matrix.cpp:114-128; products src/shared/parallel.cpp).
//
// Vectors and the scalar recurrences are float, as in the reference build
// (real(dp) = c_float).  The products run on the device: y += A x with one
// wavefront per row, x += A^T y with one wavefront per column of a transpose
// built once per solve (a stable radix sort of the entries by column keeps
// each column in row order, so the sums are deterministic).  Norms and the
// reorthogonalisation dot products are fixed-order block reductions in
// double; the reorthogonalisation coefficient goes from its reduction to its
// update on the device.  The scalar recurrences of each iteration, three
// norms read back, run on the host in the reference's order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "swd_common.hpp"

namespace {

constexpr int kRedBlocks = 512;  // blocks of the fixed-order reductions

// flag[0] = 1 if ptr[0] != 0, a row pointer decreases or a column index
// leaves [0, n); thread i checks row i and entry i
__global__ void lsmr_check(int m, int n, long long nnz, const int *__restrict__ ptr,
                           const int *__restrict__ ind, int *flag) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && ((i == 0 && ptr[0] != 0) || ptr[i + 1] < ptr[i])) flag[0] = 1;
  if (i < nnz && (ind[i] < 0 || ind[i] >= n)) flag[0] = 1;
}

// rowidx[j] = the row of entry j: one wavefront per row, 4 rows per block
__global__ void __launch_bounds__(256)
lsmr_row_index(int m, const int *__restrict__ ptr, int *rowidx) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= m) return;
  for (int j = ptr[i] + (threadIdx.x & 63); j < ptr[i + 1]; j += 64) rowidx[j] = i;
}

__global__ void lsmr_iota(long long n, int *p) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int)i;
}

// colptr[c] = the first position k with key[k] >= c, for c in [0, n], from
// keys sorted ascending: thread k (0 <= k <= nnz) writes the columns in
// (key[k-1], key[k]], taking key[-1] = -1 and key[nnz] = n, so every column
// is written exactly once
__global__ void lsmr_col_bounds(long long nnz, int n, const int *__restrict__ key, int *colptr) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nnz) return;
  const int lo = k == 0 ? 0 : key[k - 1] + 1;
  const int hi = k == nnz ? n : key[k];
  for (int c = lo; c <= hi; ++c) colptr[c] = (int)k;
}

__global__ void lsmr_gather(long long nnz, const int *__restrict__ perm,
                            const float *__restrict__ val, const int *__restrict__ rowidx,
                            float *cval, int *crow) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int p = perm[k];
  cval[k] = val[p];
  crow[k] = rowidx[p];
}

// y[i] = a * y[i] + sum_j A(i, j) x[j] with G lanes per row (256 / G rows
// per block): G = 64 for the long rows of A (thousands of entries), fewer
// for the short columns of A^T.  Each lane keeps four loads in flight; the
// lane sums are combined in a fixed order, so results repeat bit for bit.
template <int G>
__global__ void __launch_bounds__(256)
lsmr_spmv(int m, const int *__restrict__ ptr, const int *__restrict__ ind,
          const float *__restrict__ val, const float *__restrict__ x, float a, float *y) {
  const int row = blockIdx.x * (256 / G) + (int)threadIdx.x / G;
  const int lane = (int)threadIdx.x % G;
  if (row >= m) return;  // a whole group leaves together
  const int j1 = ptr[row + 1];
  int j = ptr[row] + lane;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  for (; j + 3 * G < j1; j += 4 * G) {
    const float v0 = val[j], v1 = val[j + G], v2 = val[j + 2 * G], v3 = val[j + 3 * G];
    const int i0 = ind[j], i1 = ind[j + G], i2 = ind[j + 2 * G], i3 = ind[j + 3 * G];
    s0 += v0 * x[i0];
    s1 += v1 * x[i1];
    s2 += v2 * x[i2];
    s3 += v3 * x[i3];
  }
  for (; j < j1; j += G) s0 += val[j] * x[ind[j]];
  float s = (s0 + s1) + (s2 + s3);
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (lane == 0) y[row] = a * y[row] + s;
}

// The same product with value and index read 16 B at a time (val and ind
// 16-byte aligned): the row's entries before its first 4-aligned position
// and from its last one on (at most 3 each) are taken one per lane, the
// aligned quads in between G lanes at a time, two quads in flight per lane.
template <int G>
__global__ void __launch_bounds__(256)
lsmr_spmv4(int m, const int *__restrict__ ptr, const int *__restrict__ ind,
           const float *__restrict__ val, const float *__restrict__ x, float a, float *y) {
  const int row = blockIdx.x * (256 / G) + (int)threadIdx.x / G;
  const int lane = (int)threadIdx.x % G;
  if (row >= m) return;  // a whole group leaves together
  const int j0 = ptr[row], j1 = ptr[row + 1];
  const int ja = (int)min((long long)j0 + 3 - ((long long)j0 + 3) % 4, (long long)j1);
  const int jb = max(ja, j1 - j1 % 4);
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  if (j0 + lane < ja) s0 = val[j0 + lane] * x[ind[j0 + lane]];
  if (jb + lane < j1) s1 = val[jb + lane] * x[ind[jb + lane]];
  const float4 *v4 = reinterpret_cast<const float4 *>(val);
  const int4 *i4 = reinterpret_cast<const int4 *>(ind);
  const int qe = jb / 4;
  int q = ja / 4 + lane;
  for (; q + G < qe; q += 2 * G) {
    const float4 va = v4[q], vb = v4[q + G];
    const int4 ia = i4[q], ib = i4[q + G];
    s0 += va.x * x[ia.x];
    s1 += va.y * x[ia.y];
    s2 += va.z * x[ia.z];
    s3 += va.w * x[ia.w];
    s0 += vb.x * x[ib.x];
    s1 += vb.y * x[ib.y];
    s2 += vb.z * x[ib.z];
    s3 += vb.w * x[ib.w];
  }
  if (q < qe) {
    const float4 va = v4[q];
    const int4 ia = i4[q];
    s0 += va.x * x[ia.x];
    s1 += va.y * x[ia.y];
    s2 += va.z * x[ia.z];
    s3 += va.w * x[ia.w];
  }
  float s = (s0 + s1) + (s2 + s3);
  if constexpr (G <= 64) {
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
    if (lane == 0) y[row] = a * y[row] + s;
  } else {  // a block per row: the four wavefront sums combined in order
    __shared__ float sh[4];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((lane & 63) == 0) sh[lane >> 6] = s;
    __syncthreads();
    if (lane == 0) y[row] = a * y[row] + ((sh[0] + sh[1]) + (sh[2] + sh[3]));
  }
}

// The 16 B form for a gathered vector short enough to sit in LDS (the A^T
// product of a Frechet system gathers its few thousand rows): each block of
// 1024 threads copies x into LDS once, then walks rows [b R, b R + R),
// [b R + grid R, ...) with G lanes per row; the per-row arithmetic and its
// order are those of lsmr_spmv4<G>, with x read from LDS.
template <int G>
__global__ void __launch_bounds__(1024)
lsmr_spmv4_lds(int m, int xlen, const int *__restrict__ ptr, const int *__restrict__ ind,
               const float *__restrict__ val, const float *__restrict__ x, float a, float *y) {
  extern __shared__ float xs[];
  for (int i = threadIdx.x; i < xlen; i += 1024) xs[i] = x[i];
  __syncthreads();
  constexpr int R = 1024 / G;
  const int lane = (int)threadIdx.x % G;
  const float4 *v4 = reinterpret_cast<const float4 *>(val);
  const int4 *i4 = reinterpret_cast<const int4 *>(ind);
  for (int row = blockIdx.x * R + (int)threadIdx.x / G; row < m; row += gridDim.x * R) {
    const int j0 = ptr[row], j1 = ptr[row + 1];
    const int ja = (int)min((long long)j0 + 3 - ((long long)j0 + 3) % 4, (long long)j1);
    const int jb = max(ja, j1 - j1 % 4);
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    if (j0 + lane < ja) s0 = val[j0 + lane] * xs[ind[j0 + lane]];
    if (jb + lane < j1) s1 = val[jb + lane] * xs[ind[jb + lane]];
    const int qe = jb / 4;
    int q = ja / 4 + lane;
    for (; q + G < qe; q += 2 * G) {
      const float4 va = v4[q], vb = v4[q + G];
      const int4 ia = i4[q], ib = i4[q + G];
      s0 += va.x * xs[ia.x];
      s1 += va.y * xs[ia.y];
      s2 += va.z * xs[ia.z];
      s3 += va.w * xs[ia.w];
      s0 += vb.x * xs[ib.x];
      s1 += vb.y * xs[ib.y];
      s2 += vb.z * xs[ib.z];
      s3 += vb.w * xs[ib.w];
    }
    if (q < qe) {
      const float4 va = v4[q];
      const int4 ia = i4[q];
      s0 += va.x * xs[ia.x];
      s1 += va.y * xs[ia.y];
      s2 += va.z * xs[ia.z];
      s3 += va.w * xs[ia.w];
    }
    float s = (s0 + s1) + (s2 + s3);
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
    if (lane == 0) y[row] = a * y[row] + s;
  }
}

// lanes per row: the smallest of 4..64 with 4 G >= the average number of
// work items per row (entries, or quads of entries for the 16 B form)
// (quads: 256, a block per row, for rows of more than 1024 quads)
int spmv_lanes(long long nnz, int rows, bool quads) {
  double avg = rows > 0 ? (double)nnz / rows : 0.0;
  if (quads) avg /= 4.0;
  int g = 4;
  while (g < 64 && (quads ? 4.0 : 8.0) * g < avg) g *= 2;
  if (quads && avg > 1024.0) g = 256;
  return g;
}

template <int G>
void launch_spmv_g(bool quads, dim3 grd, int rows, const int *ptr, const int *ind,
                   const float *val, const float *x, float a, float *y, hipStream_t st) {
  if (quads)
    hipLaunchKernelGGL(lsmr_spmv4<G>, grd, dim3(256), 0, st, rows, ptr, ind, val, x, a, y);
  else
    hipLaunchKernelGGL(lsmr_spmv<G>, grd, dim3(256), 0, st, rows, ptr, ind, val, x, a, y);
}

// LDS launch shape of lsmr_spmv4_lds for a gathered vector of xlen floats:
// blocks = CUs x the blocks whose LDS fits a CU, or 0 when x does not fit
// one block's LDS (then the product gathers from memory)
struct LdsShape {
  int blocks = 0;
  size_t bytes = 0;
};

LdsShape lds_shape(int xlen) {
  LdsShape sh;
  int dev = 0, per_block = 0, per_cu = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
          hipSuccess ||
      hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
          hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return sh;
  const size_t bytes = sizeof(float) * (size_t)std::max(xlen, 1);
  if (bytes > (size_t)per_block || cus <= 0) return sh;
  const int fit = std::max(1, std::min(2, per_cu / (int)bytes));
  sh.blocks = cus * fit;
  sh.bytes = bytes;
  return sh;
}

template <int G>
hipError_t launch_lds_g(const LdsShape &sh, int rows, int xlen, const int *ptr, const int *ind,
                        const float *val, const float *x, float a, float *y, hipStream_t st) {
  const hipError_t e = hipFuncSetAttribute((const void *)lsmr_spmv4_lds<G>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)sh.bytes);
  if (e != hipSuccess) return e;
  const int need = (rows + 1024 / G - 1) / (1024 / G);
  hipLaunchKernelGGL(lsmr_spmv4_lds<G>, dim3((unsigned)std::max(1, std::min(sh.blocks, need))),
                     dim3(1024), sh.bytes, st, rows, xlen, ptr, ind, val, x, a, y);
  return hipGetLastError();
}

hipError_t launch_spmv_lds(int g, const LdsShape &sh, int rows, int xlen, const int *ptr,
                           const int *ind, const float *val, const float *x, float a, float *y,
                           hipStream_t st) {
  switch (g) {
    case 4: return launch_lds_g<4>(sh, rows, xlen, ptr, ind, val, x, a, y, st);
    case 8: return launch_lds_g<8>(sh, rows, xlen, ptr, ind, val, x, a, y, st);
    case 16: return launch_lds_g<16>(sh, rows, xlen, ptr, ind, val, x, a, y, st);
    case 32: return launch_lds_g<32>(sh, rows, xlen, ptr, ind, val, x, a, y, st);
    default: return launch_lds_g<64>(sh, rows, xlen, ptr, ind, val, x, a, y, st);
  }
}

void launch_spmv(int g, bool quads, int rows, const int *ptr, const int *ind, const float *val,
                 const float *x, float a, float *y, hipStream_t st) {
  const dim3 grd((unsigned)((rows + 256 / g - 1) / (256 / g)));
  switch (g) {
    case 4: launch_spmv_g<4>(quads, grd, rows, ptr, ind, val, x, a, y, st); break;
    case 8: launch_spmv_g<8>(quads, grd, rows, ptr, ind, val, x, a, y, st); break;
    case 16: launch_spmv_g<16>(quads, grd, rows, ptr, ind, val, x, a, y, st); break;
    case 32: launch_spmv_g<32>(quads, grd, rows, ptr, ind, val, x, a, y, st); break;
    case 64: launch_spmv_g<64>(quads, grd, rows, ptr, ind, val, x, a, y, st); break;
    default: hipLaunchKernelGGL(lsmr_spmv4<256>, grd, dim3(256), 0, st, rows, ptr, ind, val, x, a, y);
  }
}

// part[block] = sum over the block's grid-stride share of x*y (x*x if y is
// null), in double; the block partials are summed in a fixed order after
__global__ void lsmr_reduce_f(int n, const float *__restrict__ x, const float *__restrict__ y,
                              double *part) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    s += (double)x[i] * (double)(y ? y[i] : x[i]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// *out = the sum of part[0..nb) in a fixed order (one block of 256)
__global__ void __launch_bounds__(256)
lsmr_reduce_final(int nb, const double *__restrict__ part, double *out) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ void lsmr_scale(int n, float a, float *x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = a * x[i];
}

// the sum of part[0..nb) in a fixed order, taken by a whole block (every
// block that calls it gets the same bits)
__device__ double block_sum_fixed(int nb, const double *__restrict__ part, double *sh) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  const double t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return t;
}

// v -= d qp with d = (float) (the fixed-order sum of the previous pass's
// partials part_in[0..nb)), then part[block] = the block's share of v . qn
// (v . v if qn is null) over the new v, in double: one pass per vector of
// the reorthogonalisation (the next dot product, or the norm of v after the
// last vector, rides on the subtraction)
__global__ void __launch_bounds__(256)
lsmr_sub_dot(int n, int nb, const double *__restrict__ part_in, const float *__restrict__ qp,
             const float *__restrict__ qn, float *v, double *part) {
  __shared__ double sh[4];
  const float d = (float)block_sum_fixed(nb, part_in, sh);
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float w = v[i] - d * qp[i];
    v[i] = w;
    s += (double)w * (double)(qn ? qn[i] : w);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// hbar = h - fh hbar;  x = x + fx hbar;  h = v - fn h  (lsmrModule_csr.f90:1003-1006),
// with part[block] = the block's share of x . x over the new x
__global__ void __launch_bounds__(256)
lsmr_update(int n, float fh, float fx, float fn, const float *__restrict__ v, float *h,
            float *hbar, float *x, double *part) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float hb = h[i] - fh * hbar[i];
    hbar[i] = hb;
    const float xi = x[i] + fx * hb;
    x[i] = xi;
    h[i] = v[i] - fn * h[i];
    s += (double)xi * (double)xi;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Cell-major renumbering of a layer-major column numbering: column
// c = l * nc + cell (l < layers) becomes cell * layers + l, so the layers of
// one cell sit side by side in x.
__device__ __forceinline__ int cell_major(int c, int nc, int layers) {
  return (c % nc) * layers + c / nc;
}

// key[k] = (row of entry k) << nb | its cell-major column: sorting these
// keys orders the entries by row, then by renumbered column
__global__ void lsmr_layer_keys(long long nnz, int nc, int layers, int nb,
                                const int *__restrict__ rowidx, const int *__restrict__ ind,
                                unsigned long long *key) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  key[k] = ((unsigned long long)rowidx[k] << nb) |
           (unsigned long long)cell_major(ind[k], nc, layers);
}

// the renumbered matrix: column from the sorted key, value from the entry
// the sort moved there
__global__ void lsmr_layer_take(long long nnz, unsigned long long mask,
                                const unsigned long long *__restrict__ key,
                                const int *__restrict__ perm, const float *__restrict__ val,
                                int *lind, float *lval) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  lind[k] = (int)(key[k] & mask);
  lval[k] = val[perm[k]];
}

// x in the caller's numbering from the cell-major solution
__global__ void lsmr_layer_out(int n, int nc, int layers, const float *__restrict__ xs, float *x) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) x[c] = xs[cell_major(c, nc, layers)];
}

float d2norm(float a, float b) {  // lsmrModule_csr.f90:1151-1163
  const float scale = fabsf(a) + fabsf(b);
  if (scale == 0.0f) return 0.0f;
  const float p = a / scale, q = b / scale;
  return scale * sqrtf(p * p + q * q);
}

struct LsmrWork {
  std::vector<void *> bufs;
  hipStream_t st;
  int err = 0;
  template <class T>
  T *get(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      err = 1;
      return nullptr;
    }
    bufs.push_back(p);
    return (T *)p;
  }
  ~LsmrWork() {
    for (void *p : bufs) (void)hipFree(p);
  }
};

}  // namespace

// The last error message of the LSMR entry points (see jdsg_lsmr_csr).
static thread_local std::string g_lsmr_err;
extern "C" const char *jdsg_lsmr_last_error(void) { return g_lsmr_err.c_str(); }

// LSMR_csr (lsmrModule_csr.f90:516-1148) with the matrix, b and x in device
// memory; layers > 1 solves in the cell-major numbering of a layer-major
// one.  Returns 0, or nonzero with a message on a bad argument or a HIP
// failure.
static int lsmr_impl(int layers, int m, int n, const float *val, const int *indices,
                     const int *indptr, const float *b, float damp, float atol, float btol,
                     float conlim, int itnlim, int localSize, float *x, int *istop_out,
                     int *itn_out, float *normA_out, float *condA_out, float *normr_out,
                     float *normAr_out, float *normx_out, void *stream) {
  auto fail = [](const char *msg) {
    g_lsmr_err = msg;
    return 1;
  };
  if (m <= 0 || n <= 0 || !val || !indices || !indptr || !b || !x)
    return fail("lsmr: empty matrix or null pointer");
  if (localSize < 0) return fail("lsmr: localSize < 0");
  if (layers < 1 || n % layers != 0) return fail("lsmr: layers must be >= 1 and divide n");
  hipStream_t st = (hipStream_t)stream;
  int nnz_i = 0;
  if (hipMemcpyAsync(&nnz_i, indptr + m, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail("lsmr: cannot read indptr[m]");
  if (nnz_i < 0) return fail("lsmr: negative entry count");
  const long long nnz = nnz_i;
  int localVecs = std::min(localSize, std::min(m, n));
  LsmrWork W;
  W.st = st;
  float *u = W.get<float>(m), *v = W.get<float>(n), *h = W.get<float>(n),
        *hbar = W.get<float>(n);
  float *lv = W.get<float>((size_t)n * std::max(localVecs, 1));
  double *part = W.get<double>(kRedBlocks), *part2 = W.get<double>(kRedBlocks),
         *red = W.get<double>(1 + (size_t)localVecs);
  int *rowidx = W.get<int>(nnz), *perm = W.get<int>(nnz), *perm2 = W.get<int>(nnz),
      *key2 = W.get<int>(nnz), *colptr = W.get<int>((size_t)n + 1), *crow = W.get<int>(nnz),
      *cnt = W.get<int>(1);
  float *cval = W.get<float>(nnz);
  if (W.err) return fail("lsmr: out of device memory");
#define LCHK(e)                                   \
  do {                                            \
    if ((e) != hipSuccess) return fail("lsmr: " #e); \
  } while (0)
  const int T = 256;
  auto grid = [&](long long k) { return dim3((unsigned)std::max<long long>(1, (k + T - 1) / T)); };
  // ---- the matrix must be a CSR one before anything indexes with it
  {
    int bad = 0;
    LCHK(hipMemsetAsync(cnt, 0, sizeof(int), st));
    hipLaunchKernelGGL(lsmr_check, grid(std::max<long long>(m, nnz)), dim3(T), 0, st, m, n, nnz,
                       indptr, indices, cnt);
    LCHK(hipMemcpyAsync(&bad, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
    LCHK(hipStreamSynchronize(st));
    if (bad)
      return fail("lsmr: indptr[0] != 0, row pointers not increasing or a column index outside "
                  "[0, n)");
  }
  hipLaunchKernelGGL(lsmr_row_index, dim3((unsigned)((m + 3) / 4)), dim3(T), 0, st, m, indptr,
                     rowidx);
  hipLaunchKernelGGL(lsmr_iota, grid(nnz), dim3(T), 0, st, nnz, perm);
  const int nbits = 32 - __builtin_clz((unsigned)std::max(n, 1));
  // ---- the matrix the iterations use: the caller's, or its cell-major
  // renumbering with each row's entries sorted by the new column (the row
  // pointers stay, and so does the row of every position)
  const float *aval = val;
  const int *aind = indices;
  float *xs = x;
  if (layers > 1) {
    const int nc = n / layers;
    const int mbits = 32 - __builtin_clz((unsigned)std::max(m, 1));
    unsigned long long *k_in = W.get<unsigned long long>(nnz),
                       *k_out = W.get<unsigned long long>(nnz);
    int *lperm = W.get<int>(nnz), *lind = W.get<int>(nnz);
    float *lval = W.get<float>(nnz);
    xs = W.get<float>(n);
    if (W.err) return fail("lsmr: out of device memory");
    hipLaunchKernelGGL(lsmr_layer_keys, grid(nnz), dim3(T), 0, st, nnz, nc, layers, nbits, rowidx,
                       indices, k_in);
    size_t tb = 0;
    LCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k_in, k_out, perm, lperm, (int)nnz, 0,
                                            nbits + mbits, st));
    void *tmp = W.get<char>(tb);
    if (W.err) return fail("lsmr: out of device memory");
    LCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k_in, k_out, perm, lperm, (int)nnz, 0,
                                            nbits + mbits, st));
    hipLaunchKernelGGL(lsmr_layer_take, grid(nnz), dim3(T), 0, st, nnz, (1ull << nbits) - 1ull,
                       k_out, lperm, val, lind, lval);
    aval = lval;
    aind = lind;
  }
  // ---- transpose: entries sorted by column, stable (row order kept)
  {
    size_t tb = 0;
    LCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, aind, key2, perm, perm2, (int)nnz, 0,
                                            nbits, st));
    void *tmp = W.get<char>(tb);
    if (W.err) return fail("lsmr: out of device memory");
    LCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, aind, key2, perm, perm2, (int)nnz, 0, nbits,
                                            st));
  }
  hipLaunchKernelGGL(lsmr_col_bounds, grid(nnz + 1), dim3(T), 0, st, nnz, n, key2, colptr);
  hipLaunchKernelGGL(lsmr_gather, grid(nnz), dim3(T), 0, st, nnz, perm2, aval, rowidx, cval, crow);
  LCHK(hipGetLastError());
  // ---- helpers
  auto dot = [&](int len, const float *p, const float *q, double *dst) {
    hipLaunchKernelGGL(lsmr_reduce_f, dim3(kRedBlocks), dim3(T), 0, st, len, p, q, part);
    hipLaunchKernelGGL(lsmr_reduce_final, dim3(1), dim3(T), 0, st, kRedBlocks, part, dst);
  };
  auto norm = [&](int len, const float *p, float *out) -> int {
    dot(len, p, nullptr, red);
    double s = 0.0;
    if (hipMemcpyAsync(&s, red, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return 1;
    *out = (float)std::sqrt(s);
    return 0;
  };
  auto scale = [&](int len, float a, float *p) {
    hipLaunchKernelGGL(lsmr_scale, grid(len), dim3(T), 0, st, len, a, p);
  };
  // the 16 B form needs 16-byte aligned value and index arrays (the
  // transpose's own are)
  const bool quads_rows = (((uintptr_t)aval | (uintptr_t)aind) & 15) == 0;
  const int g_rows = spmv_lanes(nnz, m, quads_rows), g_cols = spmv_lanes(nnz, n, true);
  // y = a y + A x;  x = a x + A^T y
  auto aprod1 = [&](float a, const float *xx, float *yy) {
    launch_spmv(g_rows, quads_rows, m, indptr, aind, aval, xx, a, yy, st);
  };
  // A^T gathers a vector of m floats: from LDS when it fits a block's LDS
  // (the columns then get at most 64 lanes)
  const LdsShape lds_cols = lds_shape(m);
  const int g_cols_lds = std::min(g_cols, 64);
  hipError_t lds_err = hipSuccess;
  auto aprod2 = [&](float a, const float *yy, float *xx) {
    if (lds_cols.blocks > 0) {
      const hipError_t e =
          launch_spmv_lds(g_cols_lds, lds_cols, n, m, colptr, crow, cval, yy, a, xx, st);
      if (e != hipSuccess) lds_err = e;
    } else {
      launch_spmv(g_cols, true, n, colptr, crow, cval, yy, a, xx, st);
    }
  };
  // the scalar a reduction left in red[k], read back
  auto read_red = [&](int k, double *out) -> int {
    if (hipMemcpyAsync(out, red + k, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return 1;
    return 0;
  };
  // ---- lsmrModule_csr.f90:829-866
  const bool damped = damp > 0.0f;
  LCHK(hipMemcpyAsync(u, b, sizeof(float) * (size_t)m, hipMemcpyDeviceToDevice, st));
  LCHK(hipMemsetAsync(xs, 0, sizeof(float) * (size_t)n, st));
  LCHK(hipMemsetAsync(v, 0, sizeof(float) * (size_t)n, st));
  LCHK(hipMemsetAsync(hbar, 0, sizeof(float) * (size_t)n, st));
  float alpha = 0.0f, beta = 0.0f;
  if (norm(m, u, &beta)) return fail("lsmr: norm failed");
  if (beta > 0.0f) {
    scale(m, 1.0f / beta, u);
    aprod2(0.0f, u, v);
    if (lds_err != hipSuccess) return fail("lsmr: LDS product launch failed");
    if (norm(n, v, &alpha)) return fail("lsmr: norm failed");
  }
  if (alpha > 0.0f) scale(n, 1.0f / alpha, v);
  float normAr = alpha * beta;
  int istop = 0, itn = 0;
  float normA = 0.0f, condA = 0.0f, normr = beta, normx = 0.0f;
  if (normAr != 0.0f) {
    int localPointer = 0;
    bool localOrtho = false, localFull = false;
    if (localVecs > 0) {
      localPointer = 1;
      localOrtho = true;
      LCHK(hipMemcpyAsync(lv, v, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, st));
    }
    float zetabar = alpha * beta, alphabar = alpha, rho = 1, rhobar = 1, cbar = 1, sbar = 0;
    LCHK(hipMemcpyAsync(h, v, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, st));
    float betadd = beta, betad = 0, rhodold = 1, tautildeold = 0, thetatilde = 0, zeta = 0,
          d = 0;
    float normA2 = alpha * alpha, maxrbar = 0.0f, minrbar = 1e30f;
    const float normb = beta;
    float ctol = 0.0f;
    if (conlim > 0.0f) ctol = 1.0f / conlim;
    for (;;) {
      itn = itn + 1;
      // :932-955 bidiagonalisation
      aprod1(-alpha, v, u);
      if (norm(m, u, &beta)) return fail("lsmr: norm failed");
      if (beta > 0.0f) {
        scale(m, 1.0f / beta, u);
        if (localOrtho) {  // localVEnqueue
          if (localPointer < localVecs) {
            localPointer = localPointer + 1;
          } else {
            localPointer = 1;
            localFull = true;
          }
          LCHK(hipMemcpyAsync(lv + (size_t)(localPointer - 1) * n, v, sizeof(float) * (size_t)n,
                              hipMemcpyDeviceToDevice, st));
        }
        aprod2(-beta, u, v);
        if (localOrtho) {  // localVOrtho: v -= (v . q_c) q_c in turn, then ||v||
          const int lim = localFull ? localVecs : localPointer;
          // the partials of v . q_c alternate between part and part2; each
          // pass sums the previous one's partials itself
          double *pin = part, *pout = part2;
          hipLaunchKernelGGL(lsmr_reduce_f, dim3(kRedBlocks), dim3(T), 0, st, n, v, lv, pin);
          for (int c = 1; c <= lim; ++c) {
            const float *qn = c < lim ? lv + (size_t)c * n : nullptr;
            hipLaunchKernelGGL(lsmr_sub_dot, dim3(kRedBlocks), dim3(T), 0, st, n, kRedBlocks, pin,
                               lv + (size_t)(c - 1) * n, qn, v, pout);
            std::swap(pin, pout);
          }
          hipLaunchKernelGGL(lsmr_reduce_final, dim3(1), dim3(T), 0, st, kRedBlocks, pin, red);
          double s = 0.0;
          if (read_red(0, &s)) return fail("lsmr: norm failed");
          alpha = (float)std::sqrt(s);
        } else if (norm(n, v, &alpha)) {
          return fail("lsmr: norm failed");
        }
        if (alpha > 0.0f) scale(n, 1.0f / alpha, v);
      }
      // :958-1000 plane rotations
      const float alphahat = d2norm(alphabar, damp);
      const float chat = alphabar / alphahat;
      const float shat = damp / alphahat;
      const float rhoold = rho;
      rho = d2norm(alphahat, beta);
      const float c = alphahat / rho;
      const float s = beta / rho;
      const float thetanew = s * alpha;
      alphabar = c * alpha;
      const float rhobarold = rhobar;
      const float zetaold = zeta;
      const float thetabar = sbar * rho;
      const float rhotemp = cbar * rho;
      rhobar = d2norm(cbar * rho, thetanew);
      cbar = cbar * rho / rhobar;
      sbar = thetanew / rhobar;
      zeta = cbar * zetabar;
      zetabar = -sbar * zetabar;
      // :1003-1006
      const float fh = thetabar * rho / (rhoold * rhobarold);
      const float fx = zeta / (rho * rhobar);
      const float fn = thetanew / rho;
      hipLaunchKernelGGL(lsmr_update, dim3(kRedBlocks), dim3(T), 0, st, n, fh, fx, fn, v, h, hbar,
                         xs, part);
      hipLaunchKernelGGL(lsmr_reduce_final, dim3(1), dim3(T), 0, st, kRedBlocks, part, red);
      // :1009-1043
      const float betaacute = chat * betadd;
      const float betacheck = -shat * betadd;
      const float betahat = c * betaacute;
      betadd = -s * betaacute;
      const float thetatildeold = thetatilde;
      const float rhotildeold = d2norm(rhodold, thetabar);
      const float ctildeold = rhodold / rhotildeold;
      const float stildeold = thetabar / rhotildeold;
      thetatilde = stildeold * rhobar;
      rhodold = ctildeold * rhobar;
      betad = -stildeold * betad + ctildeold * betahat;
      tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold;
      const float taud = (zeta - thetatilde * tautildeold) / rhodold;
      d = d + betacheck * betacheck;
      const float bd = betad - taud;
      normr = sqrtf(d + bd * bd + betadd * betadd);
      normA2 = normA2 + beta * beta;
      normA = sqrtf(normA2);
      normA2 = normA2 + alpha * alpha;
      maxrbar = std::max(maxrbar, rhobarold);
      if (itn > 1) minrbar = std::min(minrbar, rhobarold);
      condA = std::max(maxrbar, rhotemp) / std::min(minrbar, rhotemp);
      // :1046-1078
      normAr = fabsf(zetabar);
      {
        double s = 0.0;
        if (read_red(0, &s)) return fail("lsmr: norm failed");
        normx = (float)std::sqrt(s);
      }
      const float test1 = normr / normb;
      const float test2 = normAr / (normA * normr);
      const float test3 = 1.0f / condA;
      const float t1 = test1 / (1.0f + normA * normx / normb);
      const float rtol = btol + atol * normA * normx / normb;
      if (itn >= itnlim) istop = 7;
      if (1.0f + test3 <= 1.0f) istop = 6;
      if (1.0f + test2 <= 1.0f) istop = 5;
      if (1.0f + t1 <= 1.0f) istop = 4;
      if (test3 <= ctol) istop = 3;
      if (test2 <= atol) istop = 2;
      if (test1 <= rtol) istop = 1;
      if (istop != 0) break;
    }
  }
#undef LCHK
  if (lds_err != hipSuccess) return fail("lsmr: LDS product launch failed");
  if (layers > 1)
    hipLaunchKernelGGL(lsmr_layer_out, grid(n), dim3(T), 0, st, n, n / layers, layers, xs, x);
  if (damped && istop == 2) istop = 3;
  if (hipStreamSynchronize(st) != hipSuccess) return fail("lsmr: stream failed");
  if (istop_out) *istop_out = istop;
  if (itn_out) *itn_out = itn;
  if (normA_out) *normA_out = normA;
  if (condA_out) *condA_out = condA;
  if (normr_out) *normr_out = normr;
  if (normAr_out) *normAr_out = normAr;
  if (normx_out) *normx_out = normx;
  return 0;
}

extern "C" int jdsg_lsmr_csr(int m, int n, const float *val, const int *indices,
                             const int *indptr, const float *b, float damp, float atol,
                             float btol, float conlim, int itnlim, int localSize, float *x,
                             int *istop, int *itn, float *normA, float *condA, float *normr,
                             float *normAr, float *normx, void *stream) {
  return lsmr_impl(1, m, n, val, indices, indptr, b, damp, atol, btol, conlim, itnlim, localSize,
                   x, istop, itn, normA, condA, normr, normAr, normx, stream);
}

extern "C" int jdsg_lsmr_csr_layers(int m, int n, const float *val, const int *indices,
                                    const int *indptr, const float *b, float damp, float atol,
                                    float btol, float conlim, int itnlim, int localSize,
                                    float *x, int *istop, int *itn, float *normA, float *condA,
                                    float *normr, float *normAr, float *normx, int layers,
                                    void *stream) {
  return lsmr_impl(layers, m, n, val, indices, indptr, b, damp, atol, btol, conlim, itnlim,
                   localSize, x, istop, itn, normA, condA, normr, normAr, normx, stream);
}
