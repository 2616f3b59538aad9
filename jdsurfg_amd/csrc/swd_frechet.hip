// swd_frechet.hip -- Fréchet rows and the data-misfit gradient of the
// surface-wave travel times on CDNA4 (the GPU form of SurfTime::frechet_matrix
// and SurfTime::compute_grad, src/SWD/SurfaceWave.cpp:554-676 and 695-781).
//
// A data row is one source-receiver ray.  Its 2-D sensitivity fdm(ix, iy) to
// the phase or group velocity map (from the eikonal ray tracer) times the
// column's 1-D depth kernel gives the 3-D row over the interior model cells
//   c = iz (nx-2)(ny-2) + (iy-1)(nx-2) + (ix-1),   1 <= ix <= nx-2,
//   1 <= iy <= ny-2, 0 <= iz < nz-1,
//   value = float(fdm(ix, iy) * kernel(ix + iy nx, iz, idx))   (:613-619)
// and the row keeps its nonzero values in increasing c (:621-641).  The
// gradient is sum_rows value * (t - t_obs) (:754-759), then times vs at the
// cell (:767-773).
//
// The 1-D kernels stay in HBM where jdsg_surftime_kernel left them; a ray
// comes in sparse form (its grid cells n = ix + iy nx in increasing order and
// their weights), since fdm is zero away from the ray path.  One workgroup
// per row walks the row's entries iz-major in chunks of its width and
// compacts the nonzero ones in order with a block prefix sum.
#include "swd_kernels.hpp"

namespace {

constexpr int kBlock = 256;

// entry j of a row with ncell ray cells: depth iz = j / ncell, cell j % ncell
__device__ __forceinline__ bool entry_value(const RowCtx &R, const double *kernel, int idx,
                                            const int *cell, const float *w, int ncell, long long j,
                                            float *v, int *c) {
  const int iz = (int)(j / ncell), q = (int)(j % ncell);
  const int n = cell[q];
  const int ix = n % R.nx, iy = n / R.nx;
  if (ix < 1 || ix > R.nx - 2 || iy < 1 || iy > R.ny - 2) return false;
  const double k = kernel[(long long)n + (long long)iz * R.nxy + (long long)idx * R.kstride];
  *v = (float)((double)w[q] * k);
  *c = iz * R.nin + (iy - 1) * (R.nx - 2) + (ix - 1);
  return fabsf(*v) > 0.0f;   // NaN fails this test, as std::abs(tmp) > 0.0 does
}

// inclusive block sum of one flag per thread; returns the count of set flags
// below this thread and writes the block total
__device__ __forceinline__ int block_prefix(int flag, int *wsum, int *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(flag);
  const int below = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[wave] = __popcll(m);
  __syncthreads();
  int off = 0, tot = 0;
  for (int i = 0; i < kBlock / 64; ++i) {
    if (i < wave) off += wsum[i];
    tot += wsum[i];
  }
  *total = tot;
  return off + below;
}

}  // namespace

// Count the nonzero entries of each row into row_off[r + 1]; -1 marks a row
// whose period index is out of range or whose cells are not increasing grid
// indices (nothing of such a row is read from the kernel array).
extern "C" __global__ void __launch_bounds__(kBlock)
jdsg_frechet_count(RowCtx R, const double *kernel, int nrows, const int *row_idx,
                   const long long *ray_off, const int *ray_cell, const float *ray_w,
                   long long *row_off) {
  __shared__ int wsum[kBlock / 64];
  __shared__ int bad;
  const int r = blockIdx.x;
  if (r >= nrows) return;
  const long long c0 = ray_off[r];
  const int ncell = (int)(ray_off[r + 1] - c0);
  if (threadIdx.x == 0) bad = (row_idx[r] < 0 || row_idx[r] >= R.kmax || ncell < 0) ? 1 : 0;
  __syncthreads();
  const int *cell = ray_cell + c0;
  const float *w = ray_w + c0;
  for (int q = threadIdx.x; q < ncell; q += kBlock) {
    const int n = cell[q];
    if (n < 0 || n >= R.nxy || (q > 0 && cell[q - 1] >= n)) bad = 1;
  }
  __syncthreads();
  if (bad) {
    if (threadIdx.x == 0) row_off[r + 1] = -1;
    return;
  }
  const long long nent = (long long)ncell * (R.nz - 1);
  long long cnt = 0;
  for (long long j0 = 0; j0 < nent; j0 += kBlock) {
    const long long j = j0 + threadIdx.x;
    float v;
    int c;
    const int flag = (j < nent) && entry_value(R, kernel, row_idx[r], cell, w, ncell, j, &v, &c);
    int tot;
    (void)block_prefix(flag, wsum, &tot);
    cnt += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) row_off[r + 1] = cnt;
}

// In-place exclusive scan of the counts in row_off[1..nrows] (one block):
// row_off[r] becomes the start of row r, row_off[nrows] the total; row_off[0]
// is 0, or -1 when a row was marked invalid.
extern "C" __global__ void __launch_bounds__(1024)
jdsg_frechet_scan(int nrows, long long *row_off) {
  __shared__ long long part[1024];
  __shared__ int bad;
  const int t = threadIdx.x;
  if (t == 0) bad = 0;
  __syncthreads();
  const long long per = ((long long)nrows + 1023) / 1024;
  const long long a = t * per, b = (a + per < nrows) ? a + per : nrows;
  long long s = 0;
  for (long long r = a; r < b; ++r) {
    const long long x = row_off[r + 1];
    if (x < 0) bad = 1;
    s += x;
  }
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    long long run = 0;
    for (int i = 0; i < 1024; ++i) {
      const long long x = part[i];
      part[i] = run;
      run += x;
    }
  }
  __syncthreads();
  long long run = part[t];
  for (long long r = a; r < b; ++r) {
    const long long x = row_off[r + 1];
    row_off[r + 1] = run + x;
    run += x;
  }
  __syncthreads();
  if (t == 0) row_off[0] = bad ? -1 : 0;
}

// Ordered fill of the rows (after the scan).  Rows past `cap` entries are not
// written (the caller sees the total and the error).
extern "C" __global__ void __launch_bounds__(kBlock)
jdsg_frechet_fill(RowCtx R, const double *kernel, int nrows, const int *row_idx,
                  const long long *ray_off, const int *ray_cell, const float *ray_w,
                  const long long *row_off, int *col, float *val, long long cap) {
  __shared__ int wsum[kBlock / 64];
  const int r = blockIdx.x;
  if (r >= nrows) return;
  if (row_off[0] < 0) return;  // an invalid row: the offsets are not usable
  const long long base = row_off[r];
  if (row_off[r + 1] > cap) return;
  const long long c0 = ray_off[r];
  const int ncell = (int)(ray_off[r + 1] - c0);
  const int *cell = ray_cell + c0;
  const float *w = ray_w + c0;
  const long long nent = (long long)ncell * (R.nz - 1);
  long long run = base;
  for (long long j0 = 0; j0 < nent; j0 += kBlock) {
    const long long j = j0 + threadIdx.x;
    float v = 0.0f;
    int c = 0;
    const int flag = (j < nent) && entry_value(R, kernel, row_idx[r], cell, w, ncell, j, &v, &c);
    int tot;
    const int pos = block_prefix(flag, wsum, &tot);
    if (flag) {
      col[run + pos] = c;
      val[run + pos] = v;
    }
    run += tot;
    __syncthreads();
  }
}

// Misfit gradient: grad[c] += value * resid[r] for every entry of every row
// (float atomics; the reference's sum order depends on its thread count too).
extern "C" __global__ void __launch_bounds__(kBlock)
jdsg_frechet_gradrows(RowCtx R, const double *kernel, int nrows, const int *row_idx,
                      const long long *ray_off, const int *ray_cell, const float *ray_w,
                      const float *resid, float *grad) {
  const int r = blockIdx.x;
  if (r >= nrows) return;
  if (row_idx[r] < 0 || row_idx[r] >= R.kmax) return;
  const long long c0 = ray_off[r];
  const int ncell = (int)(ray_off[r + 1] - c0);
  const int *cell = ray_cell + c0;
  const float *w = ray_w + c0;
  const long long nent = (long long)ncell * (R.nz - 1);
  const float res = resid[r];
  for (long long j = threadIdx.x; j < nent; j += kBlock) {
    const int q = (int)(j % ncell);
    const int n = cell[q];
    if (n < 0 || n >= R.nxy) continue;
    float v;
    int c;
    if (entry_value(R, kernel, row_idx[r], cell, w, ncell, j, &v, &c)) atomicAdd(&grad[c], v * res);
  }
}

// grad[c] = 0 (before the rows) and grad[c] *= vs(ix, iy, iz) (after them)
extern "C" __global__ void jdsg_frechet_gradfix(RowCtx R, const float *vs, float *grad, int zero) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)R.nin * (R.nz - 1);
  if (c >= n) return;
  if (zero) {
    grad[c] = 0.0f;
    return;
  }
  const int iz = (int)(c / R.nin), rem = (int)(c % R.nin);
  const int iy = rem / (R.nx - 2) + 1, ix = rem % (R.nx - 2) + 1;
  grad[c] *= vs[(long long)ix + (long long)iy * R.nx + (long long)iz * R.nxy];
}

// ----------------------------------------------------------------------------
void jdsg_launch_frechet_rows(const RowCtx &R, const double *kernel, int nrows, const int *row_idx,
                              const long long *ray_off, const int *ray_cell, const float *ray_w,
                              long long *row_off, int *col, float *val, long long cap,
                              hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(jdsg_frechet_count, dim3(nrows), dim3(kBlock), 0, st, R, kernel, nrows,
                     row_idx, ray_off, ray_cell, ray_w, row_off);
  hipLaunchKernelGGL(jdsg_frechet_scan, dim3(1), dim3(1024), 0, st, nrows, row_off);
  hipLaunchKernelGGL(jdsg_frechet_fill, dim3(nrows), dim3(kBlock), 0, st, R, kernel, nrows,
                     row_idx, ray_off, ray_cell, ray_w, row_off, col, val, cap);
}

void jdsg_launch_frechet_grad(const RowCtx &R, const double *kernel, int nrows,
                              const int *row_idx, const long long *ray_off, const int *ray_cell,
                              const float *ray_w, const float *resid, const float *vs,
                              float *grad, hipStream_t st) {
  const long long n = (long long)R.nin * (R.nz - 1);
  const unsigned g = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(jdsg_frechet_gradfix, dim3(g), dim3(256), 0, st, R, vs, grad, 1);
  if (nrows > 0)
    hipLaunchKernelGGL(jdsg_frechet_gradrows, dim3(nrows), dim3(kBlock), 0, st, R, kernel, nrows,
                       row_idx, ray_off, ray_cell, ray_w, resid, grad);
  hipLaunchKernelGGL(jdsg_frechet_gradfix, dim3(g), dim3(256), 0, st, R, vs, grad, 0);
}
