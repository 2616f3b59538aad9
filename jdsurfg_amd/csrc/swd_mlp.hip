// swd_mlp.hip -- BF16 MLP kernels for the learned root-bracket surrogate.
//
// GEMMs run on CDNA matrix cores (v_mfma_f32_16x16x16_bf16: bf16 operands,
// fp32 accumulation) from 64x64x32 LDS tiles, one workgroup of four waves per
// output tile, each wave a 32x32 block of 2x2 MFMA tiles.  Operands may be
// read transposed so one kernel serves the forward X W^T, the input gradient
// dZ W and the weight gradient dZ^T X.  Epilogues fuse the bias and the GELU.
// Dimensions are padded by the caller: M, N multiples of 64, K of 32.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// two floats to packed bf16 (round to nearest even) in one v_cvt_pk_bf16_f32
__device__ __forceinline__ uint32_t pack_bf2(float a, float b) {
  __bf16 ha = (__bf16)a, hb = (__bf16)b;
  return (uint32_t)__builtin_bit_cast(uint16_t, ha) |
         ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
}

// GELU in its tanh form, 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))),
// with tanh(u) = 1 - 2 / (1 + 2^(2u log2 e)) on the hardware exp2 and
// reciprocal: a handful of instructions per element, so the epilogue of a
// GEMM tile costs less than its MFMA loop (the erf form cost more).  It is
// within 5e-4 of the erf form everywhere, below bf16 resolution.
constexpr float kG0 = 0.7978845608028654f, kG1 = 0.044715f;
__device__ __forceinline__ float tanh_fast(float u) {
  float t = __builtin_amdgcn_exp2f(2.8853900817779268f * u);  // exp(2u)
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(t + 1.0f);
}
// The same GELU written as z sigmoid(2u) = z / (1 + 2^(-2u log2 e)), with
// 2u log2 e = z (kS0 + kS1 z^2): one exp2 and one reciprocal in about seven
// instructions (the epilogues of the inference kernels are VALU-bound next
// to their MFMA work).  Equal to the tanh form up to fp32 rounding.
constexpr float kS0 = 2.302208198144325f, kS1 = 0.1029432395800235f;
__device__ __forceinline__ float gelu(float z) {
  const float t = z * fmaf(kS1, z * z, kS0);
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-t));
}
__device__ __forceinline__ float gelu_grad(float z) {
  float u = kG0 * (z + kG1 * z * z * z);
  float th = tanh_fast(u);
  return 0.5f * (1.0f + th) + 0.5f * z * (1.0f - th * th) * kG0 * (1.0f + 3.0f * kG1 * z * z);
}

constexpr int BM = 64, BN = 64, BK = 32, LDSK = BK + 4;

}  // namespace

// C(i,j) = sum_k A(i,k) B(k,j);  A(i,k) = TA ? A[k*lda+i] : A[i*lda+k];
// B(k,j) = TB ? B[j*ldb+k] : B[k*ldb+j].
// epi 0: C fp32;  1: Z = C + bias (fp32) and Y = gelu(Z) (bf16);
// epi 2: C + bias fp32;  3: C fp32 accumulated into C (+=);
// epi 4: Y = gelu(C + bias) bf16 only (inference).
template <bool TA, bool TB>
__global__ void __launch_bounds__(256)
jdsg_gemm_bf16(int M, int N, int K, const uint16_t *__restrict__ A, int lda,
               const uint16_t *__restrict__ B, int ldb, float *C, int ldc, int epi,
               const float *__restrict__ bias, uint16_t *Y, int ldy) {
  __shared__ uint16_t As[BM][LDSK];
  __shared__ uint16_t Bs[BN][LDSK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int bm = blockIdx.y * BM, bn = blockIdx.x * BN;
  v4f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += BK) {
    // stage 64x32 tiles of A and B^T (k contiguous) through LDS
#pragma unroll
    for (int r = 0; r < (BM * BK) / 256; ++r) {
      int e = r * 256 + tid;
      int i, k;
      if (!TA) { i = e / BK; k = e % BK; } else { k = e / BM; i = e % BM; }
      size_t ga = TA ? (size_t)(k0 + k) * lda + (bm + i) : (size_t)(bm + i) * lda + (k0 + k);
      As[i][k] = A[ga];
      int j, kb;
      if (TB) { j = e / BK; kb = e % BK; } else { kb = e / BN; j = e % BN; }
      size_t gb = TB ? (size_t)(bn + j) * ldb + (k0 + kb) : (size_t)(k0 + kb) * ldb + (bn + j);
      Bs[j][kb] = B[gb];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      v4s af[2], bfr[2];
      const int kq = kk + 4 * (lane >> 4);
#pragma unroll
      for (int mi = 0; mi < 2; ++mi) {
        const uint16_t *p = &As[wm * 32 + mi * 16 + (lane & 15)][kq];
        af[mi] = v4s{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
      }
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) {
        const uint16_t *p = &Bs[wn * 32 + ni * 16 + (lane & 15)][kq];
        bfr[ni] = v4s{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
      }
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[mi], bfr[ni], acc[mi][ni],
                                                                 0, 0, 0);
    }
    __syncthreads();
  }
  // D(i, j): lane holds rows 4*(lane/16)+r, column lane%16 of each 16x16 tile
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int i = bm + wm * 32 + mi * 16 + 4 * (lane >> 4) + r;
        int j = bn + wn * 32 + ni * 16 + (lane & 15);
        if (i >= M || j >= N) continue;
        float v = acc[mi][ni][r];
        size_t o = (size_t)i * ldc + j;
        if (epi == 0) {
          C[o] = v;
        } else if (epi == 1) {
          float z = v + bias[j];
          C[o] = z;
          Y[(size_t)i * ldy + j] = f2bf(gelu(z));
        } else if (epi == 2) {
          C[o] = v + bias[j];
        } else if (epi == 4) {
          Y[(size_t)i * ldy + j] = f2bf(gelu(v + bias[j]));
        } else {
          C[o] += v;
        }
      }
}

// Large-tile form for M, N multiples of 128 (every hidden layer) and K of
// 64: 128x128 output tiles, four waves each owning a 64x64 block of 4x4 MFMA tiles on the
// gfx950 K=32 instruction (v_mfma_f32_16x16x32_bf16), 16-byte global loads
// with the next 64-wide K step held in registers while the current one
// computes.  A
// lane's A fragment is 8 consecutive k of one row and its B fragment the same
// 8 k of one column, so the product is the same whatever order the hardware
// sums k in.  Workgroups are numbered so that consecutive tiles of one M row
// band run on one XCD and share its L2.
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
constexpr int BM2 = 128, BN2 = 128, BK2 = 64, LDK2 = BK2 + 8;

// 16-byte chunk c (0..1023) of a 128x64 operand tile.  Row-major along k
// (KMAJOR false: element (r, k) at P[r*ld + k]) or stored k-major (element
// (r, k) at P[k*ld + r]), the latter scattered into the k-contiguous LDS tile.
template <bool KMAJOR>
__device__ __forceinline__ uint4 gload_tile(const uint16_t *__restrict__ P, int ld, int r0,
                                            int k0, int c) {
  if (!KMAJOR) {
    int r = c >> 3, kc = (c & 7) * 8;
    return *reinterpret_cast<const uint4 *>(P + (size_t)(r0 + r) * ld + k0 + kc);
  }
  int k = c >> 4, rc = (c & 15) * 8;
  return *reinterpret_cast<const uint4 *>(P + (size_t)(k0 + k) * ld + r0 + rc);
}
template <bool KMAJOR>
__device__ __forceinline__ void lstore_tile(uint16_t (*S)[BK2 + 8], uint4 v, int c) {
  if (!KMAJOR) {
    int r = c >> 3, kc = (c & 7) * 8;
    *reinterpret_cast<uint4 *>(&S[r][kc]) = v;
  } else {
    int k = c >> 4, rc = (c & 15) * 8;
    S[rc + 0][k] = (uint16_t)(v.x & 0xFFFFu);
    S[rc + 1][k] = (uint16_t)(v.x >> 16);
    S[rc + 2][k] = (uint16_t)(v.y & 0xFFFFu);
    S[rc + 3][k] = (uint16_t)(v.y >> 16);
    S[rc + 4][k] = (uint16_t)(v.z & 0xFFFFu);
    S[rc + 5][k] = (uint16_t)(v.z >> 16);
    S[rc + 6][k] = (uint16_t)(v.w & 0xFFFFu);
    S[rc + 7][k] = (uint16_t)(v.w >> 16);
  }
}

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
jdsg_gemm_bf16_128(int M, int N, int K, const uint16_t *__restrict__ A, int lda,
                   const uint16_t *__restrict__ B, int ldb, float *C, int ldc, int epi,
                   const float *__restrict__ bias, uint16_t *Y, int ldy) {
  __shared__ __attribute__((aligned(16))) uint16_t As[BM2][LDK2];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN2][LDK2];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int nN = N / BN2, ntile = nN * (M / BM2);
  int t = blockIdx.x;
  if ((ntile & 7) == 0) t = (t & 7) * (ntile >> 3) + (t >> 3);
  const int bm = (t / nN) * BM2, bn = (t % nN) * BN2;
  v4f acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  uint4 ra[4], rb[4];
#define JDSG_GLOAD(k0)                                                                      \
  do {                                                                                      \
    _Pragma("unroll") for (int r = 0; r < 4; ++r) {                                         \
      ra[r] = gload_tile<TA>(A, lda, bm, (k0), r * 256 + tid);                              \
      rb[r] = gload_tile<!TB>(B, ldb, bn, (k0), r * 256 + tid);                             \
    }                                                                                       \
  } while (0)
  JDSG_GLOAD(0);
  for (int k0 = 0; k0 < K; k0 += BK2) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      lstore_tile<TA>(As, ra[r], r * 256 + tid);
      lstore_tile<!TB>(Bs, rb[r], r * 256 + tid);
    }
    __syncthreads();
    if (k0 + BK2 < K) JDSG_GLOAD(k0 + BK2);
#undef JDSG_GLOAD
#pragma unroll
    for (int kk = 0; kk < BK2; kk += 32) {
      bf8v af[4], bfv[4];
      const int kq = kk + 8 * (lane >> 4);
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
        af[mi] = *reinterpret_cast<const bf8v *>(&As[wm * 64 + mi * 16 + (lane & 15)][kq]);
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
        bfv[ni] = *reinterpret_cast<const bf8v *>(&Bs[wn * 64 + ni * 16 + (lane & 15)][kq]);
      // operands swapped (W^T X^T): a lane's accumulator holds output
      // (row lane%16, columns 4 (lane/16) .. +3), four consecutive columns
      // of one row, so the epilogue stores them as one vector
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfv[ni], af[mi], acc[mi][ni], 0,
                                                                0, 0);
    }
    __syncthreads();
  }
  // epilogue: per lane 4 rows (mi) x 4 column quads (ni); the bias quad is
  // loaded once per ni
#pragma unroll
  for (int ni = 0; ni < 4; ++ni) {
    const int j0 = bn + wn * 64 + ni * 16 + 4 * (lane >> 4);
    v4f bq = v4f{0.f, 0.f, 0.f, 0.f};
    if (epi == 1 || epi == 2 || epi == 4) bq = *reinterpret_cast<const v4f *>(bias + j0);
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      const int i = bm + wm * 64 + mi * 16 + (lane & 15);
      v4f v = acc[mi][ni];
      v4f *cp = reinterpret_cast<v4f *>(C + (size_t)i * ldc + j0);
      if (epi == 0) {
        *cp = v;
      } else if (epi == 3) {
        *cp = *cp + v;
      } else {
        v = v + bq;
        if (epi != 4) *cp = v;
        if (epi != 2) {
          uint2 h;
          h.x = pack_bf2(gelu(v[0]), gelu(v[1]));
          h.y = pack_bf2(gelu(v[2]), gelu(v[3]));
          *reinterpret_cast<uint2 *>(Y + (size_t)i * ldy + j0) = h;
        }
      }
    }
  }
}

// dZ = dY * gelu'(Z) -> bf16 (M x N row-major, ld = N)
__global__ void jdsg_gelu_bwd(long long n, const float *__restrict__ dY,
                              const float *__restrict__ Z, uint16_t *dZ, float *dZf) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float g = dY[i] * gelu_grad(Z[i]);
  dZ[i] = f2bf(g);
  if (dZf) dZf[i] = g;
}

// column sums of an M x N fp32 matrix (bias gradient): one block per column
// strip of 64, 256 threads reduce over rows
__global__ void __launch_bounds__(256)
jdsg_colsum(int M, int N, const float *__restrict__ X, float *out) {
  __shared__ float part[4][64];
  int j = blockIdx.x * 64 + (threadIdx.x & 63);
  int r0 = threadIdx.x >> 6;
  float s = 0.f;
  if (j < N)
    for (int i = r0; i < M; i += 4) s += X[(size_t)i * N + j];
  part[r0][threadIdx.x & 63] = s;
  __syncthreads();
  if (r0 == 0 && j < N) out[j] = part[0][threadIdx.x] + part[1][threadIdx.x] +
                                 part[2][threadIdx.x] + part[3][threadIdx.x];
}

// Loss on column 0 of the output (ld = ldo): squared error, or with
// delta > 0 the Huber loss (quadratic inside |d| <= delta, linear outside),
// which keeps rare far-off targets from dominating the fit.  Writes dY (fp32,
// ld ldo, zero elsewhere) and per-block partial losses.
__global__ void jdsg_mse(int M, int ldo, const float *__restrict__ out,
                         const float *__restrict__ target, const float *__restrict__ wt,
                         float delta, float *dY, float *loss_part) {
  __shared__ float red[256];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < M) {
    float w = wt ? wt[i] : 1.0f;
    float d = out[(size_t)i * ldo] - target[i];
    float ad = fabsf(d);
    if (delta > 0.f && ad > delta) {
      l = w * delta * (2.0f * ad - delta);
      dY[(size_t)i * ldo] = 2.0f * w * copysignf(delta, d) / (float)M;
    } else {
      l = w * d * d;
      dY[(size_t)i * ldo] = 2.0f * w * d / (float)M;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Adam with fp32 master weights; refreshes the bf16 copy
__global__ void jdsg_adam(long long n, float *w, float *m, float *v, const float *__restrict__ g,
                          float lr, float b1, float b2, float eps, float c1, float c2, float wd,
                          uint16_t *wbf) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i] + wd * w[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wbf) wbf[i] = f2bf(wi);
}

__global__ void jdsg_f32_to_bf16(long long n, const float *__restrict__ x, uint16_t *y) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// Output head of the network at inference: only output 0 of the padded last
// layer is a prediction, so instead of a GEMM over 64 padded outputs a
// fixed-order dot product of the last hidden layer Y (bf16, row stride H) with
// the first row w of the last weight matrix, fp32 accumulation, plus bias b,
// de-standardised: pred[e0 + i] = exp((Y_i . w + b) * scale + shift).  Eight
// lanes per row, each over H/8 consecutive inputs (16-byte loads), reduced by
// shuffles in a fixed order.
// Partial dot product of the output head: n8 groups of eight bf16 inputs y
// and weights w, fp32, in order.  Shared by the head kernel and the fused
// network kernel so both sum the same way.
__device__ __forceinline__ float head_part(const uint16_t *y, const uint16_t *__restrict__ w,
                                           int n8) {
  float s = 0.f;
  for (int q = 0; q < n8; ++q) {
    const uint4 a = reinterpret_cast<const uint4 *>(y)[q];
    const uint4 c = reinterpret_cast<const uint4 *>(w)[q];
    const uint32_t av[4] = {a.x, a.y, a.z, a.w}, cv[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      s += bf2f((uint16_t)(av[i] & 0xFFFFu)) * bf2f((uint16_t)(cv[i] & 0xFFFFu));
      s += bf2f((uint16_t)(av[i] >> 16)) * bf2f((uint16_t)(cv[i] >> 16));
    }
  }
  return s;
}
// the eight parts of a row sit in eight consecutive lanes
__device__ __forceinline__ float head_reduce8(float s) {
  s += __shfl_xor(s, 1);
  s += __shfl_xor(s, 2);
  s += __shfl_xor(s, 4);
  return s;
}

__global__ void __launch_bounds__(256)
jdsg_mlp_head(long long n, int H, const uint16_t *__restrict__ Y, const uint16_t *__restrict__ w,
              float b, double scale, double shift, double *pred, long long e0) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long row = t >> 3;
  const int part = (int)(t & 7), seg = H >> 3;
  float s = 0.f;
  if (row < n) s = head_part(Y + row * H + part * seg, w + part * seg, seg / 8);
  s = head_reduce8(s);
  if (row < n && part == 0) pred[e0 + row] = exp((double)(s + b) * scale + shift);
}

// First layer of the network at inference, split by what varies.  Of the 192
// inputs of an entry only ln T, the wave and the mode one-hots depend on the
// entry; the rest are its column's profile and half-space features.  So the
// pre-activation is Acol(col) + bf16(ln T) w_144 + w_wave + w_mode + b, with
// Acol = (column feature row) W^T from one GEMM over the columns (bf16
// operands, fp32 accumulation, like the full GEMM up to summation order).
// Row of column features: [0,144) profile, 151-154 half-space, rest zero.
__global__ void jdsg_mlp_column_rows(int ncol, const float *__restrict__ colf, uint16_t *X) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const float *cf = colf + (size_t)col * 148;
  uint4 *dst = reinterpret_cast<uint4 *>(X + (size_t)col * 192);
  for (int g = 0; g < 24; ++g) {
    float v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int f = g * 8 + q;
      v[q] = f < 144 ? cf[f] : (f >= 151 && f < 155) ? cf[144 + (f - 151)] : 0.f;
    }
    dst[g] = make_uint4(pack_bf2(v[0], v[1]), pack_bf2(v[2], v[3]), pack_bf2(v[4], v[5]),
                        pack_bf2(v[6], v[7]));
  }
}

// Column of entry e, searched in [lo, hi] (the entries of column c are
// ent_off[c] .. ent_off[c+1]-1).
__device__ __forceinline__ int entry_column(long long e, int lo, int hi,
                                            const long long *__restrict__ ent_off) {
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (ent_off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// What the first layer needs of entry e: its column (in [lo, hi]), wave,
// mode and the bf16-rounded ln T input (the entry order of the plans: per
// column, waves R then L, modes, periods fastest).
struct EntryIn {
  int col, wave, mode;
  float lt;
};
__device__ __forceinline__ EntryIn entry_inputs(long long e, int lo, int hi,
                                                const long long *__restrict__ ent_off,
                                                const int *__restrict__ per_off,
                                                const double *__restrict__ periods,
                                                const int *__restrict__ nmode) {
  lo = entry_column(e, lo, hi, ent_off);
  EntryIn r;
  r.col = lo;
  const int K = per_off[lo + 1] - per_off[lo];
  const long long el = e - ent_off[lo];
  const int wm = (int)(el / K), k = (int)(el % K);
  const int nR = nmode[2 * lo];
  r.wave = wm < nR ? 0 : 1;
  r.mode = r.wave == 0 ? wm : wm - nR;
  r.lt = bf2f(f2bf((float)((log(periods[per_off[lo] + k]) - 2.0) / 1.5)));
  return r;
}
// Four first-layer outputs j..j+3 of an entry as packed bf16 from its
// column's pre-activations a = acol(col, j..j+3):
// gelu(a + lt ext[0] + ext[1 + wave] + ext[3 + mode] + ext[7]).  ext is the
// LDS copy of the eight input rows with a ninth row of zeros, the mode term
// of modes above 3 (they have no one-hot input); j is a multiple of 4, so
// every row is read as one 16-byte vector.
__device__ __forceinline__ uint2 first_layer4(const EntryIn &in, const float4 a,
                                              const float *ext, int H, int j) {
  const int mrow = in.mode < 4 ? 3 + in.mode : 8;
  const float4 e0 = *reinterpret_cast<const float4 *>(ext + j);
  const float4 ew = *reinterpret_cast<const float4 *>(ext + (1 + in.wave) * H + j);
  const float4 em = *reinterpret_cast<const float4 *>(ext + mrow * H + j);
  const float4 e7 = *reinterpret_cast<const float4 *>(ext + 7 * H + j);
  uint2 r;
  r.x = pack_bf2(gelu(a.x + in.lt * e0.x + ew.x + em.x + e7.x),
                 gelu(a.y + in.lt * e0.y + ew.y + em.y + e7.y));
  r.y = pack_bf2(gelu(a.z + in.lt * e0.z + ew.z + em.z + e7.z),
                 gelu(a.w + in.lt * e0.w + ew.w + em.w + e7.w));
  return r;
}
// ext (8 x H, global) into LDS rows 0..7 and zeros into row 8, by the
// workgroup's nt threads.
__device__ __forceinline__ void stage_ext(float *sx, const float *__restrict__ ext, int H, int tid,
                                          int nt) {
  const int n4 = 2 * H;  // float4s in eight rows
  for (int q = tid; q < n4; q += nt)
    reinterpret_cast<float4 *>(sx)[q] = reinterpret_cast<const float4 *>(ext)[q];
  for (int q = tid; q < H / 4; q += nt)
    reinterpret_cast<float4 *>(sx)[n4 + q] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Y(row, j) = first layer of entry e0 + row (row < nrow).  A workgroup takes
// a tile of kExpandRows rows: it stages ext in LDS (dynamic, 36 H bytes),
// two of its threads find the columns of the tile's first and last rows
// (binary searches over the columns, in two waves at once), one thread per
// row then forms that row's inputs (column, wave, mode, ln T; its column
// searched only between the tile's two, since the entries of a column are
// contiguous) into LDS, and a thread then writes features 4g..4g+3 and
// H/2+4g..H/2+4g+3 of a row per step, so that the lanes of a wave read and
// write contiguous 16- and 8-byte pieces.
constexpr int kExpandRows = 256;
__global__ void __launch_bounds__(256)
jdsg_mlp_expand(long long e0, long long nrow, int ncol, const long long *__restrict__ ent_off,
                const int *__restrict__ per_off, const double *__restrict__ periods,
                const int *__restrict__ nmode, const float *__restrict__ acol, int lda,
                const float *__restrict__ ext, int H, uint16_t *Y) {
  extern __shared__ __attribute__((aligned(16))) float sx[];
  __shared__ int s_lo, s_hi;
  __shared__ EntryIn s_in[kExpandRows];
  const int tid = threadIdx.x;
  const long long r0 = (long long)blockIdx.x * kExpandRows;
  const long long r1 = (r0 + kExpandRows < nrow ? r0 + kExpandRows : nrow) - 1;
  if (tid == 0) s_lo = entry_column(e0 + r0, 0, ncol - 1, ent_off);
  if (tid == 64) s_hi = entry_column(e0 + r1, 0, ncol - 1, ent_off);
  stage_ext(sx, ext, H, tid, blockDim.x);
  __syncthreads();
  const int nr = (int)(r1 - r0 + 1);
  if (tid < nr)
    s_in[tid] = entry_inputs(e0 + r0 + tid, s_lo, s_hi, ent_off, per_off, periods, nmode);
  __syncthreads();
  const int groups = H >> 3, half = H >> 1;
  const int ntask = nr * groups;
  for (int task = tid; task < ntask; task += blockDim.x) {
    const int r = task / groups;
    const int j = (task - r * groups) * 4;
    const EntryIn in = s_in[r];
    const float *ac = acol + (size_t)in.col * lda;
    const float4 a0 = *reinterpret_cast<const float4 *>(ac + j);
    const float4 a1 = *reinterpret_cast<const float4 *>(ac + half + j);
    uint16_t *y = Y + (r0 + r) * H;
    *reinterpret_cast<uint2 *>(y + j) = first_layer4(in, a0, sx, H, j);
    *reinterpret_cast<uint2 *>(y + half + j) = first_layer4(in, a1, sx, H, half + j);
  }
}

// The whole network at inference for 128 entries per workgroup, hidden width
// 256: first layer from acol (as jdsg_mlp_expand), three hidden layers on
// MFMA and the output head (as jdsg_mlp_head), with the activations kept in
// LDS between layers instead of going through HBM.  Wave w computes
// features 64w..64w+63 of all 128 rows: acc[ft][rt] is the 16x16 tile of
// features 64w+16ft.. and rows 16rt..; the weight fragments (16 features x
// 32 inputs) stream from L2, one K step ahead; the activation fragments are
// LDS rows (stride 264 bf16 so the sixteen rows of a fragment fall in
// different banks).  The MFMA operands, their K order and the epilogue are
// those of jdsg_gemm_bf16_128's inference epilogue, and the first layer and
// the head share their code with the unfused kernels, so the predictions are
// the same bits as the unfused path's.  tcol[t] is the column of the first
// entry of tile t (tcol[ntile]: of the last entry), so a row searches its
// column among the few columns its tile spans.
constexpr int FH = 256, FLD = FH + 8;
constexpr int kFusedRows = 128;  // rows per workgroup (two workgroups per CU)

template <int FR>
__global__ void jdsg_mlp_tile_cols(long long e0, long long nrow, int ncol,
                                   const long long *__restrict__ ent_off, int ntile, int *tcol) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntile) return;
  long long r = (long long)t * FR;
  if (r > nrow - 1) r = nrow - 1;
  tcol[t] = entry_column(e0 + r, 0, ncol - 1, ent_off);
}

template <int FR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FR == 128 ? 2 : 3)))
jdsg_mlp_fused(long long e0, long long nrow, int ncol, const long long *__restrict__ ent_off,
               const int *__restrict__ per_off, const double *__restrict__ periods,
               const int *__restrict__ nmode, const float *__restrict__ acol, int lda,
               const float *__restrict__ ext, const uint16_t *__restrict__ W2,
               const uint16_t *__restrict__ W3, const uint16_t *__restrict__ W4,
               const float *__restrict__ b2, const float *__restrict__ b3,
               const float *__restrict__ b4, const uint16_t *__restrict__ w5, float b5,
               double scale, double shift, double *pred, long long p0,
               const int *__restrict__ tcol) {
  __shared__ __attribute__((aligned(16))) uint16_t X[FR][FLD];
  __shared__ __attribute__((aligned(16))) float sext[9 * FH];
  __shared__ EntryIn sin_[FR];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long long r0 = (long long)blockIdx.x * FR;
  stage_ext(sext, ext, FH, tid, 256);
  if (tid < FR) {
    long long row = r0 + tid;
    if (row >= nrow) row = nrow - 1;  // tail rows repeat the last entry, not stored
    sin_[tid] = entry_inputs(e0 + row, tcol[blockIdx.x], tcol[blockIdx.x + 1], ent_off, per_off,
                             periods, nmode);
  }
  // weight fragments two K steps ahead; step s is layer s/8, K step s%8
  const uint16_t *const Wl[3] = {W2, W3, W4};
  const float *const bl[3] = {b2, b3, b4};
  const size_t wofs = (size_t)(64 * wave + (lane & 15)) * FH + 8 * (lane >> 4);
  const int xr = lane & 15, xk = 8 * (lane >> 4);
  bf8v w0[4], w1[4];
#pragma unroll
  for (int ft = 0; ft < 4; ++ft) {
    w0[ft] = *reinterpret_cast<const bf8v *>(W2 + wofs + (size_t)ft * 16 * FH);
    w1[ft] = *reinterpret_cast<const bf8v *>(W2 + wofs + (size_t)ft * 16 * FH + 32);
  }
  __syncthreads();
  {
    // a thread keeps features j..j+3 and FH/2+j..FH/2+j+3 and walks rows
    // tid/32 + 8i; rows of one column share acol, loaded again only when
    // the column changes
    const int j = (tid & 31) * 4;
    int pc = -1;
    float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0;
    for (int i = 0; i < FR / 8; ++i) {
      const int row = i * 8 + (tid >> 5);
      const EntryIn in = sin_[row];
      if (in.col != pc) {
        const float *ac = acol + (size_t)in.col * lda;
        a0 = *reinterpret_cast<const float4 *>(ac + j);
        a1 = *reinterpret_cast<const float4 *>(ac + FH / 2 + j);
        pc = in.col;
      }
      *reinterpret_cast<uint2 *>(&X[row][j]) = first_layer4(in, a0, sext, FH, j);
      *reinterpret_cast<uint2 *>(&X[row][FH / 2 + j]) = first_layer4(in, a1, sext, FH, FH / 2 + j);
    }
  }
  __syncthreads();

  for (int L = 0; L < 3; ++L) {
    constexpr int RT = FR / 16;
    v4f acc[4][RT];
#pragma unroll
    for (int ft = 0; ft < 4; ++ft)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) acc[ft][rt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int ks = 0; ks < 8; ++ks) {
      const int s2 = L * 8 + ks + 2;
      bf8v w2[4];
      if (s2 < 24) {
        const uint16_t *np = Wl[s2 >> 3] + wofs + (s2 & 7) * 32;
#pragma unroll
        for (int ft = 0; ft < 4; ++ft)
          w2[ft] = *reinterpret_cast<const bf8v *>(np + (size_t)ft * 16 * FH);
      }
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        const bf8v xf = *reinterpret_cast<const bf8v *>(&X[rt * 16 + xr][ks * 32 + xk]);
#pragma unroll
        for (int ft = 0; ft < 4; ++ft)
          acc[ft][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0[ft], xf, acc[ft][rt], 0, 0, 0);
      }
#pragma unroll
      for (int ft = 0; ft < 4; ++ft) {
        w0[ft] = w1[ft];
        if (s2 < 24) w1[ft] = w2[ft];
      }
    }
    __syncthreads();  // every wave has read this layer's inputs
#pragma unroll
    for (int ft = 0; ft < 4; ++ft) {
      const int j0 = 64 * wave + 16 * ft + 4 * (lane >> 4);
      const v4f bq = *reinterpret_cast<const v4f *>(bl[L] + j0);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        const v4f v = acc[ft][rt] + bq;
        uint2 h;
        h.x = pack_bf2(gelu(v[0]), gelu(v[1]));
        h.y = pack_bf2(gelu(v[2]), gelu(v[3]));
        *reinterpret_cast<uint2 *>(&X[rt * 16 + xr][j0]) = h;
      }
    }
    __syncthreads();
  }
  // head: eight lanes per row, FR/32 rows per thread
#pragma unroll
  for (int i = 0; i < FR / 32; ++i) {
    const int row = i * 32 + (tid >> 3), part = tid & 7;
    float s = head_part(&X[row][part * 32], w5 + part * 32, 4);
    s = head_reduce8(s);
    if (part == 0 && r0 + row < nrow)
      pred[p0 + r0 + row] = exp((double)(s + b5) * scale + shift);
  }
}

long long jdsg_mlp_fused_work_len(long long nrow) { return (nrow + kFusedRows - 1) / kFusedRows + 1; }

void jdsg_launch_mlp_fused(long long e0, long long nrow, int ncol, const long long *ent_off,
                           const int *per_off, const double *periods, const int *nmode,
                           const float *acol, int lda, const float *ext, const uint16_t *W2,
                           const uint16_t *W3, const uint16_t *W4, const float *b2,
                           const float *b3, const float *b4, const uint16_t *w5, float b5,
                           double scale, double shift, double *pred, long long p0, int *tcol,
                           hipStream_t st) {
  if (nrow <= 0) return;
  constexpr int FR = kFusedRows;
  const int ntile = (int)((nrow + FR - 1) / FR);
  hipLaunchKernelGGL(jdsg_mlp_tile_cols<FR>, dim3((ntile + 256) / 256), dim3(256), 0, st, e0,
                     nrow, ncol, ent_off, ntile, tcol);
  hipLaunchKernelGGL(jdsg_mlp_fused<FR>, dim3((unsigned)ntile), dim3(256), 0, st, e0, nrow, ncol,
                     ent_off, per_off, periods, nmode, acol, lda, ext, W2, W3, W4, b2, b3, b4, w5,
                     b5, scale, shift, pred, p0, tcol);
}

void jdsg_launch_mlp_expand(long long e0, long long nrow, int ncol, const long long *ent_off,
                            const int *per_off, const double *periods, const int *nmode,
                            const float *acol, int lda, const float *ext, int H, uint16_t *Y,
                            hipStream_t st) {
  if (nrow <= 0 || H <= 0) return;
  hipLaunchKernelGGL(jdsg_mlp_expand, dim3((unsigned)((nrow + kExpandRows - 1) / kExpandRows)),
                     dim3(256), (size_t)36 * H, st, e0,
                     nrow, ncol, ent_off, per_off, periods, nmode, acol, lda, ext, H, Y);
}

// ----------------------------------------------------------------------------
extern "C" {

const char *jdsg_last_error(void);

// C = op(A) op(B) with the epilogue codes above; hipStream_t as void*.
int jdsg_gemm_bf16(int ta, int tb, int M, int N, int K, const void *A, int lda, const void *B,
                   int ldb, float *C, int ldc, int epi, const float *bias, void *Y, int ldy,
                   void *stream) {
  if (M % BM || N % BN || K % BK) return 1;
  hipStream_t st = (hipStream_t)stream;
  const uint16_t *a = (const uint16_t *)A, *b = (const uint16_t *)B;
  uint16_t *y = (uint16_t *)Y;
  // 16-byte loads need 8-element aligned leading dimensions
  // the big kernel also stores 4-column vectors: C rows 16-byte, Y rows
  // 8-byte and the bias 16-byte aligned
  bool big = M % BM2 == 0 && N % BN2 == 0 && K % BK2 == 0 && lda % 8 == 0 && ldb % 8 == 0 &&
             ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0 &&
             (C == nullptr || (ldc % 4 == 0 && ((uintptr_t)C & 15) == 0)) &&
             (y == nullptr || (ldy % 4 == 0 && ((uintptr_t)y & 7) == 0)) &&
             (bias == nullptr || ((uintptr_t)bias & 15) == 0);
  if (big) {
    dim3 g((unsigned)((M / BM2) * (N / BN2)));
    if (!ta && tb)
      hipLaunchKernelGGL((jdsg_gemm_bf16_128<false, true>), g, dim3(256), 0, st, M, N, K, a, lda,
                         b, ldb, C, ldc, epi, bias, y, ldy);
    else if (!ta && !tb)
      hipLaunchKernelGGL((jdsg_gemm_bf16_128<false, false>), g, dim3(256), 0, st, M, N, K, a,
                         lda, b, ldb, C, ldc, epi, bias, y, ldy);
    else if (ta && !tb)
      hipLaunchKernelGGL((jdsg_gemm_bf16_128<true, false>), g, dim3(256), 0, st, M, N, K, a, lda,
                         b, ldb, C, ldc, epi, bias, y, ldy);
    else
      hipLaunchKernelGGL((jdsg_gemm_bf16_128<true, true>), g, dim3(256), 0, st, M, N, K, a, lda,
                         b, ldb, C, ldc, epi, bias, y, ldy);
    return hipGetLastError() != hipSuccess;
  }
  dim3 grid(N / BN, M / BM);
  if (!ta && tb)
    hipLaunchKernelGGL((jdsg_gemm_bf16<false, true>), grid, dim3(256), 0, st, M, N, K, a, lda, b,
                       ldb, C, ldc, epi, bias, y, ldy);
  else if (!ta && !tb)
    hipLaunchKernelGGL((jdsg_gemm_bf16<false, false>), grid, dim3(256), 0, st, M, N, K, a, lda,
                       b, ldb, C, ldc, epi, bias, y, ldy);
  else if (ta && !tb)
    hipLaunchKernelGGL((jdsg_gemm_bf16<true, false>), grid, dim3(256), 0, st, M, N, K, a, lda, b,
                       ldb, C, ldc, epi, bias, y, ldy);
  else
    hipLaunchKernelGGL((jdsg_gemm_bf16<true, true>), grid, dim3(256), 0, st, M, N, K, a, lda, b,
                       ldb, C, ldc, epi, bias, y, ldy);
  return hipGetLastError() != hipSuccess;
}

int jdsg_gelu_bwd(long long n, const float *dY, const float *Z, void *dZ, float *dZf,
                  void *stream) {
  hipLaunchKernelGGL(jdsg_gelu_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n, dY, Z, (uint16_t *)dZ, dZf);
  return hipGetLastError() != hipSuccess;
}

int jdsg_colsum(int M, int N, const float *X, float *out, void *stream) {
  hipLaunchKernelGGL(jdsg_colsum, dim3((N + 63) / 64), dim3(256), 0, (hipStream_t)stream, M, N, X,
                     out);
  return hipGetLastError() != hipSuccess;
}

int jdsg_mse(int M, int ldo, const float *out, const float *target, const float *wt, float delta,
             float *dY, float *loss_part, void *stream) {
  hipLaunchKernelGGL(jdsg_mse, dim3((M + 255) / 256), dim3(256), 0, (hipStream_t)stream, M, ldo,
                     out, target, wt, delta, dY, loss_part);
  return hipGetLastError() != hipSuccess;
}

int jdsg_adam(long long n, float *w, float *m, float *v, const float *g, float lr, float b1,
              float b2, float eps, int step, float wd, void *wbf, void *stream) {
  float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(jdsg_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n, w, m, v, g, lr, b1, b2, eps, c1, c2, wd,
                     (uint16_t *)wbf);
  return hipGetLastError() != hipSuccess;
}

int jdsg_mlp_column_rows(int ncol, const float *colf, void *X, void *stream) {
  if (ncol <= 0) return 0;
  hipLaunchKernelGGL(jdsg_mlp_column_rows, dim3((ncol + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, ncol, colf, (uint16_t *)X);
  return hipGetLastError() != hipSuccess;
}

int jdsg_mlp_head(long long n, int H, const void *Y, const void *w, float b, double scale,
                  double shift, double *pred, long long e0, void *stream) {
  if (H % 64 || ((uintptr_t)Y & 15) || ((uintptr_t)w & 15)) return 1;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(jdsg_mlp_head, dim3((unsigned)((8 * n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n, H, (const uint16_t *)Y, (const uint16_t *)w, b,
                     scale, shift, pred, e0);
  return hipGetLastError() != hipSuccess;
}

int jdsg_f32_to_bf16(long long n, const float *x, void *y, void *stream) {
  hipLaunchKernelGGL(jdsg_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n, x, (uint16_t *)y);
  return hipGetLastError() != hipSuccess;
}

}  // extern "C"
