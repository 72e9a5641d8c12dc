// mlp.hip -- BF16 matrix-core kernels of the chemistry surrogate MLP
// (12 -> 256 -> 256 -> 256 -> 10, GELU) for CDNA4 (gfx950): forward with
// fused bias + GELU, backward (data and weight gradients, bias sums),
// loss gradient and Adam.  Operands are BF16 (raw 16-bit storage),
// accumulation FP32 on v_mfma_f32_16x16x32_bf16.  The surrogate has no
// counterpart in the reference; it replaces PSR_CHEMKIN/DVODE substeps
// (chemkin/psr_chemkin.f:1) for the cells its error indicator accepts.
#include "izem_ctx.h"

#include <stdint.h>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16;

#define MT 64   // tile rows (M)
#define NT 64   // tile columns (N)
#define KT 32   // tile depth (K), one MFMA 16x16x32 per wave sub-tile
#define LDT 40  // LDS row stride in bf16 (80 B: 16-byte aligned rows)

enum {
  EPI_F32 = 0,
  EPI_BIAS_GELU = 1,
  EPI_BIAS_F32 = 2,
  EPI_GELU_BWD = 3,
  EPI_GELU_ONLY = 5,  // inference: gelu(AB + bias) as bf16, no pre-activation kept
  EPI_PARTIAL = 4
};

__device__ inline float bf2f(u16 h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ inline u16 f2bf(float f) {
  // round to nearest even: one v_cvt_pk_bf16_f32 on gfx950
  return __builtin_bit_cast(u16, (__bf16)f);
}
// erf by Abramowitz & Stegun, Handbook of Mathematical Functions, 7.1.26:
// erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2) with
// t = 1 / (1 + p x), absolute error below 1.5e-7 for x >= 0, odd extension.
// Branch-free: the library erff takes a different polynomial on either side
// of |x| = 1 and a wave whose lanes straddle it runs both; its extra accuracy
// is lost in the BF16 activations anyway.
__device__ inline float erf_as(float x) {
  const float a = fabsf(x);
  const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * a);
  const float poly =
      t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f +
                                                                         t * 1.061405429f))));
  return copysignf(1.0f - poly * __expf(-a * a), x);
}
__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erf_as(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
  return 0.5f * (1.0f + erf_as(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

// Stage a 64 x 32 tile of the logical matrix X[R][Kd] into lds[r][k]:
// TR = 0: X stored row-major X[r*ld + k] (k contiguous);
// TR = 1: X stored transposed X[k*ld + r] (r contiguous).
// Out-of-range elements are zero.
template <int TR>
__device__ inline void stage_tile(u16 (*lds)[LDT], const u16 *__restrict__ X, long ld, int R,
                                  int Kd, int r0, int k0) {
  const int t = threadIdx.x;
  if (TR == 0) {
    const int r = t >> 2, kc = (t & 3) * 8;
    const int gr = r0 + r, gk = k0 + kc;
    u16 v[8];
    if (gr < R && gk + 7 < Kd && (ld & 7) == 0 && (((uintptr_t)X) & 15) == 0) {
      const uint4 q = *(const uint4 *)(X + (long)gr * ld + gk);
      const u16 *p = (const u16 *)&q;
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = p[i];
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = (gr < R && gk + i < Kd) ? X[(long)gr * ld + gk + i] : 0;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) lds[r][kc + i] = v[i];
  } else {
    const int k = t >> 3, rc = (t & 7) * 8;
    const int gk = k0 + k, gr = r0 + rc;
    u16 v[8];
    if (gk < Kd && gr + 7 < R && (ld & 7) == 0 && (((uintptr_t)X) & 15) == 0) {
      const uint4 q = *(const uint4 *)(X + (long)gk * ld + gr);
      const u16 *p = (const u16 *)&q;
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = p[i];
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = (gk < Kd && gr + i < R) ? X[(long)gk * ld + gr + i] : 0;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) lds[rc + i][k] = v[i];
  }
}

// C[M][N] = A[M][K] . B[K][N] (+ epilogue).  A: TA = 0 row-major, 1 stored
// as [K][M]; B: TB = 0 stored as [N][K] (weights transposed), 1 stored as
// [K][N].  Block: 256 threads = 4 waves on a 64 x 64 tile, each wave a
// 32 x 32 quarter as 2 x 2 MFMA 16x16x32 accumulators.  Split-K: blockIdx.z
// takes K range [z * kslice, (z + 1) * kslice) and (EPI_PARTIAL) writes its
// partial product to Cf + z * M * ldcf.
template <int TA, int TB, int EPI>
__global__ void __launch_bounds__(256)
k_gemm(int M, int N, int K, const u16 *__restrict__ A, long lda, const u16 *__restrict__ B,
       long ldb, const float *__restrict__ bias, const float *__restrict__ aux, long ldaux,
       float *__restrict__ Cf, long ldcf, u16 *__restrict__ Cb, long ldcb, int kslice) {
  __shared__ __attribute__((aligned(16))) u16 As[MT][LDT];
  __shared__ __attribute__((aligned(16))) u16 Bs[NT][LDT];
  const int m0 = blockIdx.y * MT, n0 = blockIdx.x * NT;
  const int kb = blockIdx.z * kslice, ke = min(K, kb + kslice);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = w >> 1, wn = w & 1;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += KT) {
    stage_tile<TA>(As, A, lda, M, ke, m0, k0);
    stage_tile<TB == 0 ? 0 : 1>(Bs, B, ldb, N, ke, n0, k0);
    __syncthreads();
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const uint4 q = *(const uint4 *)&As[wm * 32 + i * 16 + (lane & 15)][(lane >> 4) * 8];
      af[i] = __builtin_bit_cast(bf16x8, q);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const uint4 q = *(const uint4 *)&Bs[wn * 32 + j * 16 + (lane & 15)][(lane >> 4) * 8];
      bfr[j] = __builtin_bit_cast(bf16x8, q);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  float *Cz = Cf;
  if (EPI == EPI_PARTIAL) Cz = Cf + (long)blockIdx.z * M * ldcf;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + r;
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m >= M || n >= N) continue;
        const float a = acc[i][j][r];
        if (EPI == EPI_F32 || EPI == EPI_PARTIAL) {
          Cz[(long)m * ldcf + n] = a;
        } else if (EPI == EPI_BIAS_GELU) {
          const float z = a + bias[n];
          Cf[(long)m * ldcf + n] = z;
          Cb[(long)m * ldcb + n] = f2bf(gelu(z));
        } else if (EPI == EPI_BIAS_F32) {
          Cf[(long)m * ldcf + n] = a + bias[n];
        } else if (EPI == EPI_GELU_ONLY) {
          Cb[(long)m * ldcb + n] = f2bf(gelu(a + bias[n]));
        } else {  // EPI_GELU_BWD
          Cb[(long)m * ldcb + n] = f2bf(a * gelu_d(aux[(long)m * ldaux + n]));
        }
      }
}

// out[i] = sum over s of part[s * n + i] (fixed order: deterministic)
__global__ void k_sum_slices(long n, int ns, const float *__restrict__ part,
                             float *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < ns; z++) s += part[(long)z * n + i];
  out[i] = s;
}

// column sums of a BF16 matrix G[M][N] over row slices: part[s][n]
__global__ void k_colsum(int M, int N, const u16 *__restrict__ G, long ld, int rows,
                         float *__restrict__ part) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (n >= N) return;
  const int r0 = s * rows, r1 = min(M, r0 + rows);
  float a = 0.f;
  for (int r = r0; r < r1; r++) a += bf2f(G[(long)r * ld + n]);
  part[(long)s * N + n] = a;
}

// the same for N and ld multiples of 8 (16-byte rows of 8 columns): each
// thread sums 8 columns of every (256 / (N / 8))-th row of its slice, and
// the row groups of a block are added in a fixed order, so the sums do
// not depend on timing
__global__ void __launch_bounds__(256) k_colsum8(int M, int N, const u16 *__restrict__ G, long ld,
                                                 int rows, float *__restrict__ part) {
  __shared__ float red[256 * 8];
  const int cpr = N / 8;               // 16-byte chunks per row
  const int rg = 256 / cpr;            // rows per block iteration
  const int t = threadIdx.x;
  const int ch = t % cpr, g = t / cpr;
  const int s = blockIdx.x;
  const int r0 = s * rows, r1 = min(M, r0 + rows);
  float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (g < rg) {
    for (int r = r0 + g; r < r1; r += rg) {
      const uint4 w = *reinterpret_cast<const uint4 *>(G + (long)r * ld + ch * 8);
      const unsigned u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int q = 0; q < 4; q++) {
        a[2 * q] += __uint_as_float(u[q] << 16);
        a[2 * q + 1] += __uint_as_float(u[q] & 0xffff0000u);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 8; q++) red[t * 8 + q] = a[q];
  __syncthreads();
  // thread (chunk, column) adds the row groups in order
  if (t < N) {
    const int c = t / 8, q = t % 8;
    float sum = 0.f;
    for (int gg = 0; gg < rg; gg++) sum += red[(gg * cpr + c) * 8 + q];
    part[(long)s * N + t] = sum;
  }
}

// Adam (Kingma & Ba 2015) on an FP32 master matrix P[rows][cols] (row-major)
// with gradient G, moments m, v; writes BF16 copies P^T (ldt) and P (ldw)
// when given.  Step t >= 1.
__global__ void k_adam(int rows, int cols, float *__restrict__ P, const float *__restrict__ G,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2, u16 *__restrict__ Pt, long ldt,
                       u16 *__restrict__ Pw, long ldw) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)rows * cols) return;
  const float g = G[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float p = P[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  P[i] = p;
  const int r = (int)(i / cols), c = (int)(i % cols);
  if (Pt) Pt[(long)c * ldt + r] = f2bf(p);
  if (Pw) Pw[(long)r * ldw + c] = f2bf(p);
}

// write BF16 copies of an FP32 matrix (initialisation)
__global__ void k_to_bf16(int rows, int cols, const float *__restrict__ P, u16 *__restrict__ Pt,
                          long ldt, u16 *__restrict__ Pw, long ldw) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)rows * cols) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  if (Pt) Pt[(long)c * ldt + r] = f2bf(P[i]);
  if (Pw) Pw[(long)r * ldw + c] = f2bf(P[i]);
}

// mean-square loss over nout of the ldp output columns: dY = 2 (pred -
// target) / (B nout) as BF16 (padding columns zero); loss sum of squares
// accumulated into *loss (one atomic per block)
__global__ void k_mse(int Bn, int nout, int npad, const float *__restrict__ pred, long ldp,
                      const float *__restrict__ tgt, long ldt, float scale,
                      u16 *__restrict__ dY, long ldd, float *__restrict__ loss) {
  __shared__ float red[4];
  const long n = (long)Bn * npad;
  float sq = 0.f;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long b = i / npad;
    const int j = (int)(i % npad);
    float g = 0.f;
    if (j < nout) {
      const float d = pred[b * ldp + j] - tgt[b * ldt + j];
      sq += d * d;
      g = 2.f * d * scale;
    }
    dY[b * ldd + j] = f2bf(g);
  }
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += red[q];
    atomicAdd(loss, s);
  }
}

// ------------------------------------------------------------------ C ABI
#define LAUNCH_GEMM(TA_, TB_, EPI_)                                                           \
  hipLaunchKernelGGL((k_gemm<TA_, TB_, EPI_>), grid, dim3(256), 0, ctx->stream, M, N, K,      \
                     (const u16 *)A, lda, (const u16 *)B, ldb, bias, aux, ldaux, Cf, ldcf,     \
                     (u16 *)Cb, ldcb, kslice)

template <int EPI>
static void gemm_dispatch(izem_ctx *ctx, dim3 grid, int ta, int tb, int M, int N, int K,
                          const void *A, long lda, const void *B, long ldb, const float *bias,
                          const float *aux, long ldaux, float *Cf, long ldcf, void *Cb,
                          long ldcb, int kslice) {
  if (!ta && !tb) LAUNCH_GEMM(0, 0, EPI);
  else if (!ta && tb) LAUNCH_GEMM(0, 1, EPI);
  else if (ta && !tb) LAUNCH_GEMM(1, 0, EPI);
  else LAUNCH_GEMM(1, 1, EPI);
}

extern "C" int izem_mlp_gemm(izem_ctx *ctx, int ta, int tb, int epi, int M, int N, int K,
                             const void *A, long lda, const void *B, long ldb,
                             const float *bias, const float *aux, long ldaux, float *Cf,
                             long ldcf, void *Cb, long ldcb) {
  if (!ctx || !A || !B || M <= 0 || N <= 0 || K <= 0 || epi < 0 || epi > 5 || epi == 4)
    return IZEM_E_ARG;
  if ((epi == EPI_BIAS_GELU || epi == EPI_BIAS_F32 || epi == EPI_GELU_ONLY) && !bias)
    return IZEM_E_ARG;
  if ((epi == EPI_BIAS_GELU || epi == EPI_GELU_BWD || epi == EPI_GELU_ONLY) && !Cb)
    return IZEM_E_ARG;
  if (epi == EPI_GELU_BWD && !aux) return IZEM_E_ARG;
  if ((epi == EPI_F32 || epi == EPI_BIAS_GELU || epi == EPI_BIAS_F32) && !Cf) return IZEM_E_ARG;
  dim3 grid((N + NT - 1) / NT, (M + MT - 1) / MT, 1);
  const int ks = ((K + KT - 1) / KT) * KT;
  int pk = iz_prof_begin(ctx, IZP_OTHER);
  switch (epi) {
    case EPI_F32: gemm_dispatch<EPI_F32>(ctx, grid, ta, tb, M, N, K, A, lda, B, ldb, bias, aux, ldaux, Cf, ldcf, Cb, ldcb, ks); break;
    case EPI_BIAS_GELU: gemm_dispatch<EPI_BIAS_GELU>(ctx, grid, ta, tb, M, N, K, A, lda, B, ldb, bias, aux, ldaux, Cf, ldcf, Cb, ldcb, ks); break;
    case EPI_BIAS_F32: gemm_dispatch<EPI_BIAS_F32>(ctx, grid, ta, tb, M, N, K, A, lda, B, ldb, bias, aux, ldaux, Cf, ldcf, Cb, ldcb, ks); break;
    case EPI_GELU_ONLY: gemm_dispatch<EPI_GELU_ONLY>(ctx, grid, ta, tb, M, N, K, A, lda, B, ldb, bias, aux, ldaux, Cf, ldcf, Cb, ldcb, ks); break;
    default: gemm_dispatch<EPI_GELU_BWD>(ctx, grid, ta, tb, M, N, K, A, lda, B, ldb, bias, aux, ldaux, Cf, ldcf, Cb, ldcb, ks); break;
  }
  iz_prof_end(ctx, pk);
  IZ_CHECK(hipGetLastError());
  return IZEM_OK;
}

extern "C" int izem_mlp_gemm_splitk(izem_ctx *ctx, int ta, int tb, int M, int N, int K,
                                    const void *A, long lda, const void *B, long ldb,
                                    int nslices, float *work, float *C) {
  if (!ctx || !A || !B || !work || !C || M <= 0 || N <= 0 || K <= 0 || nslices < 1)
    return IZEM_E_ARG;
  int ks = (K + nslices - 1) / nslices;
  ks = ((ks + KT - 1) / KT) * KT;
  const int nz = (K + ks - 1) / ks;
  dim3 grid((N + NT - 1) / NT, (M + MT - 1) / MT, nz);
  const float *nul = nullptr;
  gemm_dispatch<EPI_PARTIAL>(ctx, grid, ta, tb, M, N, K, A, lda, B, ldb, nul, nul, 0, work,
                             N, nullptr, 0, ks);
  const long n = (long)M * N;
  hipLaunchKernelGGL(k_sum_slices, dim3(iz_blocks(n, 256)), dim3(256), 0, ctx->stream, n, nz,
                     work, C);
  IZ_CHECK(hipGetLastError());
  return IZEM_OK;
}

extern "C" int izem_mlp_colsum(izem_ctx *ctx, int M, int N, const void *G, long ld,
                               int nslices, float *work, float *out) {
  if (!ctx || !G || !work || !out || M <= 0 || N <= 0 || nslices < 1) return IZEM_E_ARG;
  const int rows = (M + nslices - 1) / nslices;
  const int ns = (M + rows - 1) / rows;
  if (N % 8 == 0 && ld % 8 == 0 && N <= 256 && ((uintptr_t)G & 15) == 0)
    hipLaunchKernelGGL(k_colsum8, dim3(ns), dim3(256), 0, ctx->stream, M, N, (const u16 *)G, ld,
                       rows, work);
  else
    hipLaunchKernelGGL(k_colsum, dim3((N + 255) / 256, ns), dim3(256), 0, ctx->stream, M, N,
                       (const u16 *)G, ld, rows, work);
  hipLaunchKernelGGL(k_sum_slices, dim3(iz_blocks(N, 256)), dim3(256), 0, ctx->stream, (long)N,
                     ns, work, out);
  IZ_CHECK(hipGetLastError());
  return IZEM_OK;
}

extern "C" int izem_mlp_adam(izem_ctx *ctx, int rows, int cols, float *P, const float *G,
                             float *m, float *v, float lr, float b1, float b2, float eps,
                             int step, void *Pt, long ldt, void *Pw, long ldw) {
  if (!ctx || !P || !G || !m || !v || rows <= 0 || cols <= 0 || step < 1) return IZEM_E_ARG;
  const float c1 = 1.f / (1.f - powf(b1, (float)step));
  const float c2 = 1.f / (1.f - powf(b2, (float)step));
  const long n = (long)rows * cols;
  hipLaunchKernelGGL(k_adam, dim3(iz_blocks(n, 256)), dim3(256), 0, ctx->stream, rows, cols, P,
                     G, m, v, lr, b1, b2, eps, c1, c2, (u16 *)Pt, ldt, (u16 *)Pw, ldw);
  IZ_CHECK(hipGetLastError());
  return IZEM_OK;
}

extern "C" int izem_mlp_to_bf16(izem_ctx *ctx, int rows, int cols, const float *P, void *Pt,
                                long ldt, void *Pw, long ldw) {
  if (!ctx || !P || rows <= 0 || cols <= 0) return IZEM_E_ARG;
  const long n = (long)rows * cols;
  hipLaunchKernelGGL(k_to_bf16, dim3(iz_blocks(n, 256)), dim3(256), 0, ctx->stream, rows, cols,
                     P, (u16 *)Pt, ldt, (u16 *)Pw, ldw);
  IZ_CHECK(hipGetLastError());
  return IZEM_OK;
}

extern "C" int izem_mlp_mse(izem_ctx *ctx, int Bn, int nout, int npad, const float *pred,
                            long ldp, const float *tgt, long ldt, void *dY, long ldd,
                            float *loss) {
  if (!ctx || !pred || !tgt || !dY || !loss || Bn <= 0 || nout <= 0 || npad < nout)
    return IZEM_E_ARG;
  const long n = (long)Bn * npad;
  const float scale = 1.f / ((float)Bn * (float)nout);
  // a few blocks per CU: one same-address atomic per block
  hipLaunchKernelGGL(k_mse, dim3(iz_blocks_capped(ctx, n, 256, 2)), dim3(256), 0, ctx->stream,
                     Bn, nout, npad,
                     pred, ldp, tgt, ldt, scale, (u16 *)dY, ldd, loss);
  IZ_CHECK(hipGetLastError());
  return IZEM_OK;
}

// -------------------------------------------------------------------
// Fused inference: all four layers of the surrogate for a tile of 64
// cells in one block.  The activations stay in one LDS buffer (64 x 256
// BF16), overwritten in place: every wave reads the whole input tile, so
// the block synchronises once after the products and once after the
// epilogue has written the layer's output.  The weights are read as MFMA
// operands straight from L2 (every block uses the same 273 KB), the
// fragments of the next 32-deep slice loaded while the current one is in
// the matrix cores.  Persistent grid over row tiles.  Weights Wt_l are
// stored [N][K] (the B operand of a 16x16x32 MFMA is then one 16-byte load
// per lane); layer widths K = 32 | 256 | 256 | 256, N = 256 | 256 | 256 | 64.
// -------------------------------------------------------------------
#define FX 264  // LDS row stride of an activation tile (bf16): 528 B, 16-byte aligned

template <int K, int N, bool LAST, int RB>
__device__ inline void fused_layer(u16 (*X)[FX], const u16 *__restrict__ Wt,
                                   const float *__restrict__ b, float *__restrict__ pred,
                                   int row0, int M) {
  constexpr int NW = N / 4;        // columns per wave
  constexpr int CB = NW / 16;      // 16-column blocks per wave
  constexpr int KS = K / 32;       // 32-deep slices
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  f32x4 acc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; i++)
#pragma unroll
    for (int j = 0; j < CB; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const u16 *wrow[CB];
#pragma unroll
  for (int j = 0; j < CB; j++)
    wrow[j] = Wt + (long)(w * NW + j * 16 + (lane & 15)) * K + (lane >> 4) * 8;
  bf16x8 bcur[CB];
#pragma unroll
  for (int j = 0; j < CB; j++) bcur[j] = __builtin_bit_cast(bf16x8, *(const uint4 *)wrow[j]);
#pragma unroll 1
  for (int ks = 0; ks < KS; ks++) {
    const int k0 = ks * 32;
    // next slice's weight fragments, in flight during this slice's products
    bf16x8 bnext[CB];
    if (ks + 1 < KS) {
#pragma unroll
      for (int j = 0; j < CB; j++)
        bnext[j] = __builtin_bit_cast(bf16x8, *(const uint4 *)(wrow[j] + k0 + 32));
    }
#pragma unroll
    for (int i = 0; i < RB; i++) {
      const uint4 q = *(const uint4 *)&X[i * 16 + (lane & 15)][k0 + (lane >> 4) * 8];
      const bf16x8 af = __builtin_bit_cast(bf16x8, q);
#pragma unroll
      for (int j = 0; j < CB; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bcur[j], acc[i][j], 0, 0, 0);
    }
    if (ks + 1 < KS) {
#pragma unroll
      for (int j = 0; j < CB; j++) bcur[j] = bnext[j];
    }
  }
  if (!LAST) __syncthreads();  // every wave has read the input tile
#pragma unroll
  for (int i = 0; i < RB; i++)
#pragma unroll
    for (int j = 0; j < CB; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = i * 16 + (lane >> 4) * 4 + r;
        const int n = w * NW + j * 16 + (lane & 15);
        const float z = acc[i][j][r] + b[n];
        if (LAST) {
          if (row0 + m < M) pred[(long)(row0 + m) * N + n] = z;
        } else {
          X[m][n] = f2bf(gelu(z));
        }
      }
  if (!LAST) __syncthreads();  // the output tile is complete
}

// resident 256-thread blocks per CU (one wave per SIMD each)
#ifndef IZ_MLP_WAVES
#define IZ_MLP_WAVES 2
#endif
// rows of an activation tile (16 per row block)
#ifndef IZ_MLP_ROWS
#define IZ_MLP_ROWS 128
#endif

__global__ void __attribute__((amdgpu_flat_work_group_size(256, 256),
                               amdgpu_waves_per_eu(IZ_MLP_WAVES, IZ_MLP_WAVES)))
k_mlp_fused(int M, const u16 *__restrict__ X, const u16 *__restrict__ W0,
            const u16 *__restrict__ W1, const u16 *__restrict__ W2, const u16 *__restrict__ W3,
            const float *__restrict__ b0, const float *__restrict__ b1,
            const float *__restrict__ b2, const float *__restrict__ b3,
            float *__restrict__ pred) {
  constexpr int R = IZ_MLP_ROWS, RB = R / 16;
  __shared__ __attribute__((aligned(16))) u16 A[R][FX];
  const int ntiles = (M + R - 1) / R;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int row0 = tile * R;
    // input tile: R rows x 32 features (16 B per thread and pass)
#pragma unroll
    for (int p = 0; p < R / 64; p++) {
      const int t = threadIdx.x, r = p * 64 + (t >> 2), c = (t & 3) * 8;
      uint4 q = make_uint4(0, 0, 0, 0);
      if (row0 + r < M) q = *(const uint4 *)(X + (long)(row0 + r) * 32 + c);
      *(uint4 *)&A[r][c] = q;
    }
    __syncthreads();
    fused_layer<32, 256, false, RB>(A, W0, b0, nullptr, row0, M);
    fused_layer<256, 256, false, RB>(A, W1, b1, nullptr, row0, M);
    fused_layer<256, 256, false, RB>(A, W2, b2, nullptr, row0, M);
    fused_layer<256, 64, true, RB>(A, W3, b3, pred, row0, M);
    __syncthreads();  // the next tile's input overwrites A
  }
}

extern "C" int izem_mlp_infer(izem_ctx *ctx, int M, const void *X, const void *const *Wt,
                              const float *const *b, float *pred) {
  if (!ctx || M <= 0 || !X || !Wt || !b || !pred) return IZEM_E_ARG;
  for (int l = 0; l < 4; l++)
    if (!Wt[l] || !b[l] || (((uintptr_t)Wt[l]) & 15)) return IZEM_E_ARG;
  if (((uintptr_t)X) & 15) return IZEM_E_ARG;
  const long ntiles = (M + IZ_MLP_ROWS - 1) / IZ_MLP_ROWS;
  long g = (long)ctx->ncu * IZ_MLP_WAVES;
  if (g > ntiles) g = ntiles;
  int pk = iz_prof_begin(ctx, IZP_CHEM);
  hipLaunchKernelGGL(k_mlp_fused, dim3((unsigned)g), dim3(256), 0, ctx->stream, M,
                     (const u16 *)X, (const u16 *)Wt[0], (const u16 *)Wt[1], (const u16 *)Wt[2],
                     (const u16 *)Wt[3], b[0], b[1], b[2], b[3], pred);
  iz_prof_end(ctx, pk);
  IZ_CHECK(hipGetLastError());
  return IZEM_OK;
}
