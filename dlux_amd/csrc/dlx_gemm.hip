// Batched complex "NT" GEMM on CDNA4 matrix cores (gfx950).
//
//   C[b](m, n) = sum_k A[b](m, k) * B[b](n, k)        (complex64 operands)
//
// Both operands are row-major with k contiguous, which is the layout of every
// contraction in the MFT once the transfer matrix is stored N_out x N_in:
//   stage 1:  U^T = (P  . W_x^T)^T   A = P   (N_in x N_in),  B = W_x (N_out x N_in)
//   stage 2:  E   =  W_y . U         A = W_y (N_out x N_in), B = U^T (N_out x N_in)
// (utils/propagation.py:243 computes the same product as (Y^T E) X.)
//
// This is the complex64 / FP32 form (v_mfma_f32_32x32x2_f32, four real
// products per complex MAC, FP32 accumulation).  The split-BF16x3 form lives
// in dlx_gemm_bf16.hip.
//
// Workgroup = 4 waves (2 x 2), each wave owns TM x TN blocks of 32 x 32
// outputs, so the workgroup tile is (64 TM) x (64 TN).  One LDS stage holds a
// BM x BK tile of A and a BN x BK tile of B; the next stage's global loads are
// issued into registers before the current stage's MFMAs (register
// double-buffering), so HBM latency overlaps matrix-core work.
#include "dlx_internal.h"

namespace dlx {
namespace {

typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// Operand staging: a ROWS x BK tile of a row-major complex64 matrix.
// Thread t handles column k = t % BK of rows t / BK + r * (256 / BK).
// ---------------------------------------------------------------------------
template <int ROWS, int BK>
struct Stage {
  static constexpr int kRowStep = kThreads / BK;
  static constexpr int kPer = ROWS / kRowStep;   // float2 loads per thread
  float2 v[kPer];

  __device__ __forceinline__ void load(const float2* __restrict__ base, long ld,
                                       int row0, int nrows, int k0, int K) {
    const int t = threadIdx.x;
    const int kk = k0 + (t % BK);
    const int r0 = t / BK;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int r = row0 + r0 + i * kRowStep;
      v[i] = (r < nrows && kk < K) ? base[(long)r * ld + kk] : make_float2(0.f, 0.f);
    }
  }
};

// ---------------------------------------------------------------------------
// FP32 form
// ---------------------------------------------------------------------------
template <int TM, int TN, int EPI>
__global__ void __launch_bounds__(kThreads)
cgemm_nt_fp32(GemmBatch g) {
  constexpr int BK = 16;
  constexpr int BM = 64 * TM, BN = 64 * TN;
  constexpr int LD = BK + 1;  // float2 stride: 34 dwords -> conflict-free frags
  __shared__ float2 As[BM * LD];
  __shared__ float2 Bs[BN * LD];

  const int b = blockIdx.z;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const float2* __restrict__ A = g.A + (long)b * g.sA;
  const float2* __restrict__ B = g.B + (long)b * g.sB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32 * TM, wn = (wave & 1) * 32 * TN;
  const int lr = lane & 31, lk = lane >> 5;

  f32x16_t accR[TM][TN], accI[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) { accR[i][j][r] = 0.f; accI[i][j][r] = 0.f; }

  Stage<BM, BK> sa;
  Stage<BN, BK> sb;
  sa.load(A, g.lda, m0, g.M, 0, g.K);
  sb.load(B, g.ldb, n0, g.N, 0, g.K);

  for (int k0 = 0; k0 < g.K; k0 += BK) {
    __syncthreads();
    {
      const int kk = threadIdx.x % BK, r0 = threadIdx.x / BK;
#pragma unroll
      for (int i = 0; i < Stage<BM, BK>::kPer; ++i)
        As[(r0 + i * Stage<BM, BK>::kRowStep) * LD + kk] = sa.v[i];
#pragma unroll
      for (int i = 0; i < Stage<BN, BK>::kPer; ++i)
        Bs[(r0 + i * Stage<BN, BK>::kRowStep) * LD + kk] = sb.v[i];
    }
    __syncthreads();
    if (k0 + BK < g.K) {
      sa.load(A, g.lda, m0, g.M, k0 + BK, g.K);
      sb.load(B, g.ldb, n0, g.N, k0 + BK, g.K);
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 2) {
      float2 fa[TM], fb[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) fa[i] = As[(wm + 32 * i + lr) * LD + ks + lk];
#pragma unroll
      for (int j = 0; j < TN; ++j) fb[j] = Bs[(wn + 32 * j + lr) * LD + ks + lk];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          accR[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i].x, fb[j].x, accR[i][j], 0, 0, 0);
          accR[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(-fa[i].y, fb[j].y, accR[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i].x, fb[j].y, accI[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i].y, fb[j].x, accI[i][j], 0, 0, 0);
        }
    }
  }

  // Epilogue: 32x32 MFMA output layout -- lane holds column (lane % 32),
  // register r holds row 8*(r/4) + 4*(lane/32) + r%4.
  const float alpha = g.alpha[b];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm + 32 * i + 8 * (r >> 2) + 4 * lk + (r & 3);
        const int n = n0 + wn + 32 * j + lr;
        if (m < g.M && n < g.N) {
          const float re = alpha * accR[i][j][r], im = alpha * accI[i][j][r];
          if (EPI == EPI_STORE) {
            ((float2*)g.C)[(long)b * g.sC + (long)m * g.ldc + n] = make_float2(re, im);
          } else if (EPI == EPI_STORE_T) {
            ((float2*)g.C)[(long)b * g.sC + (long)n * g.ldc + m] = make_float2(re, im);
          } else if (EPI == EPI_VJP) {
            const float gm = 2.f * g.G[(long)b * g.g_batch + (long)m * g.ldg + n];
            ((float2*)g.C)[(long)b * g.sC + (long)m * g.ldc + n] = make_float2(gm * re, gm * im);
          } else if (EPI == EPI_GRAD) {
            const float2 pv = ((const float2*)g.P)[(long)b * g.p_batch + (long)m * g.p_ld + n];
            const long o = (long)b * g.sC + (long)m * g.ldc + n;
            ((float*)g.C)[o] = pv.x * im - pv.y * re;
            if (g.C2) {
              const float2 uv =
                  ((const float2*)g.U)[(long)b * g.p_batch + (long)m * g.p_ld + n];
              g.C2[o] = uv.x * re + uv.y * im;
            }
          } else {
            ((float*)g.C)[(long)b * g.sC + (long)m * g.ldc + n] = re * re + im * im;
          }
        }
      }
}

template <int TM, int TN>
hipError_t dispatch(const GemmBatch& g, hipStream_t s) {
  dim3 grid((g.N + 64 * TN - 1) / (64 * TN), (g.M + 64 * TM - 1) / (64 * TM), g.batch);
  dim3 block(kThreads);
  switch (g.epilogue) {
    case EPI_STORE: hipLaunchKernelGGL((cgemm_nt_fp32<TM, TN, EPI_STORE>), grid, block, 0, s, g); break;
    case EPI_STORE_T: hipLaunchKernelGGL((cgemm_nt_fp32<TM, TN, EPI_STORE_T>), grid, block, 0, s, g); break;
    case EPI_VJP: hipLaunchKernelGGL((cgemm_nt_fp32<TM, TN, EPI_VJP>), grid, block, 0, s, g); break;
    case EPI_GRAD: hipLaunchKernelGGL((cgemm_nt_fp32<TM, TN, EPI_GRAD>), grid, block, 0, s, g); break;
    default: hipLaunchKernelGGL((cgemm_nt_fp32<TM, TN, EPI_INTENSITY>), grid, block, 0, s, g); break;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_cgemm_nt(const GemmBatch& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.batch <= 0) return hipSuccess;
  if (g.batch > GemmBatch::kMaxBatch) return hipErrorInvalidValue;
  // Small problems: 64x64 workgroup tiles keep enough workgroups in flight.
  const long tiles128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128) * g.batch;
  if (g.M <= 64 || g.N <= 64 || tiles128 < 512) return dispatch<1, 1>(g, s);
  return dispatch<2, 2>(g, s);
}

}  // namespace dlx
