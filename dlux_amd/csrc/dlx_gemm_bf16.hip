// Split-BF16x3 complex NT GEMM on split4 operands (gfx950 matrix cores).
//
//   C[b](m, n) = alpha_b * sum_k A[b](m, k) * B[b](n, k)
//
// A and B are split4 matrices (dlx_split.h): each real product is
// hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_bf16 with FP32 accumulation,
// 12 MFMAs per complex 32x32x16 block step.
//
// Workgroup = 4 waves (2 x 2), each wave owns TM x TN blocks of 32 x 32
// complex outputs.  A K-stage is 32 k (4 chunks): every row of a tile is 16
// pieces of 16 bytes, staged global -> registers -> LDS with 16-byte accesses.
// LDS is double-buffered so the next stage's LDS writes follow the current
// stage's MFMAs without a barrier in between; there is one barrier per stage.
// Workgroups are numbered so that all tiles of one instance run on the same
// XCD and share its L2 (the hardware deals workgroups round-robin over 8 XCDs).
#include "dlx_internal.h"
#include "dlx_phasor.h"
#include "dlx_split.h"

#include <cstdlib>
#include <string>

namespace dlx {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4_t;

constexpr int kThreads = 256;
constexpr int kStageChunks = 4;                        // 32 k per stage
constexpr int kRowLds = kStageChunks * kChunkBytes + 16;  // 272 B: conflict-free 16 B reads
constexpr int kXcds = 8;

__device__ __forceinline__ bf16x8_t as_bf(u32x4_t v) { return __builtin_bit_cast(bf16x8_t, v); }
__device__ __forceinline__ u32x4_t neg_bf(u32x4_t v) {
  return v ^ (u32x4_t){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
}

// Workgroup barrier for LDS hand-over: waits for this wave's LDS accesses
// only (lgkmcnt), not for outstanding global loads, which __syncthreads()'s
// fence would drain.  The memory clobber keeps the compiler from moving
// memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Hardware workgroup id -> logical id such that consecutive logical ids land
// on the same XCD (XCD = hw % 8).
__device__ __forceinline__ int xcd_remap(int hw, int total) {
  const int per = total / kXcds;
  const int rem = total - per * kXcds;
  if (hw >= per * kXcds) return hw;  // tail that does not fill all XCDs evenly
  (void)rem;
  return (hw % kXcds) * per + hw / kXcds;
}

template <int ROWS>
struct Stage16 {
  static constexpr int kPer = ROWS / 16;   // 16 threads per row (16 pieces)
  uint4 v[kPer];

  __device__ __forceinline__ void load(const uint8_t* __restrict__ base, long row_bytes,
                                       int row0, int nrows, int kc0, int kchunks) {
    const int t = threadIdx.x;
    const int piece = t & 15;
    const int chunk = kc0 + (piece >> 2);
    const int plane = piece & 3;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      // Branch-free bounds handling: out-of-range lanes read row 0 / chunk 0 of
      // the operand (always valid) and zero the value.
      const int r = row0 + (t >> 4) + 16 * i;
      const bool ok = r < nrows && chunk < kchunks;
      const long off = ok ? (long)r * row_bytes + (long)chunk * kChunkBytes + plane * 16 : 0;
      const uint4 x = *(const uint4*)(base + off);
      v[i] = ok ? x : make_uint4(0u, 0u, 0u, 0u);
    }
  }

  __device__ __forceinline__ void store(uint8_t* lds) const {
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < kPer; ++i)
      *(uint4*)(lds + ((t >> 4) + 16 * i) * kRowLds + (t & 15) * 16) = v[i];
  }
};

// Epilogue shared by the GEMM variants.  32x32 MFMA layout: lane holds column
// lane % 32, register r holds row 8*(r/4) + 4*(lane/32) + r%4.
template <int TM, int TN, int EPI>
__device__ __forceinline__ void epilogue_split(const GemmSplit& g, int b, int m0, int n0, int wm,
                                               int wn, int lr, int lk,
                                               f32x16_t (&accR)[TM][TN],
                                               f32x16_t (&accI)[TM][TN]) {
  const float alpha = g.alpha[b];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wn + 32 * j + lr;
      if (EPI == EPI_SPLIT_T || EPI == EPI_VJP_SPLIT_T) {
        // rows n of a split4 matrix whose k is this GEMM's m: 4 consecutive m
        // per register group -> half a chunk (8 bytes) per plane.  The VJP form
        // scales element (m, n) by 2 G(n, m).
        if (n >= g.N) continue;
        const int m_pad = (g.M + 7) & ~7;   // the last chunk is zero-padded
        uint8_t* crow = (uint8_t*)g.C + (long)b * g.c_batch + (long)n * g.ldc;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int mq = m0 + wm + 32 * i + 8 * q + 4 * lk;   // first of 4 rows
          if (mq >= m_pad) continue;
          float re[4], im[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const bool ok = mq + e < g.M;
            float f = alpha;
            if (EPI == EPI_VJP_SPLIT_T && ok)
              f *= 2.f * g.G[(long)b * g.g_batch + (long)n * g.ldg + mq + e];
            re[e] = ok ? f * accR[i][j][4 * q + e] : 0.f;
            im[e] = ok ? f * accI[i][j][4 * q + e] : 0.f;
          }
          uint32_t rh[2], rl[2], ih[2], il[2];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            split2_bf16(re[2 * h], re[2 * h + 1], rh[h], rl[h]);
            split2_bf16(im[2 * h], im[2 * h + 1], ih[h], il[h]);
          }
          uint8_t* dst = crow + (long)(mq >> 3) * kChunkBytes + ((mq & 7) >> 2) * 8;
          *(uint2*)(dst + 0) = make_uint2(rh[0], rh[1]);
          *(uint2*)(dst + 16) = make_uint2(rl[0], rl[1]);
          *(uint2*)(dst + 32) = make_uint2(ih[0], ih[1]);
          *(uint2*)(dst + 48) = make_uint2(il[0], il[1]);
        }
      } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int m = m0 + wm + 32 * i + 8 * (r >> 2) + 4 * lk + (r & 3);
          if (m < g.M && n < g.N) {
            const float re = alpha * accR[i][j][r], im = alpha * accI[i][j][r];
            if (EPI == EPI_STORE) {
              ((float2*)g.C)[(long)b * g.c_batch + (long)m * g.ldc + n] = make_float2(re, im);
            } else if (EPI == EPI_GRAD) {
              // P(m, n) from its split4 row m: value = hi + lo per component
              const uint8_t* pr = (const uint8_t*)g.P + (long)b * g.p_batch + (long)m * g.p_ld +
                                  (long)(n >> 3) * kChunkBytes + (n & 7) * 2;
              const float pre = __uint_as_float((uint32_t)*(const uint16_t*)(pr) << 16) +
                                __uint_as_float((uint32_t)*(const uint16_t*)(pr + 16) << 16);
              const float pim = __uint_as_float((uint32_t)*(const uint16_t*)(pr + 32) << 16) +
                                __uint_as_float((uint32_t)*(const uint16_t*)(pr + 48) << 16);
              const long o = (long)b * g.c_batch + (long)m * g.ldc + n;
              ((float*)g.C)[o] = pre * im - pim * re;
              if (g.C2) {   // the phasor U in P's layout
                const uint8_t* ur = (const uint8_t*)g.U + (pr - (const uint8_t*)g.P);
                const float ure = __uint_as_float((uint32_t)*(const uint16_t*)(ur) << 16) +
                                  __uint_as_float((uint32_t)*(const uint16_t*)(ur + 16) << 16);
                const float uim = __uint_as_float((uint32_t)*(const uint16_t*)(ur + 32) << 16) +
                                  __uint_as_float((uint32_t)*(const uint16_t*)(ur + 48) << 16);
                g.C2[o] = ure * re + uim * im;
              }
            } else {
              ((float*)g.C)[(long)b * g.c_batch + (long)m * g.ldc + n] = re * re + im * im;
            }
          }
        }
      }
    }
}

// EPI_SPLIT_T for an interior tile, staged through LDS: the lanes of the
// direct epilogue each write 8-byte pieces of 32 rows 4 KiB apart; here the
// tile's rows are assembled in LDS, 32 at a time (rows padded by 16 bytes, so
// the 32 rows of a lane group fall on different banks), and the workgroup
// writes each row's BM * 8 contiguous bytes with 16-byte stores.  Same bytes
// as epilogue_split.  `lds`: the main loop's buffers (>= 32 * (BM * 8 + 16)
// bytes), free once every wave has passed the first barrier below.
template <int TM, int TN, int EPI = EPI_SPLIT_T>
__device__ __forceinline__ void epilogue_split_t_lds(const GemmSplit& g, int b, int m0, int n0,
                                                     int wm, int wn, int lr, int lk,
                                                     f32x16_t (&accR)[TM][TN],
                                                     f32x16_t (&accI)[TM][TN], uint8_t* lds) {
  static_assert(EPI == EPI_SPLIT_T || EPI == EPI_VJP_SPLIT_T, "split4 transposed epilogues");
  constexpr int BM = 64 * TM;
  constexpr int kSeg = BM * 8;          // one Uᵀ row of the tile: BM / 8 chunks of 64 B
  constexpr int kRS = kSeg + 16;        // LDS row stride
  const float alpha = g.alpha[b];
  const int wave = threadIdx.x >> 6;
  uint8_t* crow0 = (uint8_t*)g.C + (long)b * g.c_batch + (long)(m0 >> 3) * kChunkBytes;
#pragma unroll
  for (int pw = 0; pw < 2; ++pw)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      __syncthreads();   // the previous pass's rows (or the main loop's reads) are done
      if ((wave & 1) == pw) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int mq = wm + 32 * i + 8 * q + 4 * lk;   // tile-local, first of 4
            float re[4], im[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              float f = alpha;
              if (EPI == EPI_VJP_SPLIT_T)   // element (m, n) scaled by 2 G(n, m)
                f *= 2.f * g.G[(long)b * g.g_batch + (long)(n0 + wn + 32 * j + lr) * g.ldg +
                               m0 + mq + e];
              re[e] = f * accR[i][j][4 * q + e];
              im[e] = f * accI[i][j][4 * q + e];
            }
            uint32_t rh[2], rl[2], ih[2], il[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              split2_bf16(re[2 * h], re[2 * h + 1], rh[h], rl[h]);
              split2_bf16(im[2 * h], im[2 * h + 1], ih[h], il[h]);
            }
            uint8_t* dst = lds + lr * kRS + (mq >> 3) * kChunkBytes + ((mq & 7) >> 2) * 8;
            *(uint2*)(dst + 0) = make_uint2(rh[0], rh[1]);
            *(uint2*)(dst + 16) = make_uint2(rl[0], rl[1]);
            *(uint2*)(dst + 32) = make_uint2(ih[0], ih[1]);
            *(uint2*)(dst + 48) = make_uint2(il[0], il[1]);
          }
      }
      __syncthreads();
      // rows n0 + pw * 32 TN + 32 j + (0..31): 16-byte pieces, kSeg / 16 per row
      const int nbase = n0 + pw * 32 * TN + 32 * j;
#pragma unroll
      for (int t = threadIdx.x; t < 32 * (kSeg / 16); t += kThreads) {
        const int r = t / (kSeg / 16), piece = t - r * (kSeg / 16);
        const uint4 v = *(const uint4*)(lds + r * kRS + piece * 16);
        *(uint4*)(crow0 + (long)(nbase + r) * g.ldc + piece * 16) = v;
      }
    }
}

template <int TM, int TN, int EPI>
__global__ void __launch_bounds__(kThreads)
cgemm_split4_kernel(GemmSplit g) {
  constexpr int BM = 64 * TM, BN = 64 * TN;
  constexpr int kBuf = (BM + BN) * kRowLds;
  __shared__ __attribute__((aligned(16))) uint8_t lds[2 * kBuf];

  const int tiles_n = (g.N + BN - 1) / BN, tiles_m = (g.M + BM - 1) / BM;
  const int total = tiles_n * tiles_m * g.batch;
  const int id = xcd_remap(blockIdx.x, total);
  const int tn = id % tiles_n;
  const int tm = (id / tiles_n) % tiles_m;
  const int b = id / (tiles_n * tiles_m);
  const int m0 = tm * BM, n0 = tn * BN;

  const uint8_t* __restrict__ A = g.A + (long)b * g.a_batch_bytes;
  const uint8_t* __restrict__ B = g.B + (long)b * g.b_batch_bytes;
  const int kchunks = (int)split4_chunks(g.K);
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

  // Two register stages in flight: the loads of K-stage s + 2 are issued
  // while stage s runs on the matrix cores, so a load has two stages of MFMA
  // work to cover its latency before its data is written to LDS.
  Stage16<BM> sa0, sa1;
  Stage16<BN> sb0, sb1;
  const int nst = (kchunks + kStageChunks - 1) / kStageChunks;
  sa0.load(A, g.a_row_bytes, m0, g.M, 0, kchunks);
  sb0.load(B, g.b_row_bytes, n0, g.N, 0, kchunks);
  if (nst > 1) {
    sa1.load(A, g.a_row_bytes, m0, g.M, kStageChunks, kchunks);
    sb1.load(B, g.b_row_bytes, n0, g.N, kStageChunks, kchunks);
  }
  sa0.store(lds);
  sb0.store(lds + BM * kRowLds);
  lds_barrier();

  auto mma_stage = [&](const uint8_t* la) {
    const uint8_t* lb = la + BM * kRowLds;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int coff = (2 * ks + lk) * kChunkBytes;
      u32x4_t fa[TM][4], fb[TN][4];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int p = 0; p < 4; ++p)
          fa[i][p] = *(const u32x4_t*)(la + (wm + 32 * i + lr) * kRowLds + coff + p * 16);
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int p = 0; p < 4; ++p)
          fb[j][p] = *(const u32x4_t*)(lb + (wn + 32 * j + lr) * kRowLds + coff + p * 16);
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const bf16x8_t arh = as_bf(fa[i][0]), arl = as_bf(fa[i][1]);
        const bf16x8_t aih = as_bf(fa[i][2]), ail = as_bf(fa[i][3]);
        const bf16x8_t naih = as_bf(neg_bf(fa[i][2])), nail = as_bf(neg_bf(fa[i][3]));
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const bf16x8_t brh = as_bf(fb[j][0]), brl = as_bf(fb[j][1]);
          const bf16x8_t bih = as_bf(fb[j][2]), bil = as_bf(fb[j][3]);
          f32x16_t r = accR[i][j], im = accI[i][j];
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, brh, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, bih, im, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brl, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bil, im, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nail, bih, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, brh, im, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bil, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brl, im, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brh, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bih, im, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bih, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brh, im, 0, 0, 0);
          accR[i][j] = r;
          accI[i][j] = im;
        }
      }
    }
  };

  // One K-stage: MFMAs on LDS buffer `cur`; reload `ld` with stage s + 2;
  // write `st` (holding stage s + 1) into the other buffer.
  auto step = [&](int s, Stage16<BM>& lda, Stage16<BN>& ldb, Stage16<BM>& sta,
                  Stage16<BN>& stb) {
    const uint8_t* cur = lds + (s & 1) * kBuf;
    if (s + 2 < nst) {
      lda.load(A, g.a_row_bytes, m0, g.M, (s + 2) * kStageChunks, kchunks);
      ldb.load(B, g.b_row_bytes, n0, g.N, (s + 2) * kStageChunks, kchunks);
    }
    mma_stage(cur);
    if (s + 1 < nst) {
      uint8_t* nb = lds + ((s + 1) & 1) * kBuf;
      sta.store(nb);
      stb.store(nb + BM * kRowLds);
    }
    lds_barrier();
  };

  for (int s = 0; s < nst; s += 2) {
    step(s, sa0, sb0, sa1, sb1);           // stage s+1 lives in set 1
    if (s + 1 < nst) step(s + 1, sa1, sb1, sa0, sb0);   // stage s+2 lives in set 0
  }

  epilogue_split<TM, TN, EPI>(g, b, m0, n0, wm, wn, lr, lk, accR, accI);
}

// ---------------------------------------------------------------------------
// Direct-to-LDS variant: global_load_lds_dwordx4 moves 16-byte pieces from HBM
// straight into LDS (no VGPR staging, no ds_write).  One wave instruction fills
// a contiguous 1 KiB = 4 tile rows x 16 slots; a row's 16 pieces are placed at
// slot = piece ^ (row & 15), which keeps the 16-lane fragment reads of 16
// consecutive rows on distinct banks without row padding.  Lanes outside the
// operand read a zeroed line.
// ---------------------------------------------------------------------------
__device__ uint4 g_zero_line[4];

typedef __attribute__((address_space(3))) void lds_void_t;
typedef const __attribute__((address_space(1))) void gbl_void_t;

constexpr int kRowDl = kStageChunks * kChunkBytes;   // 256 B, unpadded

template <int ROWS>
__device__ __forceinline__ void stage_dl(uint8_t* lds_tile, const uint8_t* __restrict__ base,
                                         long row_bytes, int row0, int nrows, int kc0,
                                         int kchunks) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < ROWS / 16; ++q) {
    const int rb = 16 * q + 4 * wave;                 // first of this instruction's 4 rows
    const int r = rb + (lane >> 4);
    const int piece = (lane & 15) ^ (r & 15);
    const int chunk = kc0 + (piece >> 2);
    const int rg = row0 + r;
    const uint8_t* src = (rg < nrows && chunk < kchunks)
                             ? base + (long)rg * row_bytes + (long)chunk * kChunkBytes +
                                   (piece & 3) * 16
                             : (const uint8_t*)g_zero_line;
    __builtin_amdgcn_global_load_lds((gbl_void_t*)src, (lds_void_t*)(lds_tile + rb * kRowDl),
                                     16, 0, 0);
  }
}

template <int TM, int TN, int EPI>
__global__ void __launch_bounds__(kThreads)
cgemm_split4_dl_kernel(GemmSplit g) {
  constexpr int BM = 64 * TM, BN = 64 * TN;
  constexpr int kBuf = (BM + BN) * kRowDl;
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * kBuf];

  const int tiles_n = (g.N + BN - 1) / BN, tiles_m = (g.M + BM - 1) / BM;
  const int total = tiles_n * tiles_m * g.batch;
  const int id = xcd_remap(blockIdx.x, total);
  const int tn = id % tiles_n;
  const int tm = (id / tiles_n) % tiles_m;
  const int b = id / (tiles_n * tiles_m);
  const int m0 = tm * BM, n0 = tn * BN;

  const uint8_t* __restrict__ A = g.A + (long)b * g.a_batch_bytes;
  const uint8_t* __restrict__ B = g.B + (long)b * g.b_batch_bytes;
  const int kchunks = (int)split4_chunks(g.K);
  const int nst = (kchunks + kStageChunks - 1) / kStageChunks;
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

  stage_dl<BM>(lds, A, g.a_row_bytes, m0, g.M, 0, kchunks);
  stage_dl<BN>(lds + BM * kRowDl, B, g.b_row_bytes, n0, g.N, 0, kchunks);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  for (int s = 0; s < nst; ++s) {
    if (s + 1 < nst) {
      uint8_t* nb = lds + ((s + 1) & 1) * kBuf;
      stage_dl<BM>(nb, A, g.a_row_bytes, m0, g.M, (s + 1) * kStageChunks, kchunks);
      stage_dl<BN>(nb + BM * kRowDl, B, g.b_row_bytes, n0, g.N, (s + 1) * kStageChunks,
                   kchunks);
    }
    const uint8_t* la = lds + (s & 1) * kBuf;
    const uint8_t* lb = la + BM * kRowDl;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      u32x4_t fa[TM][4], fb[TN][4];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int r = wm + 32 * i + lr;
#pragma unroll
        for (int p = 0; p < 4; ++p)
          fa[i][p] = *(const u32x4_t*)(la + r * kRowDl + ((((2 * ks + lk) << 2) | p) ^ (r & 15)) * 16);
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int r = wn + 32 * j + lr;
#pragma unroll
        for (int p = 0; p < 4; ++p)
          fb[j][p] = *(const u32x4_t*)(lb + r * kRowDl + ((((2 * ks + lk) << 2) | p) ^ (r & 15)) * 16);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const bf16x8_t arh = as_bf(fa[i][0]), arl = as_bf(fa[i][1]);
        const bf16x8_t aih = as_bf(fa[i][2]), ail = as_bf(fa[i][3]);
        const bf16x8_t naih = as_bf(neg_bf(fa[i][2])), nail = as_bf(neg_bf(fa[i][3]));
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const bf16x8_t brh = as_bf(fb[j][0]), brl = as_bf(fb[j][1]);
          const bf16x8_t bih = as_bf(fb[j][2]), bil = as_bf(fb[j][3]);
          f32x16_t r = accR[i][j], im = accI[i][j];
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, brh, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, bih, im, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brl, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bil, im, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nail, bih, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, brh, im, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bil, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brl, im, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brh, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bih, im, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bih, r, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brh, im, 0, 0, 0);
          accR[i][j] = r;
          accI[i][j] = im;
        }
      }
    }
    // this wave's loads of stage s+1 have landed; the barrier makes every
    // wave's loads visible and retires every read of buffer s & 1
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  epilogue_split<TM, TN, EPI>(g, b, m0, n0, wm, wn, lr, lk, accR, accI);
}

// ---------------------------------------------------------------------------
// Occupancy-2 direct-to-LDS variant: 16-k stages (2 chunks, 128 B per tile
// row) in two LDS buffers = 64 KiB per workgroup, so two workgroups share a
// CU and each SIMD interleaves two waves: while one waits for its next stage
// to land, the other issues MFMAs.  Row r stores piece p at slot
// p ^ ((r >> 1) & 7): 16 consecutive rows reading the same piece cover all 16
// bank groups.
// ---------------------------------------------------------------------------
constexpr int kDeepChunks = 2;
constexpr int kRowDeep = kDeepChunks * kChunkBytes;   // 128 B

template <int ROWS>
__device__ __forceinline__ void stage_deep(uint8_t* lds_tile, const uint8_t* __restrict__ base,
                                           long row_bytes, int row0, int nrows, int kc0,
                                           int kchunks) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < ROWS / 32; ++q) {
    const int rb = 32 * q + 8 * wave;                 // first of this instruction's 8 rows
    const int r = rb + (lane >> 3);
    const int piece = (lane & 7) ^ ((r >> 1) & 7);
    const int chunk = kc0 + (piece >> 2);
    const int rg = row0 + r;
    const uint8_t* src = (rg < nrows && chunk < kchunks)
                             ? base + (long)rg * row_bytes + (long)chunk * kChunkBytes +
                                   (piece & 3) * 16
                             : (const uint8_t*)g_zero_line;
    __builtin_amdgcn_global_load_lds((gbl_void_t*)src, (lds_void_t*)(lds_tile + rb * kRowDeep),
                                     16, 0, 0);
  }
}

// Main loop of the occupancy-2 kernel for one BM x BN output tile at (m0, n0):
// accumulates sum_k A(m, k) B(n, k) into accR / accI.  `lds` is the
// workgroup's 2 x (BM + BN) x 128 B staging area.
// Bounds-free loader for tiles entirely inside both operands with an even
// number of k-chunks: each thread's source addresses are formed once per tile
// and advanced by one k-slice (128 B) per stage.
template <int ROWS>
struct FastRows {
  const uint8_t* src[ROWS / 32];
  __device__ __forceinline__ void init(const uint8_t* __restrict__ base, long row_bytes,
                                       int row0) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < ROWS / 32; ++q) {
      const int r = 32 * q + 8 * wave + (lane >> 3);
      const int piece = (lane & 7) ^ ((r >> 1) & 7);
      src[q] = base + (long)(row0 + r) * row_bytes + (piece >> 2) * kChunkBytes +
               (piece & 3) * 16;
    }
  }
  __device__ __forceinline__ void issue(uint8_t* lds_tile, int kc0) const {
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < ROWS / 32; ++q)
      __builtin_amdgcn_global_load_lds((gbl_void_t*)(src[q] + (long)kc0 * kChunkBytes),
                                       (lds_void_t*)(lds_tile + (32 * q + 8 * wave) * kRowDeep),
                                       16, 0, 0);
  }
};

template <int TM, int TN, bool FAST, bool PRIO = true>
__device__ __forceinline__ void deep_mainloop_impl(uint8_t* lds, const uint8_t* __restrict__ A,
                                                   long a_row_bytes, int M,
                                                   const uint8_t* __restrict__ B,
                                                   long b_row_bytes, int N, int K, int m0,
                                                   int n0, f32x16_t (&accR)[TM][TN],
                                                   f32x16_t (&accI)[TM][TN]) {
  constexpr int BM = 64 * TM, BN = 64 * TN;
  constexpr int kBuf = (BM + BN) * kRowDeep;
  const int kchunks = (int)split4_chunks(K);
  const int nst = (kchunks + kDeepChunks - 1) / kDeepChunks;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32 * TM, wn = (wave & 1) * 32 * TN;
  const int lr = lane & 31, lk = lane >> 5;

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) { accR[i][j][r] = 0.f; accI[i][j][r] = 0.f; }

  FastRows<BM> fa_rows;
  FastRows<BN> fb_rows;
  if (FAST) {
    fa_rows.init(A, a_row_bytes, m0);
    fb_rows.init(B, b_row_bytes, n0);
  }
  auto issue = [&](int st) {
    uint8_t* buf = lds + (st & 1) * kBuf;
    if (FAST) {
      fa_rows.issue(buf, st * kDeepChunks);
      fb_rows.issue(buf + BM * kRowDeep, st * kDeepChunks);
    } else {
      stage_deep<BM>(buf, A, a_row_bytes, m0, M, st * kDeepChunks, kchunks);
      stage_deep<BN>(buf + BM * kRowDeep, B, b_row_bytes, n0, N, st * kDeepChunks, kchunks);
    }
  };
  issue(0);

  for (int s = 0; s < nst; ++s) {
    __builtin_amdgcn_s_waitcnt(0);   // this wave's loads of stage s have landed
    __syncthreads();                 // every wave's have; stage s-1 reads retired
    const uint8_t* la = lds + (s & 1) * kBuf;
    const uint8_t* lb = la + BM * kRowDeep;
    u32x4_t fa[TM][4], fb[TN][4];
    // fragments in the order they are consumed: A row block 0, all of B, the
    // rest of A, so the first row's products wait only for their own reads
    auto read_a = [&](int i) {
      const int r = wm + 32 * i + lr;
#pragma unroll
      for (int p = 0; p < 4; ++p)
        fa[i][p] = *(const u32x4_t*)(la + r * kRowDeep + (((lk << 2) | p) ^ ((r >> 1) & 7)) * 16);
    };
    read_a(0);
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int r = wn + 32 * j + lr;
#pragma unroll
      for (int p = 0; p < 4; ++p)
        fb[j][p] = *(const u32x4_t*)(lb + r * kRowDeep + (((lk << 2) | p) ^ ((r >> 1) & 7)) * 16);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 1; i < TM; ++i) read_a(i);
    // the next stage's loads go out behind this stage's LDS reads (measured
    // 1.5-2 % faster than issuing them first at 512 -> 256)
    if (s + 1 < nst) issue(s + 1);
    // raise this wave's issue priority while it feeds the matrix cores, so the
    // SIMD's other wave does its loads and LDS reads in the gaps
    if (PRIO) __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      __builtin_amdgcn_sched_barrier(0);
      const bf16x8_t arh = as_bf(fa[i][0]), arl = as_bf(fa[i][1]);
      const bf16x8_t aih = as_bf(fa[i][2]), ail = as_bf(fa[i][3]);
      const bf16x8_t naih = as_bf(neg_bf(fa[i][2])), nail = as_bf(neg_bf(fa[i][3]));
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const bf16x8_t brh = as_bf(fb[j][0]), brl = as_bf(fb[j][1]);
        const bf16x8_t bih = as_bf(fb[j][2]), bil = as_bf(fb[j][3]);
        f32x16_t r = accR[i][j], im = accI[i][j];
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, brh, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, bih, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brl, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bil, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nail, bih, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, brh, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bil, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brl, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brh, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bih, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bih, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brh, im, 0, 0, 0);
        accR[i][j] = r;
        accI[i][j] = im;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (PRIO) __builtin_amdgcn_s_setprio(0);
  }
}

template <int TM, int TN>
__device__ __forceinline__ void deep_mainloop(uint8_t* lds, const uint8_t* __restrict__ A,
                                              long a_row_bytes, int M,
                                              const uint8_t* __restrict__ B, long b_row_bytes,
                                              int N, int K, int m0, int n0,
                                              f32x16_t (&accR)[TM][TN],
                                              f32x16_t (&accI)[TM][TN]) {
  constexpr int BM = 64 * TM, BN = 64 * TN;
  const bool interior = m0 + BM <= M && n0 + BN <= N && (K % (8 * kDeepChunks)) == 0;
  if (interior)
    deep_mainloop_impl<TM, TN, true>(lds, A, a_row_bytes, M, B, b_row_bytes, N, K, m0, n0,
                                     accR, accI);
  else
    deep_mainloop_impl<TM, TN, false>(lds, A, a_row_bytes, M, B, b_row_bytes, N, K, m0, n0,
                                      accR, accI);
}

template <int TM, int TN, int EPI>
__global__ void __launch_bounds__(kThreads, 2)
cgemm_split4_deep_kernel(GemmSplit g) {
  constexpr int BM = 64 * TM, BN = 64 * TN;
  constexpr int kBuf = (BM + BN) * kRowDeep;
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * kBuf];

  const int tiles_n = (g.N + BN - 1) / BN, tiles_m = (g.M + BM - 1) / BM;
  const int total = tiles_n * tiles_m * g.batch;
  const int id = xcd_remap(blockIdx.x, total);
  const int tn = id % tiles_n;
  const int tm = (id / tiles_n) % tiles_m;
  const int b = id / (tiles_n * tiles_m);
  const int m0 = tm * BM, n0 = tn * BN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32 * TM, wn = (wave & 1) * 32 * TN;
  const int lr = lane & 31, lk = lane >> 5;

  f32x16_t accR[TM][TN], accI[TM][TN];
  deep_mainloop<TM, TN>(lds, g.A + (long)b * g.a_batch_bytes, g.a_row_bytes, g.M,
                        g.B + (long)b * g.b_batch_bytes, g.b_row_bytes, g.N, g.K, m0, n0,
                        accR, accI);
  if constexpr (EPI == EPI_SPLIT_T || EPI == EPI_VJP_SPLIT_T) {
    if (m0 + BM <= g.M && n0 + BN <= g.N) {
      epilogue_split_t_lds<TM, TN, EPI>(g, b, m0, n0, wm, wn, lr, lk, accR, accI, lds);
      return;
    }
  }
  epilogue_split<TM, TN, EPI>(g, b, m0, n0, wm, wn, lr, lk, accR, accI);
}

// ---------------------------------------------------------------------------
// Stage 1 with the pupil generated in the loader: the A operand (pupil rows y,
// k = x) is never read from HBM.  Each thread owns one row of the 128-row A
// tile and one 8-pixel chunk of a stage's 16 k.  It loads the next stage's 8
// OPDs and 8 amplitudes (the caller's planes) before this stage's products are
// issued and turns them into split4 phasors (path, amplitude and phasor as the
// pupil kernels form them, dlx_phasor.h) after the products are issued, so the
// phasor work overlaps the matrix cores; with no tilt the path is the OPD
// itself (tx X + ty Y is +-0 and adds nothing to the phase), the loader skips
// the coordinate arithmetic; the
// four 16-byte planes go straight into the next stage's A buffer with the
// deep kernel's swizzle.  B (the transfer matrix) is staged direct-to-LDS as
// in the deep kernel; the epilogue is EPI_SPLIT_T.  Interior tiles only.
// ---------------------------------------------------------------------------
template <int TN>
__global__ void __launch_bounds__(kThreads, 2)
cgemm_pupil_kernel(GemmSplit g, PupilGen pg) {
  constexpr int TM = 2, BM = 128, BN = 64 * TN;
  constexpr int kBuf = (BM + BN) * kRowDeep;
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * kBuf];

  const int tiles_n = g.N / BN, tiles_m = g.M / BM;
  const int total = tiles_n * tiles_m * g.batch;
  const int id = xcd_remap(blockIdx.x, total);
  const int tn = id % tiles_n;
  const int tm = (id / tiles_n) % tiles_m;
  const int b = id / (tiles_n * tiles_m);
  const int m0 = tm * BM, n0 = tn * BN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32 * TM, wn = (wave & 1) * 32 * TN;
  const int lr = lane & 31, lk = lane >> 5;
  const int nst = g.K / (8 * kDeepChunks);

  f32x16_t accR[TM][TN], accI[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) { accR[i][j][r] = 0.f; accI[i][j][r] = 0.f; }

  FastRows<BN> fb_rows;
  fb_rows.init(g.B + (long)b * g.b_batch_bytes, g.b_row_bytes, n0);

  // this thread's pupil row (tile row pr) and chunk (pc) of every stage
  const int pr = threadIdx.x >> 1, pc = threadIdx.x & 1;
  const double inv_wl = pg.inv_wl[b];
  const long pix0 = (long)(m0 + pr) * pg.n + pc * 8;
  const int sw = (pr >> 1) & 7;
  const bool tilted = pg.tx != 0.0 || pg.ty != 0.0;
  const double tyY = pg.ty * pupil_coord(m0 + pr, pg.dx, pg.x0);
  float4 ro[2], ra[2];
  auto load_raw = [&](int st) {
    const long p = pix0 + st * (8 * kDeepChunks);
#pragma unroll
    for (int i = 0; i < 2; ++i)
      ro[i] = pg.opd ? *(const float4*)(pg.opd + p + 4 * i) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int i = 0; i < 2; ++i)
      ra[i] = pg.amp ? *(const float4*)(pg.amp + p + 4 * i) : make_float4(1.f, 1.f, 1.f, 1.f);
  };
  auto convert = [&](int st) {
    const float ov[8] = {ro[0].x, ro[0].y, ro[0].z, ro[0].w, ro[1].x, ro[1].y, ro[1].z, ro[1].w};
    const float av[8] = {ra[0].x, ra[0].y, ra[0].z, ra[0].w, ra[1].x, ra[1].y, ra[1].z, ra[1].w};
    const int xc = st * (8 * kDeepChunks) + pc * 8;
    float2 v[8];
    if (tilted) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const double path = pupil_path(pg.tx, pupil_coord(xc + e, pg.dx, pg.x0), tyY,
                                       (double)ov[e]);
        v[e] = pupil_phasor(path, inv_wl, 0.0, pg.scale * av[e]);
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = pupil_phasor((double)ov[e], inv_wl, 0.0, pg.scale * av[e]);
    }
    uint4 p[4];
    pack_chunk(v, p);
    uint8_t* dst = lds + (st & 1) * kBuf + pr * kRowDeep;
#pragma unroll
    for (int q = 0; q < 4; ++q) *(uint4*)(dst + ((((pc << 2) | q) ^ sw) << 4)) = p[q];
  };

  load_raw(0);
  fb_rows.issue(lds + BM * kRowDeep, 0);
  convert(0);

  for (int s = 0; s < nst; ++s) {
    __builtin_amdgcn_s_waitcnt(0);   // B of stage s landed, A of stage s written
    __syncthreads();                 // by every wave; stage s-1 reads retired
    const uint8_t* la = lds + (s & 1) * kBuf;
    const uint8_t* lb = la + BM * kRowDeep;
    u32x4_t fa[TM][4], fb[TN][4];
    auto read_a = [&](int i) {
      const int r = wm + 32 * i + lr;
#pragma unroll
      for (int p = 0; p < 4; ++p)
        fa[i][p] = *(const u32x4_t*)(la + r * kRowDeep + (((lk << 2) | p) ^ ((r >> 1) & 7)) * 16);
    };
    read_a(0);
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int r = wn + 32 * j + lr;
#pragma unroll
      for (int p = 0; p < 4; ++p)
        fb[j][p] = *(const u32x4_t*)(lb + r * kRowDeep + (((lk << 2) | p) ^ ((r >> 1) & 7)) * 16);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 1; i < TM; ++i) read_a(i);
    const bool more = s + 1 < nst;
    if (more) {
      load_raw(s + 1);
      fb_rows.issue(lds + ((s + 1) & 1) * kBuf + BM * kRowDeep, (s + 1) * kDeepChunks);
    }
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      __builtin_amdgcn_sched_barrier(0);
      const bf16x8_t arh = as_bf(fa[i][0]), arl = as_bf(fa[i][1]);
      const bf16x8_t aih = as_bf(fa[i][2]), ail = as_bf(fa[i][3]);
      const bf16x8_t naih = as_bf(neg_bf(fa[i][2])), nail = as_bf(neg_bf(fa[i][3]));
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const bf16x8_t brh = as_bf(fb[j][0]), brl = as_bf(fb[j][1]);
        const bf16x8_t bih = as_bf(fb[j][2]), bil = as_bf(fb[j][3]);
        f32x16_t r = accR[i][j], im = accI[i][j];
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, brh, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, bih, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brl, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bil, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nail, bih, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, brh, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bil, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brl, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brh, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bih, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bih, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brh, im, 0, 0, 0);
        accR[i][j] = r;
        accI[i][j] = im;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(0);
    // the next stage's pupil chunk, into the A buffer stage s-1 used
    if (more) convert(s + 1);
  }
  epilogue_split_t_lds<TM, TN>(g, b, m0, n0, wm, wn, lr, lk, accR, accI, lds);
}

// ---------------------------------------------------------------------------
// Wide variant: one 8-wave workgroup per CU on a 256 x 128 tile (waves 4 x 2,
// each 64 x 64), three 48 KiB LDS buffers (144 KiB), so the loads of stage
// s + 2 are in flight while stage s is computed.  The per-stage wait keeps the
// six newest loads of this wave (stage s + 1) outstanding.
// ---------------------------------------------------------------------------
constexpr int kWideThreads = 512;
constexpr int kWideBufs = 3;

template <int ROWS>
__device__ __forceinline__ void stage_wide(uint8_t* lds_tile, const uint8_t* __restrict__ base,
                                           long row_bytes, int row0, int nrows, int kc0,
                                           int kchunks) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < ROWS / 64; ++q) {
    const int rb = 64 * q + 8 * wave;                 // first of this instruction's 8 rows
    const int r = rb + (lane >> 3);
    const int piece = (lane & 7) ^ ((r >> 1) & 7);
    const int chunk = kc0 + (piece >> 2);
    const int rg = row0 + r;
    const uint8_t* src = (rg < nrows && chunk < kchunks)
                             ? base + (long)rg * row_bytes + (long)chunk * kChunkBytes +
                                   (piece & 3) * 16
                             : (const uint8_t*)g_zero_line;
    __builtin_amdgcn_global_load_lds((gbl_void_t*)src, (lds_void_t*)(lds_tile + rb * kRowDeep),
                                     16, 0, 0);
  }
}

template <int EPI>
__global__ void __launch_bounds__(kWideThreads, 1)
cgemm_split4_wide_kernel(GemmSplit g) {
  constexpr int TM = 2, TN = 2, BM = 256, BN = 128;
  constexpr int kBuf = (BM + BN) * kRowDeep;
  constexpr int kLoads = BM / 64 + BN / 64;           // global_load_lds per wave per stage
  static_assert(kLoads == 6, "the stage wait below encodes six loads");
  __shared__ __attribute__((aligned(1024))) uint8_t lds[kWideBufs * kBuf];

  const int tiles_n = (g.N + BN - 1) / BN, tiles_m = (g.M + BM - 1) / BM;
  const int total = tiles_n * tiles_m * g.batch;
  const int id = xcd_remap(blockIdx.x, total);
  const int tn = id % tiles_n;
  const int tm = (id / tiles_n) % tiles_m;
  const int b = id / (tiles_n * tiles_m);
  const int m0 = tm * BM, n0 = tn * BN;
  const uint8_t* __restrict__ A = g.A + (long)b * g.a_batch_bytes;
  const uint8_t* __restrict__ B = g.B + (long)b * g.b_batch_bytes;
  const int kchunks = (int)split4_chunks(g.K);
  const int nst = (kchunks + kDeepChunks - 1) / kDeepChunks;
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

  auto issue = [&](int st) {
    uint8_t* buf = lds + (st % kWideBufs) * kBuf;
    stage_wide<BM>(buf, A, g.a_row_bytes, m0, g.M, st * kDeepChunks, kchunks);
    stage_wide<BN>(buf + BM * kRowDeep, B, g.b_row_bytes, n0, g.N, st * kDeepChunks, kchunks);
  };
  issue(0);
  if (nst > 1) issue(1);

  for (int s = 0; s < nst; ++s) {
    // stage s has landed for this wave (stage s + 1 may stay in flight); the
    // barrier makes every wave's loads visible and retires stage s - 1 reads
    if (s + 1 < nst)
      asm volatile("s_waitcnt vmcnt(6)\n\ts_barrier" ::: "memory");
    else
      asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
    if (s + 2 < nst) issue(s + 2);
    const uint8_t* la = lds + (s % kWideBufs) * kBuf;
    const uint8_t* lb = la + BM * kRowDeep;
    u32x4_t fa[TM][4], fb[TN][4];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int r = wm + 32 * i + lr;
#pragma unroll
      for (int p = 0; p < 4; ++p)
        fa[i][p] = *(const u32x4_t*)(la + r * kRowDeep + (((lk << 2) | p) ^ ((r >> 1) & 7)) * 16);
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int r = wn + 32 * j + lr;
#pragma unroll
      for (int p = 0; p < 4; ++p)
        fb[j][p] = *(const u32x4_t*)(lb + r * kRowDeep + (((lk << 2) | p) ^ ((r >> 1) & 7)) * 16);
    }
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const bf16x8_t arh = as_bf(fa[i][0]), arl = as_bf(fa[i][1]);
      const bf16x8_t aih = as_bf(fa[i][2]), ail = as_bf(fa[i][3]);
      const bf16x8_t naih = as_bf(neg_bf(fa[i][2])), nail = as_bf(neg_bf(fa[i][3]));
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const bf16x8_t brh = as_bf(fb[j][0]), brl = as_bf(fb[j][1]);
        const bf16x8_t bih = as_bf(fb[j][2]), bil = as_bf(fb[j][3]);
        f32x16_t r = accR[i][j], im = accI[i][j];
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, brh, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, bih, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brl, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bil, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nail, bih, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, brh, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bil, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brl, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brh, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bih, im, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bih, r, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brh, im, 0, 0, 0);
        accR[i][j] = r;
        accI[i][j] = im;
      }
    }
  }
  epilogue_split<TM, TN, EPI>(g, b, m0, n0, wm, wn, lr, lk, accR, accI);
}

// ---------------------------------------------------------------------------
// Grouped persistent variant for uneven batches: one launch covers the
// 128 x 128 tiles of many problems of different shapes.  The tile list is cut
// into 8 contiguous ranges, one per XCD (hardware workgroup w runs on XCD
// w % 8), so all tiles of a problem share one L2; the workgroups of an XCD
// stride over its range.  Problems are ordered by the host (largest first
// within each range), and every workgroup leaves its loop after its range.
// ---------------------------------------------------------------------------
template <int EPI, int TN>
__global__ void __launch_bounds__(kThreads, 2)
cgemm_split4_group_kernel(GroupLaunch gl) {
  constexpr int TM = 2, BM = 128, BN = 64 * TN;
  constexpr int kBuf = (BM + BN) * kRowDeep;
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * kBuf];
  const int x = blockIdx.x % kXcds;
  const int per = gridDim.x / kXcds;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32 * TM, wn = (wave & 1) * 32 * TN;
  const int lr = lane & 31, lk = lane >> 5;
  const int t_end = gl.xcd_start[x + 1];
  for (int t = gl.xcd_start[x] + (int)blockIdx.x / kXcds; t < t_end; t += per) {
    // problem owning tile t: the last one whose tile_start <= t
    int lo = 0, hi = gl.nprobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (gl.probs[mid].tile_start <= t) lo = mid; else hi = mid - 1;
    }
    const GroupProblem& pr = gl.probs[lo];
    const int local = t - pr.tile_start;
    const int tn = local % pr.tiles_n, tm = local / pr.tiles_n;
    const int m0 = tm * BM, n0 = tn * BN;
    __syncthreads();   // the previous tile's LDS reads have retired
    f32x16_t accR[TM][TN], accI[TM][TN];
    // the generic loader: the persistent loop's own state leaves no registers
    // for the bounds-free one
    deep_mainloop_impl<TM, TN, false, false>(lds, pr.A, pr.a_row_bytes, pr.M, pr.B,
                                             pr.b_row_bytes, pr.N, pr.K, m0, n0, accR, accI);
    GemmSplit v{};
    v.C = pr.C; v.ldc = pr.ldc; v.c_batch = 0;
    v.M = pr.M; v.N = pr.N; v.K = pr.K; v.batch = 1; v.epilogue = EPI;
    v.alpha[0] = pr.alpha;
    if (EPI == EPI_SPLIT_T && m0 + BM <= pr.M && n0 + BN <= pr.N)
      epilogue_split_t_lds<TM, TN>(v, 0, m0, n0, wm, wn, lr, lk, accR, accI, lds);
    else
      epilogue_split<TM, TN, EPI>(v, 0, m0, n0, wm, wn, lr, lk, accR, accI);
  }
}

// Staging variants: 0 = deep direct-to-LDS (default), 1 = direct-to-LDS with
// two 32-k buffers, 2 = register staging.  DLUX_AMD_GEMM_STAGING=deep|dl|regs
// selects one for comparison runs.
int staging_mode() {
  static const int mode = [] {
    const char* v = getenv("DLUX_AMD_GEMM_STAGING");
    if (v && std::string(v) == "wide") return 3;
    if (v && std::string(v) == "regs") return 2;
    if (v && std::string(v) == "dl") return 1;
    return 0;
  }();
  return mode;
}

template <int TM, int TN, int EPI>
void launch_one(const GemmSplit& g, int tiles, int mode, hipStream_t s) {
  if (mode == 0)
    hipLaunchKernelGGL((cgemm_split4_deep_kernel<TM, TN, EPI>), dim3(tiles), dim3(kThreads), 0,
                       s, g);
  else if (mode == 1)
    hipLaunchKernelGGL((cgemm_split4_dl_kernel<TM, TN, EPI>), dim3(tiles), dim3(kThreads), 0, s,
                       g);
  else
    hipLaunchKernelGGL((cgemm_split4_kernel<TM, TN, EPI>), dim3(tiles), dim3(kThreads), 0, s, g);
}

template <int TM, int TN>
hipError_t dispatch_split(const GemmSplit& g, hipStream_t s, int mode) {
  const int tiles = ((g.N + 64 * TN - 1) / (64 * TN)) * ((g.M + 64 * TM - 1) / (64 * TM)) *
                    g.batch;
  switch (g.epilogue) {
    case EPI_STORE: launch_one<TM, TN, EPI_STORE>(g, tiles, mode, s); break;
    case EPI_SPLIT_T: launch_one<TM, TN, EPI_SPLIT_T>(g, tiles, mode, s); break;
    case EPI_VJP_SPLIT_T: launch_one<TM, TN, EPI_VJP_SPLIT_T>(g, tiles, mode, s); break;
    case EPI_GRAD: launch_one<TM, TN, EPI_GRAD>(g, tiles, mode, s); break;
    default: launch_one<TM, TN, EPI_INTENSITY>(g, tiles, mode, s); break;
  }
  return hipGetLastError();
}

}  // namespace

template <int TN>
hipError_t launch_group_tn(const GroupLaunch& gl, int epilogue, const dim3& grid, hipStream_t s) {
  const dim3 block(kThreads);
  switch (epilogue) {
    case EPI_SPLIT_T: hipLaunchKernelGGL((cgemm_split4_group_kernel<EPI_SPLIT_T, TN>), grid, block, 0, s, gl); break;
    case EPI_INTENSITY: hipLaunchKernelGGL((cgemm_split4_group_kernel<EPI_INTENSITY, TN>), grid, block, 0, s, gl); break;
    case EPI_STORE: hipLaunchKernelGGL((cgemm_split4_group_kernel<EPI_STORE, TN>), grid, block, 0, s, gl); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_cgemm_split4_group(const GroupLaunch& gl, int epilogue, hipStream_t s) {
  const int total = gl.xcd_start[8];
  if (total <= 0 || gl.nprobs <= 0) return hipSuccess;
  // two workgroups per CU, 32 CUs per XCD; no more per XCD than its largest range
  int most = 0;
  for (int x = 0; x < 8; ++x) {
    const int n = gl.xcd_start[x + 1] - gl.xcd_start[x];
    most = n > most ? n : most;
  }
  const int per = most < 64 ? most : 64;
  const dim3 grid(8 * per);
  return gl.tile_n == 64 ? launch_group_tn<1>(gl, epilogue, grid, s)
                         : launch_group_tn<2>(gl, epilogue, grid, s);
}

namespace {
// Tile width of the pupil-prologue kernel for this launch (128 or 64), or 0
// when the shape is not covered: interior 128-row tiles, whole 16-k stages,
// 16-byte aligned planes, at most two tile columns, and (as for the deep
// kernel) enough tiles for two workgroups per CU.
int pupil_tile_n(const GemmSplit& g, const PupilGen& pg) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0 || g.batch > GemmSplit::kMaxBatch) return 0;
  const bool aligned =
      ((reinterpret_cast<uintptr_t>(pg.opd) | reinterpret_cast<uintptr_t>(pg.amp)) & 15) == 0;
  if (g.epilogue != EPI_SPLIT_T || g.K != g.M || pg.n != g.M || g.M % 128 != 0 ||
      g.K % (8 * kDeepChunks) != 0 || !aligned || g.b_row_bytes != split4_row_bytes(g.K) ||
      g.ldc != split4_row_bytes(g.M))
    return 0;
  // every pupil element is formed once per tile column: at most two (beyond
  // that the phasor work costs more than the pupil array's write and read)
  const long tiles128 = (long)(g.M / 128) * (g.N / 128) * g.batch;
  const long tiles64 = (long)(g.M / 128) * (g.N / 64) * g.batch;
  if (g.N % 128 == 0 && g.N <= 256 && tiles128 >= 512) return 128;
  if (g.N % 64 == 0 && g.N <= 128 && tiles64 >= 512) return 64;
  return 0;
}
}  // namespace

bool cgemm_pupil_supported(const GemmSplit& g, const PupilGen& pg) {
  return pupil_tile_n(g, pg) != 0;
}

hipError_t launch_cgemm_pupil(const GemmSplit& g, const PupilGen& pg, hipStream_t s) {
  const int tn = pupil_tile_n(g, pg);
  if (tn == 0) return hipErrorNotSupported;
  const unsigned tiles = (unsigned)((g.M / 128) * (g.N / tn) * g.batch);
  if (tn == 128)
    hipLaunchKernelGGL(cgemm_pupil_kernel<2>, dim3(tiles), dim3(kThreads), 0, s, g, pg);
  else
    hipLaunchKernelGGL(cgemm_pupil_kernel<1>, dim3(tiles), dim3(kThreads), 0, s, g, pg);
  return hipGetLastError();
}

template <int EPI>
void launch_wide(const GemmSplit& g, hipStream_t s) {
  const int tiles = ((g.N + 127) / 128) * ((g.M + 255) / 256) * g.batch;
  hipLaunchKernelGGL(cgemm_split4_wide_kernel<EPI>, dim3(tiles), dim3(kWideThreads), 0, s, g);
}

hipError_t launch_cgemm_split4(const GemmSplit& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.batch <= 0) return hipSuccess;
  if (g.batch > GemmSplit::kMaxBatch) return hipErrorInvalidValue;
  // one 256 x 128 tile per CU: only when there are enough tiles for every CU
  const long tiles_wide = (long)((g.N + 127) / 128) * ((g.M + 255) / 256) * g.batch;
  if (staging_mode() == 3 && g.M > 128 && g.N > 64 && tiles_wide >= 256) {
    switch (g.epilogue) {
      case EPI_STORE: launch_wide<EPI_STORE>(g, s); break;
      case EPI_SPLIT_T: launch_wide<EPI_SPLIT_T>(g, s); break;
      case EPI_VJP_SPLIT_T: launch_wide<EPI_VJP_SPLIT_T>(g, s); break;
      case EPI_GRAD: launch_wide<EPI_GRAD>(g, s); break;
      default: launch_wide<EPI_INTENSITY>(g, s); break;
    }
    return hipGetLastError();
  }
  // 128 x 128 tiles: the occupancy-2 kernel when they give two workgroups per
  // CU (256 CUs); else 128 x 64 tiles on it when those do (7 % faster than
  // one 128 x 128 workgroup per CU for the 64-lambda 256^2 output stage); else
  // the one-workgroup-per-CU direct-to-LDS kernel; 64 x 64 tiles for small
  // problems.
  const long tiles128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128) * g.batch;
  int mode = staging_mode();
  if (mode == 3) mode = 0;
  if (mode == 0 && tiles128 < 512) {
    const long tiles_2x1 = (long)((g.M + 127) / 128) * ((g.N + 63) / 64) * g.batch;
    if (g.M > 64 && g.N > 64 && tiles_2x1 >= 512) return dispatch_split<2, 1>(g, s, 0);
    mode = 1;
  }
  if (g.M <= 64 || g.N <= 64 || tiles128 < 128) return dispatch_split<1, 1>(g, s, mode);
  return dispatch_split<2, 2>(g, s, mode);
}

}  // namespace dlx
