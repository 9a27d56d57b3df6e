// "split4" operand format of the split-BF16x3 contractions.
//
// A complex matrix R x K (row-major, k contiguous) is stored as rows of
// Kc = ceil(K / 8) chunks; a chunk holds 8 consecutive k of four bf16 planes
//   plane 0 = bf16(re), 1 = bf16(re - plane0), 2 = bf16(im), 3 = bf16(im - plane2)
// as 4 x 16 bytes (64 B per chunk), zero beyond K.  Every producer of a
// contraction operand (transfer-matrix generation, pupil synthesis, the
// stage-1 epilogue, the complex64 converter) writes this format directly, so
// the GEMM stages 16-byte pieces from HBM to LDS with no conversion work.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dlx {

constexpr int kChunk = 8;          // k per chunk
constexpr int kChunkBytes = 64;    // 4 planes x 8 bf16

__host__ __device__ inline long split4_chunks(long k) { return (k + kChunk - 1) / kChunk; }
__host__ __device__ inline long split4_row_bytes(long k) { return split4_chunks(k) * kChunkBytes; }

typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2_split_t;
typedef __attribute__((__vector_size__(2 * sizeof(float)))) float f32x2_split_t;

// Two floats -> two bf16 rounded to nearest even, packed (x0 in the low half):
// one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t bf16x2_rne(float x0, float x1) {
  const f32x2_split_t v = {x0, x1};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_split_t));
}

// Split of two values: x_i = hi_i + lo_i + O(2^-17 |x_i|), each pair packed.
// The residual is formed from the rounded x_i: contraction is off, so a
// product feeding x_i is never fused into the subtraction, and every producer
// of the same values writes the same bits whatever the code around it.
__device__ __forceinline__ void split2_bf16(float x0, float x1, uint32_t& hi, uint32_t& lo) {
#pragma clang fp contract(off)
  hi = bf16x2_rne(x0, x1);
  lo = bf16x2_rne(x0 - __uint_as_float(hi << 16), x1 - __uint_as_float(hi & 0xFFFF0000u));
}

// Pack 8 complex values (one chunk) into its four 16-byte planes.
__device__ __forceinline__ void pack_chunk(const float2 (&v)[8], uint4 (&p)[4]) {
  uint32_t rh[4], rl[4], ih[4], il[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    split2_bf16(v[2 * i].x, v[2 * i + 1].x, rh[i], rl[i]);
    split2_bf16(v[2 * i].y, v[2 * i + 1].y, ih[i], il[i]);
  }
  p[0] = make_uint4(rh[0], rh[1], rh[2], rh[3]);
  p[1] = make_uint4(rl[0], rl[1], rl[2], rl[3]);
  p[2] = make_uint4(ih[0], ih[1], ih[2], ih[3]);
  p[3] = make_uint4(il[0], il[1], il[2], il[3]);
}

// Store one chunk (8 complex values at k = 8*chunk .. 8*chunk+7 of `row`).
__device__ __forceinline__ void store_chunk(uint8_t* base, long row, long row_bytes, long chunk,
                                            const float2 (&v)[8]) {
  uint4 p[4];
  pack_chunk(v, p);
  uint4* dst = (uint4*)(base + row * row_bytes + chunk * kChunkBytes);
  dst[0] = p[0];
  dst[1] = p[1];
  dst[2] = p[2];
  dst[3] = p[3];
}

// Quad-cooperative chunk store: the four lanes q = 0..3 of a quad hold
// elements 2q and 2q+1 of one chunk (v0, v1).  Lane q receives plane q of the
// chunk by a 4 x 4 exchange inside the quad (DPP quad permutations: round m
// reads lane q ^ m) and stores it as one 16-byte write, so the lanes of a
// wave write consecutive 16-byte pieces.  All four lanes must be active.
template <int M>
__device__ __forceinline__ uint32_t quad_xor(uint32_t v) {
  // quad_perm [q ^ M for q in 0..3]
  constexpr int ctrl = M == 1 ? 0xB1 : (M == 2 ? 0x4E : 0x1B);
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xF, 0xF, false);
}

// a[i] for a lane-dependent i, as selects on values: an array indexed by a
// lane-dependent value is placed in LDS by the compiler (a per-lane copy,
// written and read back with bank conflicts)
__device__ __forceinline__ uint32_t pick4(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3,
                                          int i) {
  const uint32_t lo = (i & 1) ? a1 : a0, hi = (i & 1) ? a3 : a2;
  return (i & 2) ? hi : lo;
}

__device__ __forceinline__ void store_chunk_quad(uint8_t* base, long row, long row_bytes,
                                                 long chunk, int q, float2 v0, float2 v1) {
  uint32_t w0, w1, w2, w3;
  split2_bf16(v0.x, v1.x, w0, w1);
  split2_bf16(v0.y, v1.y, w2, w3);
  // g_m = the word lane q ^ m holds for plane q
  const uint32_t g0 = pick4(w0, w1, w2, w3, q);
  const uint32_t g1 = quad_xor<1>(pick4(w0, w1, w2, w3, q ^ 1));
  const uint32_t g2 = quad_xor<2>(pick4(w0, w1, w2, w3, q ^ 2));
  const uint32_t g3 = quad_xor<3>(pick4(w0, w1, w2, w3, q ^ 3));
  // word j of plane q comes from lane j = q ^ m
  const uint4 o = make_uint4(pick4(g0, g1, g2, g3, q), pick4(g0, g1, g2, g3, q ^ 1),
                             pick4(g0, g1, g2, g3, q ^ 2), pick4(g0, g1, g2, g3, q ^ 3));
  *(uint4*)(base + row * row_bytes + chunk * kChunkBytes + q * 16) = o;
}

}  // namespace dlx
