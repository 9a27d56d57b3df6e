// Linear canonical transforms (Collins integrals) on gfx950: the transfer
// matrices of the matrix form and the chirps of the FFT form.  The transforms
// themselves run on the MFT's complex GEMMs and on the centred FFT.
//
// Per axis, with x on the input grid and x' on the output grid,
//   u_out(x') = sqrt(dx dx' / (i lambda B)) sum_x u_in(x)
//               exp(i pi / (lambda B) (A x^2 - 2 x x' + D x'^2))
// so a 2-D field maps as u_out = K u_in K^T.  All phases are accumulated in
// FP64 turns and reduced to [-1/2, 1/2] before the hardware sine / cosine.
#include "dlx_internal.h"
#include "dlx_split.h"

namespace dlx {

__device__ __forceinline__ float2 lct_element(const LctMatChunk& c, int b, double x,
                                              double xp) {
  // each term reduced on its own: |ci2 x^2| can reach 1e6 turns
  double t1 = c.ci2[b] * x * x, t2 = c.co2[b] * xp * xp, t3 = c.cx[b] * x * xp;
  t1 -= rint(t1);
  t2 -= rint(t2);
  t3 -= rint(t3);
  double t = t1 + t2 + t3 + c.c0[b];
  t -= rint(t);
  return make_float2(__builtin_amdgcn_cosf((float)t), __builtin_amdgcn_sinf((float)t));
}

// K[b][l][k] = exp(2 pi i (ci2 x_k^2 + co2 x'_l^2 + cx x_k x'_l + c0)),
// x_k = x0_in + k dx_in, x'_l = x0_out + l dx_out; stored N_out x N_in.
__global__ void __launch_bounds__(256)
lct_matrix_kernel(LctMatChunk c, int n_in, int n_out, double x0_in, double dx_in,
                  double x0_out, double dx_out, float2* __restrict__ out, long bstride) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  const int b = blockIdx.z;
  if (k >= n_in) return;
  const double x = x0_in + (double)k * dx_in, xp = x0_out + (double)l * dx_out;
  out[(long)b * bstride + (long)l * n_in + k] = lct_element(c, b, x, xp);
}

// The same matrices written in split4 (rows l, k = n_in): a quad of lanes per
// 8-element chunk, quad-cooperative 16-byte stores; padding k is zero.
__global__ void __launch_bounds__(256)
lct_matrix_split4_kernel(LctMatChunk c, int n_in, int n_out, double x0_in, double dx_in,
                         double x0_out, double dx_out, uint8_t* __restrict__ out,
                         long batch_bytes) {
  const long kc = split4_chunks(n_in);
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (t >= (long)n_out * kc * 4) return;   // whole quads exit together
  const int q = (int)(t & 3);
  const long cidx = t >> 2;
  const int l = (int)(cidx / kc), ch = (int)(cidx - (long)l * kc);
  const double xp = x0_out + (double)l * dx_out;
  float2 v[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int k = ch * 8 + 2 * q + e;
    v[e] = k < n_in ? lct_element(c, b, x0_in + (double)k * dx_in, xp) : make_float2(0.f, 0.f);
  }
  store_chunk_quad(out + (long)b * batch_bytes, l, split4_row_bytes(n_in), ch, q, v[0], v[1]);
}

hipError_t launch_lct_matrix_split4(const LctMatChunk& c, int nb, int n_in, int n_out,
                                    double x0_in, double dx_in, double x0_out, double dx_out,
                                    uint8_t* out, long batch_bytes, hipStream_t s) {
  const long work = (long)n_out * split4_chunks(n_in) * 4;
  hipLaunchKernelGGL(lct_matrix_split4_kernel, dim3((unsigned)((work + 255) / 256), nb),
                     dim3(256), 0, s, c, n_in, n_out, x0_in, dx_in, x0_out, dx_out, out,
                     batch_bytes);
  return hipGetLastError();
}

hipError_t launch_lct_matrix(const LctMatChunk& c, int nb, int n_in, int n_out, double x0_in,
                             double dx_in, double x0_out, double dx_out, float2* out,
                             long bstride, hipStream_t s) {
  dim3 grid((n_in + 255) / 256, n_out, nb);
  hipLaunchKernelGGL(lct_matrix_kernel, grid, dim3(256), 0, s, c, n_in, n_out, x0_in, dx_in,
                     x0_out, dx_out, out, bstride);
  return hipGetLastError();
}

// Separable chirp: out[b](i, j) = in[b](i, j) exp(2 pi i (T_b(x_i) + T_b(x_j))),
// T_b(x) = q2 x^2 + q1 x + q0, x_k = x0_b + k dx_b.  In place when in == out.
__global__ void __launch_bounds__(256)
lct_chirp_kernel(LctChirpChunk c, int n, const float2* in, float2* out, long npix) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (p >= npix) return;
  const int i = (int)(p / n), j = (int)(p - (long)i * n);
  const double xi = c.x0[b] + (double)i * c.dx[b], xj = c.x0[b] + (double)j * c.dx[b];
  double a = c.q2[b] * xi * xi, bq = c.q2[b] * xj * xj;
  a -= rint(a);
  bq -= rint(bq);
  double l = c.q1[b] * (xi + xj);
  l -= rint(l);
  double t = a + bq + l + 2.0 * c.q0[b];
  t -= rint(t);
  const float s = __builtin_amdgcn_sinf((float)t), co = __builtin_amdgcn_cosf((float)t);
  const float2 e = in[(long)b * npix + p];
  out[(long)b * npix + p] = make_float2(e.x * co - e.y * s, e.x * s + e.y * co);
}

hipError_t launch_lct_chirp(const LctChirpChunk& c, int nb, int n, const float2* in,
                            float2* out, hipStream_t s) {
  const long npix = (long)n * n;
  dim3 grid((unsigned)((npix + 255) / 256), nb);
  hipLaunchKernelGGL(lct_chirp_kernel, grid, dim3(256), 0, s, c, n, in, out, npix);
  return hipGetLastError();
}

}  // namespace dlx
