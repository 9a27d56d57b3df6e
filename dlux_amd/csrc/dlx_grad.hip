// Backward-pass kernels of the polychromatic PSF (gfx950).
//
// The adjoint contractions themselves run on the complex GEMMs (dlx_gemm*.hip)
// with gradient epilogues; this file holds the two reductions around them:
//
//   grad_reduce:    dL/dOPD(p) = sum_l k_l h_l(p),  dL/dphase(p) = sum_l h_l(p)
//                   over the per-wavelength planes h_l = Im(conj(P_l) Pbar_l)
//                   written by the last adjoint GEMM (k_l = 2 pi / lambda_l);
//   project_basis:  dL/dc[s, j] = sum_p basis[j, p] g[s, p]
//                   (the transpose of eval_basis), deterministic two-pass sum.
//
// Both are HBM-bound streaming passes: float4 loads, one pass over the planes.
#include "dlx_internal.h"

namespace dlx {
namespace {

__global__ void __launch_bounds__(256)
grad_reduce_kernel(GradChunk c, int nb, long npix, const float* __restrict__ h,
                   int accumulate, int vec_ok, float* __restrict__ g_opd,
                   float* __restrict__ g_phase) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;   // float4 index
  const long p = 4 * q;
  if (p >= npix) return;
  // float4 accesses need every plane 16-byte aligned (npix % 4 == 0)
  const bool full = vec_ok && p + 4 <= npix;
  float4 so = make_float4(0.f, 0.f, 0.f, 0.f), sp = so;
  for (int b = 0; b < nb; ++b) {
    const float* hb = h + (long)b * npix + p;
    float4 v;
    if (full) {
      v = *(const float4*)hb;
    } else {
      v.x = hb[0];
      v.y = p + 1 < npix ? hb[1] : 0.f;
      v.z = p + 2 < npix ? hb[2] : 0.f;
      v.w = p + 3 < npix ? hb[3] : 0.f;
    }
    const float k = c.kw[b];
    so.x += k * v.x; so.y += k * v.y; so.z += k * v.z; so.w += k * v.w;
    sp.x += v.x; sp.y += v.y; sp.z += v.z; sp.w += v.w;
  }
  const float ov[4] = {so.x, so.y, so.z, so.w}, pv[4] = {sp.x, sp.y, sp.z, sp.w};
  if (full) {
    if (g_opd) {
      float4 o = so;
      if (accumulate) {
        const float4 a = *(const float4*)(g_opd + p);
        o.x += a.x; o.y += a.y; o.z += a.z; o.w += a.w;
      }
      *(float4*)(g_opd + p) = o;
    }
    if (g_phase) {
      float4 o = sp;
      if (accumulate) {
        const float4 a = *(const float4*)(g_phase + p);
        o.x += a.x; o.y += a.y; o.z += a.z; o.w += a.w;
      }
      *(float4*)(g_phase + p) = o;
    }
  } else {
    for (int e = 0; e < 4 && p + e < npix; ++e) {
      if (g_opd) g_opd[p + e] = ov[e] + (accumulate ? g_opd[p + e] : 0.f);
      if (g_phase) g_phase[p + e] = pv[e] + (accumulate ? g_phase[p + e] : 0.f);
    }
  }
}

// Pass 1: partial[(s * nbasis + j) * S + part] = sum over the part's pixel range.
constexpr int kProjThreads = 256;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__global__ void __launch_bounds__(kProjThreads)
project_partial_kernel(int nbasis, long npix, long span, const float* __restrict__ basis,
                       const float* __restrict__ g, int vec_ok, float* __restrict__ partial) {
  const int part = blockIdx.x, j = blockIdx.y, set = blockIdx.z;
  const int nparts = gridDim.x;
  const long p0 = (long)part * span;
  const long p1 = p0 + span < npix ? p0 + span : npix;
  const float* bj = basis + (long)j * npix;
  const float* gs = g + (long)set * npix;
  float acc = 0.f;
  if (vec_ok) {
    // p0 and npix are multiples of 4 here (span is)
    for (long p = p0 + 4 * threadIdx.x; p < p1; p += 4 * kProjThreads) {
      const float4 a = *(const float4*)(bj + p);
      const float4 b = *(const float4*)(gs + p);
      acc += a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
    }
  } else {
    for (long p = p0 + threadIdx.x; p < p1; p += kProjThreads) acc += bj[p] * gs[p];
  }
  __shared__ float red[kProjThreads / 64];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < kProjThreads / 64; ++w) t += red[w];
    partial[((long)set * nbasis + j) * nparts + part] = t;
  }
}

// Pass 2: out[s, j] = sum of the parts, in a fixed order.
__global__ void __launch_bounds__(64)
project_final_kernel(int total, int nparts, const float* __restrict__ partial,
                     float* __restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= total) return;
  float t = 0.f;
  for (int q = 0; q < nparts; ++q) t += partial[(long)i * nparts + q];
  out[i] = t;
}

}  // namespace

hipError_t launch_grad_reduce(const GradChunk& c, int nb, long npix, const float* h,
                              int accumulate, float* g_opd, float* g_phase, hipStream_t s) {
  const long quads = (npix + 3) / 4;
  const int vec_ok = (npix % 4 == 0) && ((uintptr_t)h % 16 == 0) &&
                     ((uintptr_t)g_opd % 16 == 0) && ((uintptr_t)g_phase % 16 == 0);
  hipLaunchKernelGGL(grad_reduce_kernel, dim3((unsigned)((quads + 255) / 256)), dim3(256), 0, s,
                     c, nb, npix, h, accumulate, vec_ok, g_opd, g_phase);
  return hipGetLastError();
}

int project_parts(long npix) {
  // about 8 float4 per thread per part; at most 64 parts
  const long per = 8L * 4 * kProjThreads;
  long parts = (npix + per - 1) / per;
  if (parts < 1) parts = 1;
  if (parts > 64) parts = 64;
  return (int)parts;
}

hipError_t launch_project_basis(int nbasis, long npix, int nsets, const float* basis,
                                const float* g, float* partial, float* out, hipStream_t s) {
  // npix == 0: no parts, the final pass writes zeros
  const int parts = npix > 0 ? project_parts(npix) : 0;
  if (parts > 0) {
    long span = (npix + parts - 1) / parts;
    span = (span + 3) & ~3L;
    const int vec_ok =
        (npix % 4 == 0) && ((uintptr_t)basis % 16 == 0) && ((uintptr_t)g % 16 == 0);
    hipLaunchKernelGGL(project_partial_kernel, dim3(parts, nbasis, nsets), dim3(kProjThreads),
                       0, s, nbasis, npix, span, basis, g, vec_ok, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int total = nbasis * nsets;
  hipLaunchKernelGGL(project_final_kernel, dim3((total + 63) / 64), dim3(64), 0, s, total,
                     parts, partial, out);
  return hipGetLastError();
}

}  // namespace dlx
