// Transfer-matrix generation and the lambda-independent / elementwise passes
// of the hot path (gfx950): pupil-field synthesis, basis evaluation, power and
// intensity reductions.  All are HBM-bound; they are written for coalesced
// 8/16-byte accesses with one pixel (or four) per lane.
#include "dlx_internal.h"

namespace dlx {

// ---------------------------------------------------------------------------
// Transfer matrix  W[b][l][k] = amp_b * exp(sign * 2 pi i * a_k * b_l)
// (utils/propagation.py:109-127, stored transposed: N_out x N_in).
// The product a_k*b_l is formed in FP64 and reduced to [-1/2, 1/2] turns, so
// the FP32 sincospi sees an argument with an absolute error of ~1e-17 turns.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
dft_matrix_kernel(DftChunk c, int n_in, int n_out, int inverse, float2* __restrict__ out,
                  long bstride) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  const int b = blockIdx.z;
  if (k >= n_in) return;
  const double a = (double)k * c.a_scale[b] + c.a_off[b];
  const double bb = (double)l * c.b_scale[b] + c.b_off[b];
  double t = a * bb;
  t -= rint(t);
  float s, co;
  sincospif(2.0f * (float)t, &s, &co);
  const float amp = (float)c.amp[b];
  out[(long)b * bstride + (long)l * n_in + k] = make_float2(amp * co, inverse ? amp * s : -amp * s);
}

hipError_t launch_dft_chunk(const DftChunk& c, int nb, int n_in, int n_out, int inverse,
                            float2* out, long bstride, hipStream_t s) {
  dim3 grid((n_in + 255) / 256, n_out, nb);
  hipLaunchKernelGGL(dft_matrix_kernel, grid, dim3(256), 0, s, c, n_in, n_out, inverse, out,
                     bstride);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pupil field for a chunk of wavelengths:
//   out[b](y,x) = s * amp * exp(2 pi i * ((opd + tx X + ty Y)/lambda_b + phase/(2 pi)))
//                 [* in[b](y,x)]
// The argument is accumulated in FP64 turns and reduced before the FP32 sincospi.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
pupil_field_kernel(PupilChunk c, int n, double dx, double x0, double tx, double ty,
                   const float* __restrict__ amp, const float* __restrict__ opd,
                   const float* __restrict__ phase, float scale,
                   const float2* __restrict__ in, float2* __restrict__ out, long npix) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (p >= npix) return;
  const int y = (int)(p / n), x = (int)(p - (long)y * n);
  const double X = x0 + (double)x * dx, Y = x0 + (double)y * dx;
  double path = tx * X + ty * Y;
  if (opd) path += (double)opd[p];
  double t = path * c.inv_wl[b];
  if (phase) t += (double)phase[p] * 0.15915494309189535;  // 1 / (2 pi)
  t -= rint(t);
  float s, co;
  sincospif(2.0f * (float)t, &s, &co);
  const float a = scale * (amp ? amp[p] : 1.0f);
  float2 v = make_float2(a * co, a * s);
  if (in) {
    const float2 e = in[(long)b * npix + p];
    v = make_float2(e.x * v.x - e.y * v.y, e.x * v.y + e.y * v.x);
  }
  out[(long)b * npix + p] = v;
}

hipError_t launch_pupil_chunk(const PupilChunk& c, int nb, int n, double dx, double x0,
                              double tx, double ty, const float* amp, const float* opd,
                              const float* phase, float scale, const float2* in,
                              float2* out, hipStream_t s) {
  const long npix = (long)n * n;
  dim3 grid((unsigned)((npix + 255) / 256), nb);
  hipLaunchKernelGGL(pupil_field_kernel, grid, dim3(256), 0, s, c, n, dx, x0, tx, ty, amp,
                     opd, phase, scale, in, out, npix);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Basis evaluation: out[s][p] = sum_j coeffs[s][j] * basis[j][p] (+ base[p]).
// One lane per 4 pixels (16-byte loads); the basis is streamed once for up to
// kSets coefficient sets (utils/math.py:177-196 tensordot).
// ---------------------------------------------------------------------------
constexpr int kSets = 8;

__global__ void __launch_bounds__(256)
eval_basis_kernel(int nbasis, long npix, int nsets, const float* __restrict__ basis,
                  const float* __restrict__ coeffs, const float* __restrict__ base,
                  float* __restrict__ out, int vec_ok) {
  const long p4 = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (p4 >= npix) return;
  const bool full = vec_ok && (p4 + 4 <= npix);
  float acc[kSets][4];
#pragma unroll
  for (int s = 0; s < kSets; ++s)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[s][q] = 0.f;
  for (int j = 0; j < nbasis; ++j) {
    float v[4];
    const float* row = basis + (long)j * npix + p4;
    if (full) {
      const float4 t = *(const float4*)row;
      v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = (p4 + q < npix) ? row[q] : 0.f;
    }
#pragma unroll
    for (int s = 0; s < kSets; ++s) {
      if (s < nsets) {
        const float cj = coeffs[(long)s * nbasis + j];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[s][q] = fmaf(cj, v[q], acc[s][q]);
      }
    }
  }
#pragma unroll
  for (int s = 0; s < kSets; ++s) {
    if (s >= nsets) break;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (p4 + q < npix) {
        float r = acc[s][q];
        if (base) r += base[p4 + q];
        out[(long)s * npix + p4 + q] = r;
      }
    }
  }
}

hipError_t launch_eval_basis(int nbasis, long npix, int nsets, const float* basis,
                             const float* coeffs, const float* base, float* out,
                             hipStream_t s) {
  // 16-byte loads need every basis row 16-byte aligned.
  const int vec_ok = ((npix & 3) == 0) && ((reinterpret_cast<uintptr_t>(basis) & 15) == 0);
  for (int s0 = 0; s0 < nsets; s0 += kSets) {
    const int ns = nsets - s0 < kSets ? nsets - s0 : kSets;
    const long lanes = (npix + 3) / 4;
    hipLaunchKernelGGL(eval_basis_kernel, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0,
                       s, nbasis, npix, ns, basis, coeffs + (long)s0 * nbasis, base,
                       out + (long)s0 * npix, vec_ok);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// Reductions.  One workgroup per instance sums in a fixed order, so results
// are bit-reproducible (no float atomics).
// ---------------------------------------------------------------------------
constexpr int kRedThreads = 1024;

__global__ void __launch_bounds__(kRedThreads)
power_kernel(long npix, const float2* __restrict__ f, float* __restrict__ out) {
  __shared__ float red[kRedThreads / 64];
  const int b = blockIdx.x;
  const float2* fb = f + (long)b * npix;
  float acc = 0.f;
  for (long p = threadIdx.x; p < npix; p += kRedThreads) {
    const float2 e = fb[p];
    acc = fmaf(e.x, e.x, fmaf(e.y, e.y, acc));
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < kRedThreads / 64; ++i) s += red[i];
    out[b] = s;
  }
}

hipError_t launch_power(long npix, int batch, const float2* f, float* out, hipStream_t s) {
  hipLaunchKernelGGL(power_kernel, dim3(batch), dim3(kRedThreads), 0, s, npix, f, out);
  return hipGetLastError();
}

// |E|^2 per pixel, or summed over the batch in index order (deterministic).
__global__ void __launch_bounds__(256)
intensity_kernel(long npix, int batch, const float2* __restrict__ f, int sum_batch,
                 float* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  if (sum_batch) {
    float acc = 0.f;
    for (int b = 0; b < batch; ++b) {
      const float2 e = f[(long)b * npix + p];
      acc += e.x * e.x + e.y * e.y;
    }
    out[p] = acc;
  } else {
    for (int b = 0; b < batch; ++b) {
      const float2 e = f[(long)b * npix + p];
      out[(long)b * npix + p] = e.x * e.x + e.y * e.y;
    }
  }
}

hipError_t launch_intensity(long npix, int batch, const float2* f, int sum_batch, float* out,
                            hipStream_t s) {
  hipLaunchKernelGGL(intensity_kernel, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s,
                     npix, batch, f, sum_batch, out);
  return hipGetLastError();
}

// psf[p] (=|+=) sum_b planes[b][p], index order.  accumulate != 0 adds to psf.
__global__ void __launch_bounds__(256)
sum_planes_kernel(long npix, int nplanes, const float* __restrict__ planes, int accumulate,
                  float* __restrict__ psf) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  float acc = accumulate ? psf[p] : 0.f;
  for (int b = 0; b < nplanes; ++b) acc += planes[(long)b * npix + p];
  psf[p] = acc;
}

hipError_t launch_sum_planes(long npix, int nplanes, const float* planes, int accumulate,
                             float* psf, hipStream_t s) {
  hipLaunchKernelGGL(sum_planes_kernel, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s,
                     npix, nplanes, planes, accumulate, psf);
  return hipGetLastError();
}

}  // namespace dlx
