// Transfer-matrix generation and the lambda-independent / elementwise passes
// of the hot path (gfx950): pupil-field synthesis, basis evaluation, power and
// intensity reductions.  All are HBM-bound; they are written for coalesced
// 8/16-byte accesses with one pixel (or four) per lane.
#include "dlx_internal.h"
#include "dlx_phasor.h"
#include "dlx_split.h"

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
// The argument is accumulated in FP64 turns and reduced to [-1/2, 1/2] before
// the hardware sine / cosine (which take turns).
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
  const double path = pupil_path(tx, pupil_coord(x, dx, x0), ty * pupil_coord(y, dx, x0),
                                 opd ? (double)opd[p] : 0.0);
  const double turns0 = phase ? (double)phase[p] * 0.15915494309189535 : 0.0;  // 1 / (2 pi)
  float2 v = pupil_phasor(path, c.inv_wl[b], turns0, scale * (amp ? amp[p] : 1.0f));
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
// split4 producers: one lane per 8-element chunk, four 16-byte stores.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
dft_split4_kernel(DftChunk c, int n_in, int n_out, int inverse, uint8_t* __restrict__ out,
                  long batch_bytes) {
  // a quad of lanes per chunk: lane q computes elements 2q, 2q+1 (quad-cooperative store)
  const long kc = split4_chunks(n_in);
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (t >= (long)n_out * kc * 4) return;   // whole quads exit together
  const int q = (int)(t & 3);
  const long cidx = t >> 2;
  const int l = (int)(cidx / kc), ch = (int)(cidx - (long)l * kc);
  const double bb = (double)l * c.b_scale[b] + c.b_off[b];
  const float amp = (float)c.amp[b];
  float2 v[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int k = ch * 8 + 2 * q + e;
    if (k < n_in) {
      const double a = (double)k * c.a_scale[b] + c.a_off[b];
      double ph = a * bb;
      ph -= rint(ph);
      // hardware sin/cos take turns; max abs error 1.3e-7 on [-1/2, 1/2], far below
      // the 2^-17 relative resolution of the split-BF16 pair
      const float sn = __builtin_amdgcn_sinf((float)ph), co = __builtin_amdgcn_cosf((float)ph);
      v[e] = make_float2(amp * co, inverse ? amp * sn : -amp * sn);
    } else {
      v[e] = make_float2(0.f, 0.f);
    }
  }
  store_chunk_quad(out + (long)b * batch_bytes, l, split4_row_bytes(n_in), ch, q, v[0], v[1]);
}

hipError_t launch_dft_split4(const DftChunk& c, int nb, int n_in, int n_out, int inverse,
                             uint8_t* out, long batch_bytes, hipStream_t s) {
  const long work = (long)n_out * split4_chunks(n_in) * 4;
  hipLaunchKernelGGL(dft_split4_kernel, dim3((unsigned)((work + 255) / 256), nb), dim3(256), 0,
                     s, c, n_in, n_out, inverse, out, batch_bytes);
  return hipGetLastError();
}

// Pupil field in split4 (rows y, k = x).  A quad of lanes produces one chunk
// (8 pixels): lane q loads pixels 2q, 2q+1 of amplitude / OPD / phase once and
// emits them for up to 16 wavelengths with quad-cooperative 16-byte stores.
constexpr int kPupilWl = 16;

__global__ void __launch_bounds__(256)
pupil_split4_kernel(PupilChunk c, int nb, int n, double dx, double x0, double tx, double ty,
                    const float* __restrict__ amp, const float* __restrict__ opd,
                    const float* __restrict__ phase, float scale, uint8_t* __restrict__ out,
                    long batch_bytes) {
  const long kc = split4_chunks(n);
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  // the work size is a multiple of 4 and of the quad alignment: whole quads exit together
  if (t >= (long)n * kc * 4) return;
  const int q = (int)(t & 3);
  const long cidx = t >> 2;
  const int y = (int)(cidx / kc), ch = (int)(cidx - (long)y * kc);
  const double tyY = ty * pupil_coord(y, dx, x0);
  double path[2], turns0[2];
  float a[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int x = ch * 8 + 2 * q + e;
    const bool in = x < n;
    const long p = (long)y * n + x;
    path[e] = pupil_path(tx, pupil_coord(x, dx, x0), tyY, (in && opd) ? (double)opd[p] : 0.0);
    turns0[e] = (in && phase) ? (double)phase[p] * 0.15915494309189535 : 0.0;
    a[e] = in ? scale * (amp ? amp[p] : 1.0f) : 0.f;
  }
  const int b0 = blockIdx.y * kPupilWl;
  const int b1 = b0 + kPupilWl < nb ? b0 + kPupilWl : nb;
  const long rb = split4_row_bytes(n);
  for (int b = b0; b < b1; ++b) {
    float2 v[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) v[e] = pupil_phasor(path[e], c.inv_wl[b], turns0[e], a[e]);
    store_chunk_quad(out + (long)b * batch_bytes, y, rb, ch, q, v[0], v[1]);
  }
}

hipError_t launch_pupil_split4(const PupilChunk& c, int nb, int n, double dx, double x0,
                               double tx, double ty, const float* amp, const float* opd,
                               const float* phase, float scale, uint8_t* out, long batch_bytes,
                               hipStream_t s) {
  const long work = (long)n * split4_chunks(n) * 4;
  dim3 grid((unsigned)((work + 255) / 256), (nb + kPupilWl - 1) / kPupilWl);
  hipLaunchKernelGGL(pupil_split4_kernel, grid, dim3(256), 0, s, c, nb, n, dx, x0, tx, ty, amp,
                     opd, phase, scale, out, batch_bytes);
  return hipGetLastError();
}

// complex64 rows x k -> split4 (per batch instance).
__global__ void __launch_bounds__(256)
to_split4_kernel(const float2* __restrict__ in, long rows, int k, long in_batch,
                 uint8_t* __restrict__ out, long out_batch_bytes) {
  const long kc = split4_chunks(k);
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (t >= rows * kc) return;
  const long r = t / kc;
  const int ch = (int)(t - r * kc);
  const float2* src = in + (long)b * in_batch + r * k;
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int x = ch * 8 + e;
    v[e] = x < k ? src[x] : make_float2(0.f, 0.f);
  }
  store_chunk(out + (long)b * out_batch_bytes, r, split4_row_bytes(k), ch, v);
}

hipError_t launch_to_split4(const float2* in, int batch, long rows, int k, long in_batch,
                            uint8_t* out, long out_batch_bytes, hipStream_t s) {
  const long work = rows * split4_chunks(k);
  hipLaunchKernelGGL(to_split4_kernel, dim3((unsigned)((work + 255) / 256), batch), dim3(256),
                     0, s, in, rows, k, in_batch, out, out_batch_bytes);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Basis evaluation: out[s][p] = sum_j coeffs[s][j] * basis[j][p] (+ base[p]).
// One lane per 4 pixels (16-byte loads); the basis is streamed once for up to
// kSets coefficient sets (utils/math.py:177-196 tensordot).
// ---------------------------------------------------------------------------
constexpr int kSets = 8;

// NS coefficient sets per launch; the basis loop is unrolled so several
// 16-byte loads per lane are in flight (the kernel streams the basis once).
template <int NS>
__global__ void __launch_bounds__(256)
eval_basis_kernel(int nbasis, long npix, int nsets, const float* __restrict__ basis,
                  const float* __restrict__ coeffs, const float* __restrict__ base,
                  float* __restrict__ out, int vec_ok) {
  const long p4 = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (p4 >= npix) return;
  const bool full = vec_ok && (p4 + 4 <= npix);
  float acc[NS][4];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[s][q] = 0.f;
#pragma unroll 6
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
    for (int s = 0; s < NS; ++s) {
      if (NS == 1 || s < nsets) {
        const float cj = coeffs[(long)s * nbasis + j];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[s][q] = fmaf(cj, v[q], acc[s][q]);
      }
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    if (NS > 1 && s >= nsets) break;
    if (full) {
      float4 r = make_float4(acc[s][0], acc[s][1], acc[s][2], acc[s][3]);
      if (base) {
        const float4 bb = *(const float4*)(base + p4);
        r.x += bb.x; r.y += bb.y; r.z += bb.z; r.w += bb.w;
      }
      *(float4*)(out + (long)s * npix + p4) = r;
    } else {
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
}

hipError_t launch_eval_basis(int nbasis, long npix, int nsets, const float* basis,
                             const float* coeffs, const float* base, float* out,
                             hipStream_t s) {
  // 16-byte loads and stores need 16-byte aligned rows
  const int vec_ok = ((npix & 3) == 0) && ((reinterpret_cast<uintptr_t>(basis) & 15) == 0) &&
                     ((reinterpret_cast<uintptr_t>(out) & 15) == 0) &&
                     (!base || (reinterpret_cast<uintptr_t>(base) & 15) == 0);
  const long lanes = (npix + 3) / 4;
  const dim3 grid((unsigned)((lanes + 255) / 256));
  if (nsets == 1) {
    hipLaunchKernelGGL(eval_basis_kernel<1>, grid, dim3(256), 0, s, nbasis, npix, 1, basis,
                       coeffs, base, out, vec_ok);
    return hipGetLastError();
  }
  for (int s0 = 0; s0 < nsets; s0 += kSets) {
    const int ns = nsets - s0 < kSets ? nsets - s0 : kSets;
    hipLaunchKernelGGL(eval_basis_kernel<kSets>, grid, dim3(256), 0, s, nbasis, npix, ns, basis,
                       coeffs + (long)s0 * nbasis, base, out + (long)s0 * npix, vec_ok);
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

// psf[p] (=|+=) sum_b planes[b][p] in index order (deterministic); four
// pixels per lane and a plane loop unrolled 16 deep keep 256 bytes per lane in
// flight (a 256^2 PSF has only 16 Ki lanes).
__global__ void __launch_bounds__(256)
sum_planes_kernel(long npix, int nplanes, const float* __restrict__ planes, int accumulate,
                  float* __restrict__ psf, int vec_ok) {
  const long p4 = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (p4 >= npix) return;
  if (vec_ok) {
    float4 acc = accumulate ? *(const float4*)(psf + p4) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 16
    for (int b = 0; b < nplanes; ++b) {
      const float4 v = *(const float4*)(planes + (long)b * npix + p4);
      acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
    }
    *(float4*)(psf + p4) = acc;
  } else {
    for (long p = p4; p < p4 + 4 && p < npix; ++p) {
      float acc = accumulate ? psf[p] : 0.f;
      for (int b = 0; b < nplanes; ++b) acc += planes[(long)b * npix + p];
      psf[p] = acc;
    }
  }
}

hipError_t launch_sum_planes(long npix, int nplanes, const float* planes, int accumulate,
                             float* psf, hipStream_t s) {
  const long lanes = (npix + 3) / 4;
  const int vec_ok = ((npix & 3) == 0) && ((reinterpret_cast<uintptr_t>(planes) & 15) == 0) &&
                     ((reinterpret_cast<uintptr_t>(psf) & 15) == 0);
  hipLaunchKernelGGL(sum_planes_kernel, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, s,
                     npix, nplanes, planes, accumulate, psf, vec_ok);
  return hipGetLastError();
}

}  // namespace dlx
