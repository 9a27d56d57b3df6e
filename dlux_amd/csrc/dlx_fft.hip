// Centred 2-D FFT (dLux's FFT propagator) and the angular-spectrum propagator
// on gfx950.
//
// dLux: out = fftshift(fft2(ifftshift(pad(x)))) / M      (utils/propagation.py:52-64)
//       inverse: fftshift(ifft2(ifftshift(pad(x)))) * M  (= conj-sign DFT / M)
//
// A 2-D transform is two launches of one kernel.  Each workgroup loads R
// rows of length M (power of two) into LDS -- applying the zero pad and
// ifftshift as an index map on the fly -- runs an in-place radix-2 DIT FFT on
// every row, and writes the rows out TRANSPOSED.  The second launch therefore
// reads the first's columns as rows and its transposed write restores the
// orientation; that write also applies fftshift, the 1/M scale and, for the
// angular-spectrum method, the transfer function H(fx, fy).  HBM traffic is
// one read and one write of the M x M complex64 array per launch.
#include <cmath>

#include "dlx_internal.h"

namespace dlx {
namespace {

// R * M complex64 per workgroup (128 KiB + twiddles): R = 8 rows of 2048, so the
// transposed store writes 64-byte runs; 8 waves per workgroup.
constexpr int kThreads = 512;
constexpr int kLdsElems = 16384;

struct FftParams {
  int M, log2M, R, log2R, inverse;
  int in_mode;      // 0: centred + zero-padded from an N x N source; 1: plain M x M rows
  int N, p;         // source side and pad offset (in_mode 0)
  long in_bstride, out_bstride;
  int out_mode;     // 0: transposed; 1: transposed + fftshift + scale (+ H)
  float scale;
  int asm_on;
  double asm_z, asm_inv_wl2, asm_df;
};

__global__ void __launch_bounds__(kThreads)
fft_rows_kernel(const float2* __restrict__ in, float2* __restrict__ out, FftParams P) {
  extern __shared__ float2 lds[];
  const int M = P.M, R = P.R, half_m = M >> 1;
  float2* tw = lds + (long)R * M;  // M/2 twiddles exp(sign 2 pi i q / M)
  const int b = blockIdx.y;
  const int r0 = blockIdx.x * R;
  const float sgn = P.inverse ? 1.f : -1.f;
  const float2* src = in + (long)b * P.in_bstride;

  for (int q = threadIdx.x; q < half_m; q += kThreads) {
    float s, c;
    sincospif(2.0f * (float)q / (float)M, &s, &c);
    tw[q] = make_float2(c, sgn * s);
  }
  // load with bit-reversed placement (M and R are powers of two: shifts and masks)
  const int total = R * M;
  const int log2M = P.log2M, log2H = P.log2M - 1;
  for (int t = threadIdx.x; t < total; t += kThreads) {
    const int rr = t >> log2M, nc = t & (M - 1);
    const int row = r0 + rr;
    float2 v = make_float2(0.f, 0.f);
    if (P.in_mode == 0) {
      const int sr = ((row + half_m) & (M - 1)) - P.p;
      const int sc = ((nc + half_m) & (M - 1)) - P.p;
      if (sr >= 0 && sr < P.N && sc >= 0 && sc < P.N) v = src[(long)sr * P.N + sc];
    } else {
      v = src[(long)row * M + nc];
    }
    const int rev = (int)(__brev((unsigned)nc) >> (32 - P.log2M));
    lds[rr * M + rev] = v;
  }
  __syncthreads();

  // in-place radix-2 decimation-in-time stages
  const int nbfly = R * half_m;
  for (int lh = 0, ltw = log2H; lh < log2M; ++lh, --ltw) {
    const int h = 1 << lh;   // half the butterfly span at this stage
    for (int t = threadIdx.x; t < nbfly; t += kThreads) {
      const int rr = t >> log2H, j = t & (half_m - 1);
      const int grp = j >> lh, pos = j & (h - 1);
      const int i0 = (rr << log2M) + (grp << (lh + 1)) + pos, i1 = i0 + h;
      const float2 w = tw[pos << ltw];
      const float2 a = lds[i0], x = lds[i1];
      const float2 bw = make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
      lds[i0] = make_float2(a.x + bw.x, a.y + bw.y);
      lds[i1] = make_float2(a.x - bw.x, a.y - bw.y);
    }
    __syncthreads();
  }

  // transposed store: out[k][row] (consecutive lanes -> consecutive rows)
  float2* dst = out + (long)b * P.out_bstride;
  for (int t = threadIdx.x; t < total; t += kThreads) {
    const int k = t >> P.log2R, rr = t & (R - 1);
    const int row = r0 + rr;
    float2 v = lds[rr * M + k];
    if (P.out_mode == 0) {
      dst[(long)k * M + row] = v;
    } else {
      const int oy = (k + half_m) & (M - 1), ox = (row + half_m) & (M - 1);
      v.x *= P.scale;
      v.y *= P.scale;
      if (P.asm_on) {
        const double fy = (double)(oy - half_m) * P.asm_df, fx = (double)(ox - half_m) * P.asm_df;
        const double arg = P.asm_inv_wl2 - fx * fx - fy * fy;
        if (arg > 0.0) {
          double t = P.asm_z * sqrt(arg);
          t -= rint(t);
          float s, c;
          sincospif(2.0f * (float)t, &s, &c);
          v = make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
        } else {
          v = make_float2(0.f, 0.f);
        }
      }
      dst[(long)oy * M + ox] = v;
    }
  }
}

int ilog2(int m) {
  int l = 0;
  while ((1 << l) < m) ++l;
  return l;
}

hipError_t launch_rows(const float2* in, float2* out, int batch, FftParams P, hipStream_t s) {
  P.log2M = ilog2(P.M);
  P.R = kLdsElems / P.M;
  if (P.R < 1) P.R = 1;
  if (P.R > P.M) P.R = P.M;
  P.log2R = ilog2(P.R);
  const size_t lds = ((size_t)P.R * P.M + P.M / 2) * sizeof(float2);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  if (lds > 64 * 1024) {
    // Opt in to more than 64 KiB of dynamic LDS (gfx950 has 160 KiB per CU).
    (void)hipFuncSetAttribute((const void*)fft_rows_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();
  }
  hipLaunchKernelGGL(fft_rows_kernel, dim3(P.M / P.R, batch), dim3(kThreads), lds, s, in, out,
                     P);
  return hipGetLastError();
}

}  // namespace

// Maximum transform side handled by the LDS kernel (one row + twiddles fit).
constexpr int kMaxFftSide = 8192;

bool fft_supported(int m) { return m >= 2 && m <= kMaxFftSide && (m & (m - 1)) == 0; }

// Full centred 2-D transform of `batch` N x N fields, padded to M x M.
// tmp: batch * M * M complex64 scratch.
hipError_t launch_fft2_centred(const float2* in, int N, int M, int batch, int inverse,
                               float scale, float2* tmp, float2* out, int asm_on, double z,
                               double wavelength, double dx, hipStream_t s) {
  FftParams A{};
  A.M = M;
  A.inverse = inverse;
  A.in_mode = 0;
  A.N = N;
  A.p = (M - N) / 2;
  A.in_bstride = (long)N * N;
  A.out_bstride = (long)M * M;
  A.out_mode = 0;
  hipError_t e = launch_rows(in, tmp, batch, A, s);
  if (e != hipSuccess) return e;
  FftParams B = A;
  B.in_mode = 1;
  B.in_bstride = (long)M * M;
  B.out_mode = 1;
  B.scale = scale;
  B.asm_on = asm_on;
  if (asm_on) {
    B.asm_z = z;
    B.asm_inv_wl2 = 1.0 / (wavelength * wavelength);
    B.asm_df = 1.0 / ((double)M * dx);
  }
  return launch_rows(tmp, out, batch, B, s);
}

}  // namespace dlx

using namespace dlx;

extern "C" size_t dlx_fft2_workspace_size(int32_t n, int32_t m, int32_t batch) {
  (void)n;
  if (m <= 0 || batch <= 0) return 0;
  return align_up((size_t)batch * m * m * sizeof(float2));
}

extern "C" int dlx_fft2(int32_t n, int32_t m, int32_t batch, int32_t inverse, const void* in,
                        void* out, void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  DLX_CHECK_ARG(n > 0 && m >= n && batch >= 0, "need 0 < n <= m and batch >= 0");
  DLX_CHECK_ARG(((m - n) % 2) == 0, "m - n must be even (centre-preserving pad)");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(in && out, "null argument to dlx_fft2");
  if (!fft_supported(m)) {
    set_error("dlx_fft2 supports power-of-two sides up to %d, got %d", kMaxFftSide, m);
    return DLX_ERR_UNSUPPORTED;
  }
  const size_t need = dlx_fft2_workspace_size(n, m, batch);
  if (!workspace || workspace_bytes < need) {
    set_error("dlx_fft2 workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  DLX_HIP(launch_fft2_centred((const float2*)in, n, m, batch, inverse, 1.0f / (float)m,
                              (float2*)workspace, (float2*)out, 0, 0, 0, 0,
                              (hipStream_t)stream));
  return DLX_OK;
}

extern "C" size_t dlx_asm_workspace_size(int32_t n, int32_t m, int32_t batch) {
  if (m <= 0 || batch <= 0) return 0;
  return 2 * dlx_fft2_workspace_size(n, m, batch);
}

extern "C" int dlx_asm_prop(int32_t n, int32_t m, int32_t batch, const double* wavelength,
                            double pixel_scale, double distance, const void* in, void* out,
                            void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  DLX_CHECK_ARG(n > 0 && m >= n && batch >= 0, "need 0 < n <= m and batch >= 0");
  DLX_CHECK_ARG(((m - n) % 2) == 0, "m - n must be even (centre-preserving pad)");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(in && out && wavelength, "null argument to dlx_asm_prop");
  DLX_CHECK_ARG(std::isfinite(pixel_scale) && pixel_scale > 0, "pixel_scale must be > 0");
  DLX_CHECK_ARG(std::isfinite(distance), "distance must be finite");
  if (!fft_supported(m)) {
    set_error("dlx_asm_prop supports power-of-two padded sides up to %d, got %d", kMaxFftSide,
              m);
    return DLX_ERR_UNSUPPORTED;
  }
  const size_t need = dlx_asm_workspace_size(n, m, batch);
  if (!workspace || workspace_bytes < need) {
    set_error("dlx_asm_prop workspace too small: need %zu bytes, got %zu", need,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  const size_t slab = dlx_fft2_workspace_size(n, m, batch);
  float2* tmp = (float2*)workspace;
  float2* spec = (float2*)((char*)workspace + slab);
  const float inv_m = 1.0f / (float)m;
  for (int b = 0; b < batch; ++b) {
    DLX_CHECK_ARG(std::isfinite(wavelength[b]) && wavelength[b] > 0,
                  "wavelength[%d] must be > 0", b);
    const float2* src = (const float2*)in + (long)b * n * n;
    float2* sp = spec + (long)b * m * m;
    // forward centred transform with H fused into its final store
    DLX_HIP(launch_fft2_centred(src, n, m, 1, 0, inv_m, tmp, sp, 1, distance, wavelength[b],
                                pixel_scale, s));
    // inverse centred transform back to the (padded) spatial grid: total 1/M^2
    DLX_HIP(launch_fft2_centred(sp, m, m, 1, 1, inv_m, tmp, (float2*)out + (long)b * m * m, 0,
                                0, 0, 0, s));
  }
  return DLX_OK;
}
