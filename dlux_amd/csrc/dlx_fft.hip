// Centred 2-D FFT (dLux's FFT propagator) and the angular-spectrum propagator
// on gfx950.
//
// dLux: out = fftshift(fft2(ifftshift(pad(x)))) / M      (utils/propagation.py:52-64)
//       inverse: fftshift(ifft2(ifftshift(pad(x)))) * M  (= conj-sign DFT / M)
//
// A 2-D transform is two launches of one kernel.  Each workgroup loads R
// rows of length M (power of two) into LDS -- applying the zero pad and
// ifftshift as an index map on the fly -- runs a Stockham auto-sort FFT on
// every row (radix-8 stages, then one radix-4 or radix-2 stage; each stage
// reads its butterfly inputs into registers, synchronises, and writes the
// outputs in place), and writes the rows out TRANSPOSED.  The second launch therefore
// reads the first's columns as rows and its transposed write restores the
// orientation; that write also applies fftshift, the 1/M scale and, for the
// angular-spectrum method, the transfer function H(fx, fy).  HBM traffic is
// one read and one write of the M x M complex64 array per launch.
#include <cmath>

#include "dlx_internal.h"

namespace dlx {
namespace {

// At most kLdsElems complex64 of rows per workgroup plus an M-entry twiddle
// table, within the 160 KiB of LDS; 8 waves per workgroup.
constexpr int kThreads = 512;
constexpr int kLdsElems = 16384;
constexpr int kLdsBudget = 20480;   // complex64 values in 160 KiB
// LDS row stride beyond M (complex64).  0: padded rows would spread the
// transposed store's row reads over the banks, but push a 4-row 2048-point
// block past half of the CU's LDS and halve the workgroups per CU (measured
// slower).
constexpr int kRowPad = 0;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// (a) * (sgn i)
__device__ __forceinline__ float2 mul_si(float2 a, float sgn) {
  return make_float2(-sgn * a.y, sgn * a.x);
}

// In-register DFT of size RAD: y_q = sum_r x_r exp(sgn 2 pi i r q / RAD).
template <int RAD>
__device__ __forceinline__ void dft_reg(float2 (&x)[RAD], float sgn);

template <>
__device__ __forceinline__ void dft_reg<2>(float2 (&x)[2], float) {
  const float2 a = x[0], b = x[1];
  x[0] = make_float2(a.x + b.x, a.y + b.y);
  x[1] = make_float2(a.x - b.x, a.y - b.y);
}

template <>
__device__ __forceinline__ void dft_reg<4>(float2 (&x)[4], float sgn) {
  const float2 a0 = make_float2(x[0].x + x[2].x, x[0].y + x[2].y);
  const float2 a1 = make_float2(x[0].x - x[2].x, x[0].y - x[2].y);
  const float2 a2 = make_float2(x[1].x + x[3].x, x[1].y + x[3].y);
  const float2 a3 = mul_si(make_float2(x[1].x - x[3].x, x[1].y - x[3].y), sgn);
  x[0] = make_float2(a0.x + a2.x, a0.y + a2.y);
  x[2] = make_float2(a0.x - a2.x, a0.y - a2.y);
  x[1] = make_float2(a1.x + a3.x, a1.y + a3.y);
  x[3] = make_float2(a1.x - a3.x, a1.y - a3.y);
}

template <>
__device__ __forceinline__ void dft_reg<8>(float2 (&x)[8], float sgn) {
  float2 e[4] = {x[0], x[2], x[4], x[6]};
  float2 o[4] = {x[1], x[3], x[5], x[7]};
  dft_reg<4>(e, sgn);
  dft_reg<4>(o, sgn);
  const float h = 0.70710678118654752f;
  // o_q * exp(sgn 2 pi i q / 8)
  const float2 o1 = make_float2(h * (o[1].x - sgn * o[1].y), h * (o[1].y + sgn * o[1].x));
  const float2 o2 = mul_si(o[2], sgn);
  const float2 o3 = make_float2(h * (-o[3].x - sgn * o[3].y), h * (-o[3].y + sgn * o[3].x));
  const float2 ot[4] = {o[0], o1, o2, o3};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    x[q] = make_float2(e[q].x + ot[q].x, e[q].y + ot[q].y);
    x[q + 4] = make_float2(e[q].x - ot[q].x, e[q].y - ot[q].y);
  }
}

// One Stockham stage of radix RAD on `rows` rows of length M held in LDS:
// item j (of M/RAD per row) reads x_r = row[j + r M/RAD], twiddles them by
// exp(sgn 2 pi i r k / (p RAD)) with k = j mod p, transforms, and writes
// y_q to row[(j - k) RAD + k + q p].  p is the product of earlier radices.
template <int RAD, int LOG2RAD>
__device__ __forceinline__ void stockham_stage(float2* buf, const float2* tw, int total,
                                               int log2M, int rs, int log2p, float sgn) {
  constexpr int kItems = kLdsElems / RAD / kThreads;
  const int log2_per_row = log2M - LOG2RAD;
  const int per_row = 1 << log2_per_row;
  const int p = 1 << log2p;
  float2 v[kItems][RAD];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int t = threadIdx.x + it * kThreads;
    if (t < total) {
      const int rr = t >> log2_per_row, j = t & (per_row - 1);
      const float2* row = buf + rr * rs;
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[it][r] = row[j + r * per_row];
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int t = threadIdx.x + it * kThreads;
    if (t < total) {
      const int rr = t >> log2_per_row, j = t & (per_row - 1);
      const int k = j & (p - 1);
      if (p > 1) {
        const int sh = log2M - log2p - LOG2RAD;   // twiddle index r k M / (p RAD)
#pragma unroll
        for (int r = 1; r < RAD; ++r) v[it][r] = cmul(v[it][r], tw[(r * k) << sh]);
      }
      dft_reg<RAD>(v[it], sgn);
      float2* row = buf + rr * rs;
      const int i0 = ((j - k) << LOG2RAD) + k;
#pragma unroll
      for (int q = 0; q < RAD; ++q) row[i0 + (q << log2p)] = v[it][q];
    }
  }
  __syncthreads();
}

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
  const int RS = M + kRowPad;      // LDS row stride
  float2* tw = lds + (long)R * RS;  // M twiddles exp(sign 2 pi i q / M)
  // consecutive row groups on one XCD (workgroups are dealt to the 8 XCDs
  // round-robin), so the short runs of the transposed stores that neighbouring
  // groups write into the same cache lines meet in that XCD's L2
  const int groups = gridDim.x, total_wg = gridDim.x * gridDim.y;
  int id = blockIdx.y * groups + blockIdx.x;
  if (total_wg % 8 == 0) id = (id % 8) * (total_wg / 8) + id / 8;
  const int b = id / groups;
  const int r0 = (id - b * groups) * R;
  const float sgn = P.inverse ? 1.f : -1.f;
  const float2* src = in + (long)b * P.in_bstride;

  for (int q = threadIdx.x; q < M; q += kThreads) {
    float s, c;
    sincospif(2.0f * (float)q / (float)M, &s, &c);
    tw[q] = make_float2(c, sgn * s);
  }
  // load in natural order (M and R are powers of two: shifts and masks), 16
  // values per thread in flight before any is written to LDS
  const int total = R * M;
  const int log2M = P.log2M;
  constexpr int kLoadBatch = 16;
  for (int t0 = 0; t0 < total; t0 += kLoadBatch * kThreads) {
    float2 v[kLoadBatch];
#pragma unroll
    for (int i = 0; i < kLoadBatch; ++i) {
      const int t = t0 + threadIdx.x + i * kThreads;
      const int rr = t >> log2M, nc = t & (M - 1);
      const int row = r0 + rr;
      v[i] = make_float2(0.f, 0.f);
      if (t < total) {
        if (P.in_mode == 0) {
          const int sr = ((row + half_m) & (M - 1)) - P.p;
          const int sc = ((nc + half_m) & (M - 1)) - P.p;
          if (sr >= 0 && sr < P.N && sc >= 0 && sc < P.N) v[i] = src[(long)sr * P.N + sc];
        } else {
          v[i] = src[(long)row * M + nc];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kLoadBatch; ++i) {
      const int t = t0 + threadIdx.x + i * kThreads;
      if (t < total) lds[(t >> log2M) * RS + (t & (M - 1))] = v[i];
    }
  }
  __syncthreads();

  // Stockham stages: radix 8 while it divides, then one radix-4 or radix-2 stage
  int log2p = 0;
  while (log2M - log2p >= 3) {
    stockham_stage<8, 3>(lds, tw, total >> 3, log2M, RS, log2p, sgn);
    log2p += 3;
  }
  if (log2M - log2p == 2) stockham_stage<4, 2>(lds, tw, total >> 2, log2M, RS, log2p, sgn);
  else if (log2M - log2p == 1) stockham_stage<2, 1>(lds, tw, total >> 1, log2M, RS, log2p, sgn);

  // transposed store: out[k][row] (consecutive lanes -> consecutive rows)
  float2* dst = out + (long)b * P.out_bstride;
  for (int t = threadIdx.x; t < total; t += kThreads) {
    const int k = t >> P.log2R, rr = t & (R - 1);
    const int row = r0 + rr;
    float2 v = lds[rr * RS + k];
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
          // the phase z sqrt(1/lambda^2 - f^2) in FP64 turns (tens of thousands of
          // turns at millimetre distances), reduced, then the hardware sin / cos
          // (turns; max abs error 1.3e-7 on [-1/2, 1/2])
          double t = P.asm_z * sqrt(arg);
          t -= rint(t);
          const float s = __builtin_amdgcn_sinf((float)t), c = __builtin_amdgcn_cosf((float)t);
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
  // rows per workgroup: about two workgroups per CU (512) when the rows fit,
  // within kLdsElems and the LDS left beside the M-entry twiddle table
  int cap = (kLdsBudget - P.M) / (P.M + kRowPad);
  if (cap > kLdsElems / P.M) cap = kLdsElems / P.M;
  if (cap < 1) cap = 1;
  // (at least 4 rows -- 32-byte runs in the transposed store -- while that
  // still leaves one workgroup per CU)
  int r = P.M / 512;
  const int r_min = P.M / 256 < 4 ? P.M / 256 : 4;
  if (r < r_min) r = r_min;
  if (r < 1) r = 1;
  if (r > cap) r = cap;
  P.R = 1 << (ilog2(r + 1) - 1);   // largest power of two <= r
  if (P.R > P.M) P.R = P.M;
  P.log2R = ilog2(P.R);
  const size_t lds = ((size_t)P.R * (P.M + kRowPad) + P.M) * sizeof(float2);
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
  DLX_CHECK_ARG(batch <= 65535, "batch must be <= 65535, got %d", batch);
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
  DLX_CHECK_ARG(batch <= 65535, "batch must be <= 65535, got %d", batch);
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
