// Matrix-core rate probe: the bf16 MFMA throughput this GPU sustains, with
// operands of low entropy (the nominal rate) or with full-entropy operands
// like the split-BF16x3 contractions' (each of the 16 significant bits of the
// hi and lo planes random).  The contractions' utilisation is read against
// both.  Each wave issues v_mfma_f32_32x32x16_bf16 on 8 independent
// accumulators from registers only (no memory traffic), 8 waves per CU.
#include "dlx_internal.h"

namespace dlx {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

constexpr int kChains = 8;
constexpr int kFrags = 8;

template <bool RANDOM>
__global__ void __launch_bounds__(256, 2)
mfma_rate_kernel(float* __restrict__ out, int iters) {
  bf16x8_t a[kFrags], b[kFrags];
  uint32_t st = 2654435761u * (blockIdx.x * 256u + threadIdx.x + 1u);
#pragma unroll
  for (int q = 0; q < kFrags; ++q)
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      float x = 0.25f, y = 0.5f;
      if (RANDOM) {
        st = st * 1664525u + 1013904223u;
        x = (float)(st >> 8) * (1.0f / 16777216.0f) - 0.5f;
        st = st * 1664525u + 1013904223u;
        y = (float)(st >> 8) * (1.0f / 16777216.0f) - 0.5f;
      }
      a[q][i] = (__bf16)x;
      b[q][i] = (__bf16)y;
    }
  f32x16_t acc[kChains];
  for (int c = 0; c < kChains; ++c)
    for (int r = 0; r < 16; ++r) acc[c][r] = 0.f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < kFrags; ++u)
#pragma unroll
      for (int c = 0; c < kChains; ++c)
        acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[(u + c) % kFrags],
                                                         b[(3 * u + c) % kFrags], acc[c], 0, 0,
                                                         0);
  }
  float s = 0.f;
  for (int c = 0; c < kChains; ++c)
    for (int r = 0; r < 16; ++r) s += acc[c][r];
  out[blockIdx.x * 256 + threadIdx.x] = s;   // keeps the work observable
}

}  // namespace
}  // namespace dlx

using namespace dlx;

extern "C" int dlx_mfma_rate(int32_t random_operands, double* tflops, void* workspace,
                             size_t workspace_bytes, void* stream) {
  clear_error();
  DLX_CHECK_ARG(tflops, "null argument to dlx_mfma_rate");
  int dev = 0, cus = 0;
  DLX_HIP(hipGetDevice(&dev));
  DLX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = 2 * cus;   // 8 waves per CU
  const size_t need = (size_t)blocks * 256 * sizeof(float);
  if (!workspace || workspace_bytes < need) {
    set_error("dlx_mfma_rate workspace too small: need %zu bytes, got %zu", need,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  // ~1-2 ms per launch; 30 warm-up launches let the clocks settle from idle
  const int iters = 400, reps = 10, warm = 30;
  hipEvent_t beg, end;
  DLX_HIP(hipEventCreate(&beg));
  DLX_HIP(hipEventCreate(&end));
  auto launch = [&]() {
    if (random_operands)
      hipLaunchKernelGGL(mfma_rate_kernel<true>, dim3(blocks), dim3(256), 0, s,
                         (float*)workspace, iters);
    else
      hipLaunchKernelGGL(mfma_rate_kernel<false>, dim3(blocks), dim3(256), 0, s,
                         (float*)workspace, iters);
  };
  for (int r = 0; r < warm; ++r) launch();
  DLX_HIP(hipEventRecord(beg, s));
  for (int r = 0; r < reps; ++r) launch();
  DLX_HIP(hipEventRecord(end, s));
  DLX_HIP(hipEventSynchronize(end));
  float ms = 0.f;
  DLX_HIP(hipEventElapsedTime(&ms, beg, end));
  (void)hipEventDestroy(beg);
  (void)hipEventDestroy(end);
  const double flops = (double)reps * blocks * 4 * iters * kFrags * kChains * 32768.0;
  *tflops = flops / (ms * 1e-3) / 1e12;
  return DLX_OK;
}
