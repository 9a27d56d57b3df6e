// Pupil phasor shared by every pupil producer: the complex64 and split4 pupil
// kernels and the contraction whose loader generates the pupil on chip.  One
// definition, with the fused multiply-adds written out, so all of them give
// the same bits for the same pixel and wavelength.
//
//   P(y, x) = a(y, x) exp(2 pi i t),
//   t = (opd(y, x) + tx X + ty Y) / lambda + phase(y, x) / (2 pi)   [turns]
//
// (wavefronts.py:332-349 add_phase, :351-368 add_opd, :370-393 tilt; the FP64 turn is
// reduced to [-1/2, 1/2] before the hardware sine / cosine, which take turns.)
#pragma once

#include <hip/hip_runtime.h>

namespace dlx {

// Pixel-centre coordinate x0 + i dx.
__device__ __forceinline__ double pupil_coord(int i, double dx, double x0) {
  return fma((double)i, dx, x0);
}

// Optical path of a pixel: tx X + (ty Y) + opd, with ty Y formed by the caller
// (once per row).
__device__ __forceinline__ double pupil_path(double tx, double X, double tyY, double opd) {
  return fma(tx, X, tyY) + opd;
}

// a exp(2 pi i (path / lambda + turns0)); max abs error of the hardware
// sin / cos 1.3e-7 on [-1/2, 1/2].
__device__ __forceinline__ float2 pupil_phasor(double path, double inv_wl, double turns0,
                                               float a) {
  double t = fma(path, inv_wl, turns0);
  t -= rint(t);
  const float s = __builtin_amdgcn_sinf((float)t), c = __builtin_amdgcn_cosf((float)t);
  return make_float2(a * c, a * s);
}

}  // namespace dlx
