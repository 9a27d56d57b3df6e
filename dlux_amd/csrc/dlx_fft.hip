// Centred 2-D FFT (dLux's FFT propagator) and the angular-spectrum propagator
// on gfx950.
//
// dLux: out = fftshift(fft2(ifftshift(pad(x)))) / M      (utils/propagation.py:52-64)
//       inverse: fftshift(ifft2(ifftshift(pad(x)))) * M  (= conj-sign DFT / M)
//
// A 2-D transform is two launches of one kernel, templated on log2 M so the
// index arithmetic folds.  Each workgroup loads R rows of length M (power of
// two) into LDS -- applying the zero pad and ifftshift as an index map on the
// fly -- runs a Stockham auto-sort FFT on every row (radix-8 stages, then one
// radix-4 or radix-2 stage; each stage reads its butterfly inputs into
// registers, synchronises, and writes the outputs in place), and writes the
// rows out TRANSPOSED.  The second launch therefore reads the first's columns
// as rows and its transposed write restores the orientation; that write also
// applies fftshift and the 1/M scale.  HBM traffic is one read and one write of
// the M x M complex64 array per launch.  The angular-spectrum method takes
// three launches instead of four (launch_asm): the y transforms of its forward
// and inverse 2-D transforms, with the transfer function between them, run
// back to back in LDS.
#include <cmath>

#include "dlx_internal.h"

namespace dlx {
namespace {

// At most kLdsElems complex64 of rows per workgroup plus an M-entry twiddle
// table, within the 160 KiB of LDS; 8 waves per workgroup.
constexpr int kThreads = 512;
constexpr int kLdsElems = 16384;
constexpr int kLdsBudget = 20480;   // complex64 values in 160 KiB

// One 8-byte store of a complex64 (the compiler may otherwise split a float2
// store into two 4-byte stores, doubling the write requests of the transposed
// store's short runs).
typedef float cplx_v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_c64(float2* p, float2 v) {
  *(cplx_v*)p = cplx_v{v.x, v.y};
}

// LDS position of element x of row rr (row length 2^LOG2M): x with its bits
// 3..8 XORed into bits 0..5 and a row term XORed in.  A bijection on
// [0, M) for every row, so each row keeps its own M slots; it spreads the
// Stockham stages' strided writes (stride RAD p) and the transposed
// accesses (one column of R rows, M apart) over the LDS banks.
template <int LOG2M>
__device__ __forceinline__ int lds_pos(int x, int rr) {
  return x ^ ((x >> 3) & 63) ^ ((rr * 20) & ((1 << LOG2M) - 1));
}

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

// One Stockham stage of radix RAD on rows of length M = 2^LOG2M held in LDS
// (row stride M): item j (of M/RAD per row) reads x_r = row[j + r M/RAD],
// twiddles them by exp(sgn 2 pi i r k / (p RAD)) with k = j mod p, transforms,
// and writes y_q to row[(j - k) RAD + k + q p].  p = 2^LOG2P is the product of
// earlier radices.  The sizes are template arguments, so the index arithmetic
// folds into shifts, masks and the LDS instructions' immediate offsets.
template <int RAD, int LOG2RAD, int LOG2M, int LOG2P, bool INV>
__device__ __forceinline__ void stockham_stage(float2* buf, const float2* tw, int total) {
  constexpr float sgn = INV ? 1.f : -1.f;   // the transform's sign, folded into the code
  constexpr int kItems = kLdsElems / RAD / kThreads;
  constexpr int log2_per_row = LOG2M - LOG2RAD;
  constexpr int per_row = 1 << log2_per_row;
  constexpr int p = 1 << LOG2P;
  float2 v[kItems][RAD];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int t = threadIdx.x + it * kThreads;
    if (t < total) {
      const int rr = t >> log2_per_row, j = t & (per_row - 1);
      const float2* row = buf + (rr << LOG2M);
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[it][r] = row[lds_pos<LOG2M>(j + r * per_row, rr)];
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
        constexpr int sh = LOG2M - LOG2P - LOG2RAD;   // twiddle index r k M / (p RAD)
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
          const float2 w = tw[(r * k) << sh];
          v[it][r] = cmul(v[it][r], make_float2(w.x, sgn * w.y));
        }
      }
      dft_reg<RAD>(v[it], sgn);
      float2* row = buf + (rr << LOG2M);
      const int i0 = ((j - k) << LOG2RAD) + k;
#pragma unroll
      for (int q = 0; q < RAD; ++q) row[lds_pos<LOG2M>(i0 + (q << LOG2P), rr)] = v[it][q];
    }
  }
  __syncthreads();
}

// All stages from p = 2^LOG2P: radix 8 while it divides, then one radix-4 or
// radix-2 stage.
template <int LOG2M, int LOG2P, bool INV>
__device__ __forceinline__ void stockham_from(float2* buf, const float2* tw, int total) {
  if constexpr (LOG2M - LOG2P >= 3) {
    stockham_stage<8, 3, LOG2M, LOG2P, INV>(buf, tw, total >> 3);
    stockham_from<LOG2M, LOG2P + 3, INV>(buf, tw, total);
  } else if constexpr (LOG2M - LOG2P == 2) {
    stockham_stage<4, 2, LOG2M, LOG2P, INV>(buf, tw, total >> 2);
  } else if constexpr (LOG2M - LOG2P == 1) {
    stockham_stage<2, 1, LOG2M, LOG2P, INV>(buf, tw, total >> 1);
  }
}

struct FftParams {
  int M, log2M, R, log2R, inverse;
  // 0: centred + zero-padded from an N x N source (ifftshift as an index map);
  // 1: plain M x M rows; 2: rows gathered from the columns of an M x M array
  int in_mode;
  int N, p;         // source side and pad offset (in_mode 0)
  long in_bstride, out_bstride;
  // 0: transposed; 1: transposed + fftshift + scale (+ H); 2: plain rows;
  // 3: plain rows + fftshift + scale
  int out_mode;
  float scale;
  // asm_on with out_mode 1: H(fx, fy) of the centred frequencies, fused into
  // the store.  asm_mid: the rows are y-lines of a spectrum transformed along
  // x (row index kx); each is transformed along y, multiplied by H at the
  // natural-order frequencies (kx, ky), and transformed back along y.
  int asm_on, asm_mid;
  double asm_z, asm_inv_wl2, asm_df;
  // rows_lo > 0 (in_mode 0): only the rows that hold the padded field, [0,
  // rows_lo) and [M - rows_lo, M), are transformed (the others are zero rows
  // whose transforms are zero); the launch has 2 rows_lo / R groups.
  // zlo < zhi (in_mode 1): positions [zlo, zhi) of every row load as zeros
  // (the columns the previous launch left unwritten).
  int rows_lo, zlo, zhi;
};

// The angular-spectrum transfer function at frequency indices (iy, ix) of the
// centred spectrum (-M/2 .. M/2 - 1 times df), applied to v:
// v exp(2 pi i z sqrt(1/lambda^2 - fx^2 - fy^2)) where the root is real, 0
// elsewhere.  The phase is formed in FP64 turns (tens of thousands of turns at
// millimetre distances), reduced, then taken by the hardware sin / cos
// (turns; max abs error 1.3e-7 on [-1/2, 1/2]).
__device__ __forceinline__ float2 apply_h(float2 v, int iy, int ix, const FftParams& P) {
  const double fy = (double)iy * P.asm_df, fx = (double)ix * P.asm_df;
  const double arg = P.asm_inv_wl2 - fx * fx - fy * fy;
  if (!(arg > 0.0)) return make_float2(0.f, 0.f);
  double t = P.asm_z * sqrt(arg);
  t -= rint(t);
  const float s = __builtin_amdgcn_sinf((float)t), c = __builtin_amdgcn_cosf((float)t);
  return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
}

template <int LOG2M>
__global__ void __launch_bounds__(kThreads)
fft_rows_kernel(const float2* __restrict__ in, float2* __restrict__ out, FftParams P) {
  extern __shared__ float2 lds[];
  constexpr int M = 1 << LOG2M, half_m = M >> 1;
  constexpr int RS = M;            // LDS row stride
  const int R = P.R;
  float2* tw = lds + (long)R * RS;  // M twiddles (cos, sin)(2 pi q / M)
  // consecutive row groups on one XCD (workgroups are dealt to the 8 XCDs
  // round-robin), so the short runs of the transposed stores (and loads) that
  // neighbouring groups access in the same cache lines meet in that XCD's L2
  const int groups = gridDim.x, total_wg = gridDim.x * gridDim.y;
  int id = blockIdx.y * groups + blockIdx.x;
  if (total_wg % 8 == 0) id = (id % 8) * (total_wg / 8) + id / 8;
  const int b = id / groups;
  int r0 = (id - b * groups) * R;
  if (P.rows_lo > 0 && r0 >= P.rows_lo) r0 += M - 2 * P.rows_lo;
  const float2* src = in + (long)b * P.in_bstride;

  if constexpr (M >= 4) {
    // one quadrant by sincospif, the other three by its exact symmetries
    for (int q = threadIdx.x; q < M / 4; q += kThreads) {
      float s, c;
      sincospif(2.0f * (float)q / (float)M, &s, &c);
      tw[q] = make_float2(c, s);
      tw[q + M / 4] = make_float2(-s, c);
      tw[q + M / 2] = make_float2(-c, -s);
      tw[q + 3 * M / 4] = make_float2(s, -c);
    }
  } else {
    for (int q = threadIdx.x; q < M; q += kThreads) {
      float s, c;
      sincospif(2.0f * (float)q / (float)M, &s, &c);
      tw[q] = make_float2(c, s);
    }
  }
  // load (M and R are powers of two: shifts and masks), 16 values per thread
  // in flight before any is written to LDS
  const int total = R * M;
  constexpr int kLoadBatch = 16;
  for (int t0 = 0; t0 < total; t0 += kLoadBatch * kThreads) {
    float2 v[kLoadBatch];
#pragma unroll
    for (int i = 0; i < kLoadBatch; ++i) {
      const int t = t0 + threadIdx.x + i * kThreads;
      v[i] = make_float2(0.f, 0.f);
      if (t < total) {
        if (P.in_mode == 0) {
          const int rr = t >> LOG2M, nc = t & (M - 1);
          const int sr = ((r0 + rr + half_m) & (M - 1)) - P.p;
          const int sc = ((nc + half_m) & (M - 1)) - P.p;
          if (sr >= 0 && sr < P.N && sc >= 0 && sc < P.N) v[i] = src[(long)sr * P.N + sc];
        } else if (P.in_mode == 1) {
          const int y = t & (M - 1);
          if (y < P.zlo || y >= P.zhi) v[i] = src[(long)r0 * M + t];
        } else {
          // element (rr, x) of the group is column r0 + rr of row x
          const int x = t >> P.log2R, rr = t & (R - 1);
          v[i] = src[(long)x * M + r0 + rr];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kLoadBatch; ++i) {
      const int t = t0 + threadIdx.x + i * kThreads;
      if (t < total) {
        const int rr = P.in_mode == 2 ? t & (R - 1) : t >> LOG2M;
        const int x = P.in_mode == 2 ? t >> P.log2R : t & (M - 1);
        lds[rr * RS + lds_pos<LOG2M>(x, rr)] = v[i];
      }
    }
  }
  __syncthreads();

  if (P.inverse) stockham_from<LOG2M, 0, true>(lds, tw, total);
  else stockham_from<LOG2M, 0, false>(lds, tw, total);

  if (P.asm_mid) {
    // y-line kx = r0 + rr at natural-order frequency ky: multiply by H, then
    // transform back along y
    for (int t = threadIdx.x; t < total; t += kThreads) {
      const int rr = t >> LOG2M, ky = t & (M - 1);
      const int kx = r0 + rr;
      float2* e = lds + rr * RS + lds_pos<LOG2M>(ky, rr);
      *e = apply_h(*e, ((ky + half_m) & (M - 1)) - half_m, ((kx + half_m) & (M - 1)) - half_m, P);
    }
    __syncthreads();
    if (P.inverse) stockham_from<LOG2M, 0, false>(lds, tw, total);
    else stockham_from<LOG2M, 0, true>(lds, tw, total);
  }

  float2* dst = out + (long)b * P.out_bstride;
  if (P.out_mode >= 2) {
    // plain rows (out_mode 3: fftshift on both axes and the scale)
    for (int t = threadIdx.x; t < total; t += kThreads) {
      const int rr = t >> LOG2M, k = t & (M - 1);
      float2 v = lds[rr * RS + lds_pos<LOG2M>(k, rr)];
      if (P.out_mode == 2) {
        store_c64(dst + (long)(r0 + rr) * M + k, v);
      } else {
        v.x *= P.scale;
        v.y *= P.scale;
        const int oy = (r0 + rr + half_m) & (M - 1), ox = (k + half_m) & (M - 1);
        store_c64(dst + (long)oy * M + ox, v);
      }
    }
    return;
  }
  // transposed store: out[k][row] (consecutive lanes -> consecutive rows)
  for (int t = threadIdx.x; t < total; t += kThreads) {
    const int k = t >> P.log2R, rr = t & (R - 1);
    const int row = r0 + rr;
    float2 v = lds[rr * RS + lds_pos<LOG2M>(k, rr)];
    if (P.out_mode == 0) {
      store_c64(dst + (long)k * M + row, v);
    } else {
      const int oy = (k + half_m) & (M - 1), ox = (row + half_m) & (M - 1);
      v.x *= P.scale;
      v.y *= P.scale;
      if (P.asm_on) v = apply_h(v, oy - half_m, ox - half_m, P);
      store_c64(dst + (long)oy * M + ox, v);
    }
  }
}

int ilog2(int m) {
  int l = 0;
  while ((1 << l) < m) ++l;
  return l;
}

template <int LOG2M>
hipError_t launch_rows_m(const float2* in, float2* out, int batch, const FftParams& P,
                         size_t lds, hipStream_t s) {
  if (lds > 64 * 1024) {
    // Opt in to more than 64 KiB of dynamic LDS (gfx950 has 160 KiB per CU).
    (void)hipFuncSetAttribute((const void*)fft_rows_kernel<LOG2M>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();
  }
  const int rows = P.rows_lo > 0 ? 2 * P.rows_lo : P.M;
  hipLaunchKernelGGL(fft_rows_kernel<LOG2M>, dim3(rows / P.R, batch), dim3(kThreads), lds, s, in,
                     out, P);
  return hipGetLastError();
}

// Rows per workgroup for length-M rows: about two workgroups per CU (512)
// when the rows fit, within kLdsElems and the LDS left beside the M-entry
// twiddle table, at least 4 rows (32-byte runs in the transposed store) while
// that still leaves one workgroup per CU; a power of two.
int rows_per_group(int M) {
  int cap = (kLdsBudget - M) / M;
  if (cap > kLdsElems / M) cap = kLdsElems / M;
  if (cap < 1) cap = 1;
  int r = M / 512;
  const int r_min = M / 256 < 4 ? M / 256 : 4;
  if (r < r_min) r = r_min;
  if (r < 1) r = 1;
  if (r > cap) r = cap;
  int R = 1 << (ilog2(r + 1) - 1);   // largest power of two <= r
  return R > M ? M : R;
}

hipError_t launch_rows(const float2* in, float2* out, int batch, FftParams P, hipStream_t s) {
  P.log2M = ilog2(P.M);
  P.R = rows_per_group(P.M);
  // a launch over the field's rows alone has fewer groups: halve the rows per
  // group while there are fewer than 512 groups (2048^2 padded from 1024^2:
  // 512 groups of 2 rows measured 1 % faster than 256 of 4)
  if (P.rows_lo > 0)
    while (P.R > 1 && 2 * P.rows_lo / P.R < 512 && P.rows_lo % (P.R / 2) == 0) P.R /= 2;
  P.log2R = ilog2(P.R);
  const size_t lds = ((size_t)P.R * P.M + P.M) * sizeof(float2);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  switch (P.log2M) {
    case 1: return launch_rows_m<1>(in, out, batch, P, lds, s);
    case 2: return launch_rows_m<2>(in, out, batch, P, lds, s);
    case 3: return launch_rows_m<3>(in, out, batch, P, lds, s);
    case 4: return launch_rows_m<4>(in, out, batch, P, lds, s);
    case 5: return launch_rows_m<5>(in, out, batch, P, lds, s);
    case 6: return launch_rows_m<6>(in, out, batch, P, lds, s);
    case 7: return launch_rows_m<7>(in, out, batch, P, lds, s);
    case 8: return launch_rows_m<8>(in, out, batch, P, lds, s);
    case 9: return launch_rows_m<9>(in, out, batch, P, lds, s);
    case 10: return launch_rows_m<10>(in, out, batch, P, lds, s);
    case 11: return launch_rows_m<11>(in, out, batch, P, lds, s);
    case 12: return launch_rows_m<12>(in, out, batch, P, lds, s);
    case 13: return launch_rows_m<13>(in, out, batch, P, lds, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

// Maximum transform side handled by the LDS kernel (one row + twiddles fit).
constexpr int kMaxFftSide = 8192;

bool fft_supported(int m) { return m >= 2 && m <= kMaxFftSide && (m & (m - 1)) == 0; }

// Full centred 2-D transform of `batch` N x N fields, padded to M x M.
// tmp: batch * M * M complex64 scratch.
// A launch over the rows of an N x N field padded to M x M (ifftshifted):
// the rows that hold the field are [0, N/2) and [M - N/2, M); transform only
// those when N/2 is a multiple of the rows per workgroup.
static void set_padded_rows(int N, int M, FftParams& A) {
  const int half = N / 2;
  A.rows_lo = N < M && half > 0 && half % rows_per_group(M) == 0 ? half : 0;
}

// The launch after one that skipped zero rows reads, in each of its rows, the
// positions those rows would have written as zeros.
static void set_zero_columns(const FftParams& A, FftParams& B) {
  B.zlo = A.rows_lo > 0 ? A.rows_lo : 0;
  B.zhi = A.rows_lo > 0 ? A.M - A.rows_lo : 0;
}

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
  set_padded_rows(N, M, A);
  hipError_t e = launch_rows(in, tmp, batch, A, s);
  if (e != hipSuccess) return e;
  FftParams B = A;
  B.rows_lo = 0;
  B.in_mode = 1;
  B.in_bstride = (long)M * M;
  B.out_mode = 1;
  B.scale = scale;
  set_zero_columns(A, B);
  B.asm_on = asm_on;
  if (asm_on) {
    B.asm_z = z;
    B.asm_inv_wl2 = 1.0 / (wavelength * wavelength);
    B.asm_df = 1.0 / ((double)M * dx);
  }
  return launch_rows(tmp, out, batch, B, s);
}

// Angular-spectrum propagation of one N x N field padded to M x M:
// out = fftshift(ifft2(fft2(ifftshift(pad x)) H')) with H' the transfer
// function at the natural-order frequencies -- the centred forward and
// inverse transforms with their fftshift / ifftshift pairs cancelled between
// them.  Three launches: x transforms of the padded field (transposed store);
// per y-line (fixed kx) the forward y transform, H', and the inverse y
// transform, in LDS; the inverse x transforms of rows gathered from columns,
// stored with fftshift and the 1/M^2 scale.
hipError_t launch_asm(const float2* in, int N, int M, double z, double wavelength, double dx,
                      float2* tmp, float2* tmp2, float2* out, hipStream_t s) {
  FftParams A{};
  A.M = M;
  A.in_mode = 0;
  A.N = N;
  A.p = (M - N) / 2;
  A.in_bstride = (long)N * N;
  A.out_bstride = (long)M * M;
  A.out_mode = 0;
  set_padded_rows(N, M, A);
  hipError_t e = launch_rows(in, tmp, 1, A, s);
  if (e != hipSuccess) return e;
  FftParams B = A;
  B.rows_lo = 0;
  B.in_mode = 1;
  B.in_bstride = (long)M * M;
  B.out_mode = 2;
  set_zero_columns(A, B);
  B.asm_mid = 1;
  B.asm_z = z;
  B.asm_inv_wl2 = 1.0 / (wavelength * wavelength);
  B.asm_df = 1.0 / ((double)M * dx);
  e = launch_rows(tmp, tmp2, 1, B, s);
  if (e != hipSuccess) return e;
  FftParams C = A;
  C.rows_lo = 0;
  C.inverse = 1;
  C.in_mode = 2;
  C.in_bstride = (long)M * M;
  C.out_mode = 3;
  C.scale = (float)(1.0 / ((double)M * M));
  return launch_rows(tmp2, out, 1, C, s);
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
  for (int b = 0; b < batch; ++b) {
    DLX_CHECK_ARG(std::isfinite(wavelength[b]) && wavelength[b] > 0,
                  "wavelength[%d] must be > 0", b);
    DLX_HIP(launch_asm((const float2*)in + (long)b * n * n, n, m, distance, wavelength[b],
                       pixel_scale, tmp + (long)b * m * m, spec + (long)b * m * m,
                       (float2*)out + (long)b * m * m, s));
  }
  return DLX_OK;
}
