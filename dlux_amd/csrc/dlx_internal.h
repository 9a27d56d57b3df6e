// Internal helpers shared by the dlux_amd HIP translation units (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <cstdio>
#include <string>
#include <vector>
#include <algorithm>

#include "../../include/dlux_amd.h"

namespace dlx {

// Per-thread error message returned by dlx_last_error().
void set_error(const char* fmt, ...);
void clear_error();

#define DLX_CHECK_ARG(cond, ...)          \
  do {                                    \
    if (!(cond)) {                        \
      ::dlx::set_error(__VA_ARGS__);      \
      return DLX_ERR_INVALID;             \
    }                                     \
  } while (0)

#define DLX_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::dlx::set_error("%s failed: %s", #call, hipGetErrorString(_e));       \
      return DLX_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

// Status of the last kernel launch on this thread.
inline int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("launch of %s failed: %s", what, hipGetErrorString(e));
    return DLX_ERR_HIP;
  }
  return DLX_OK;
}

inline size_t align_up(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }

// Device-side per-instance MFT parameters (one per wavelength / source).
struct MftInst {
  double a_scale;   // 1 / N_in
  double a_off;     // -(N_in-1)/2 / N_in - shift_a / N_in    (a_k = k*a_scale + a_off)
  double b_scale;   // p_out / fringe [/(f+df)]
  double b_off;     // (-(N_out-1)/2 - shift_b) * b_scale      (b_l = l*b_scale + b_off)
};

// Host-side builder: the per-axis coordinate vectors of transfer_matrix
// (utils/propagation.py:109-121), in FP64.
inline MftInst make_inst(const dlx_mft_geom& g, double wavelength, double dx_in,
                         double p_out, double shift) {
  MftInst p;
  const double n_in = g.n_in, n_out = g.n_out;
  const double fringe = wavelength / (dx_in * n_in);
  double so = p_out / fringe;
  if (g.has_focal) so /= (g.focal_length + g.focal_shift);
  p.a_scale = 1.0 / n_in;
  p.a_off = -(n_in - 1.0) / 2.0 / n_in - shift / n_in;
  p.b_scale = so;
  p.b_off = (-(n_out - 1.0) / 2.0 - shift) * so;
  return p;
}

// nfringes / (N_in N_out)  (utils/propagation.py:130-175, :246-254).
inline double mft_norm(const dlx_mft_geom& g, double wavelength, double dx_in,
                       double p_out) {
  const double diameter = g.n_in * dx_in;
  const double fringe = wavelength / diameter;
  double out_size = g.n_out * p_out;
  if (g.has_focal) out_size /= (g.focal_length + g.focal_shift);
  const double nfr = out_size / fringe;
  return nfr / ((double)g.n_in * (double)g.n_out);
}

// ---------------------------------------------------------------------------
// kernels implemented in the .hip files, launched from the host orchestration
// ---------------------------------------------------------------------------

// Complex NT GEMM C[b](m,n) = sum_k A[b](m,k) B[b](n,k), then an epilogue.
enum Epilogue {
  EPI_STORE = 0,        // C = alpha * acc (complex64), ldc
  EPI_STORE_T = 1,      // C^T = alpha * acc: written at C[n*ldc + m]
  EPI_INTENSITY = 2,    // I = |alpha * acc|^2 (float32), ldc
  EPI_SPLIT_T = 3,      // split4 matrix with rows n and k = m (ldc / c_batch in bytes)
  // gradient epilogues (G = dL/dPSF, float, ldg / g_batch in elements)
  EPI_VJP = 4,          // C(m,n) = 2 G(m,n) alpha acc (complex64, ldc)
  EPI_VJP_SPLIT_T = 5,  // as EPI_SPLIT_T with value 2 G(n,m) alpha acc (split4 rows n, k = m)
  EPI_GRAD = 6,         // C(m,n) = Im(conj(P(m,n)) alpha acc) (float, ldc); P complex64
                        // (fp32 kernels, p_ld elements) or split4 (p_ld bytes per row);
                        // C2 (if set, same layout as C) = Re(conj(U(m,n)) alpha acc),
                        // U in P's layout
};

struct GemmBatch {
  static constexpr int kMaxBatch = 64;   // instances per launch
  const float2* A; long lda; long sA;
  const float2* B; long ldb; long sB;
  void* C; long ldc; long sC;
  int M, N, K, batch;
  int epilogue;
  int precision;        // DLX_PREC_*
  float alpha[kMaxBatch];   // per-instance scale of the result
  const float* G; long ldg; long g_batch;          // EPI_VJP
  const void* P; long p_ld; long p_batch;           // EPI_GRAD
  float* C2;                                        // EPI_GRAD real part (optional)
  const void* U;                                    // EPI_GRAD: phasor for C2, P's layout
};

hipError_t launch_cgemm_nt(const GemmBatch& g, hipStream_t s);

// Split-BF16x3 complex NT GEMM on split4 operands (dlx_split.h).  Row strides
// and batch strides of A and B are in bytes; C strides are in elements
// (EPI_STORE / EPI_INTENSITY) or bytes (EPI_SPLIT_T).  alpha is per batch.
struct GemmSplit {
  static constexpr int kMaxBatch = 64;
  const uint8_t* A; long a_row_bytes; long a_batch_bytes;
  const uint8_t* B; long b_row_bytes; long b_batch_bytes;
  void* C; long ldc; long c_batch;
  int M, N, K, batch, epilogue;
  float alpha[kMaxBatch];
  const float* G; long ldg; long g_batch;          // EPI_VJP_SPLIT_T
  const void* P; long p_ld; long p_batch;           // EPI_GRAD (split4, bytes)
  float* C2;                                        // EPI_GRAD real part (optional)
  const void* U;                                    // EPI_GRAD: phasor for C2, P's layout
};

hipError_t launch_cgemm_split4(const GemmSplit& g, hipStream_t s);

// The pupil of instance b, generated in the stage-1 loader instead of read
// from A:  A[b](y, x) = a(y, x) exp(2 pi i path(y, x) inv_wl[b]) with
// path = tx X + ty Y + opd (FP64) and a = scale amp, formed from the caller's
// amplitude and OPD planes (n x n, row-major, float; either may be null) as
// the pupil kernels form them (dlx_phasor.h).
struct PupilGen {
  const float* opd; const float* amp;
  double dx, x0, tx, ty;
  float scale;
  int n;
  double inv_wl[GemmSplit::kMaxBatch];
};
// Stage 1 (EPI_SPLIT_T, K = M = pg.n) with the pupil as A.  Returns
// hipErrorNotSupported, launching nothing, for shapes it does not cover
// (cgemm_pupil_supported is false; the caller then writes the pupil and calls
// launch_cgemm_split4).
hipError_t launch_cgemm_pupil(const GemmSplit& g, const PupilGen& pg, hipStream_t s);
bool cgemm_pupil_supported(const GemmSplit& g, const PupilGen& pg);

// One problem of a grouped (uneven) split4 GEMM, in device memory.  Its
// 128 x tile_n tiles are [tile_start, tile_start + tiles_m * tiles_n) of the
// launch's tile list, row-major over (tm, tn).
struct GroupProblem {
  const uint8_t* A; const uint8_t* B; void* C;
  long a_row_bytes, b_row_bytes, ldc;
  int M, N, K, tiles_n, tiles_m, tile_start;
  float alpha; int pad_;
};
struct GroupLaunch {
  const GroupProblem* probs;   // device array, sorted by tile_start
  int nprobs;
  int tile_n;                  // 128, or 64 for launches of problems with N <= 64
  int xcd_start[9];            // tile range of XCD x: [xcd_start[x], xcd_start[x+1])
};
// Epilogues supported: EPI_STORE, EPI_SPLIT_T, EPI_INTENSITY.
hipError_t launch_cgemm_split4_group(const GroupLaunch& gl, int epilogue, hipStream_t s);

// Per-launch instance parameters passed by value (kernel arguments), so no
// device-side parameter buffer is needed.
struct DftChunk {
  static constexpr int kMax = 48;
  double a_scale[kMax], a_off[kMax], b_scale[kMax], b_off[kMax], amp[kMax];
};
struct PupilChunk {
  static constexpr int kMax = 128;
  double inv_wl[kMax];
};

struct GradChunk {
  static constexpr int kMax = 64;
  float kw[kMax];   // 2 pi / lambda per plane
};

hipError_t launch_dft_chunk(const DftChunk& c, int nb, int n_in, int n_out, int inverse,
                            float2* out, long bstride, hipStream_t s);

// Linear canonical transforms (dlx_lct.hip): per-instance phase polynomials in turns.
struct LctMatChunk {
  static constexpr int kMax = 48;
  double ci2[kMax], co2[kMax], cx[kMax], c0[kMax];
};
struct LctChirpChunk {
  static constexpr int kMax = 48;
  double q2[kMax], q1[kMax], q0[kMax], x0[kMax], dx[kMax];   // grid x_k = x0 + k dx
};
hipError_t launch_lct_matrix(const LctMatChunk& c, int nb, int n_in, int n_out, double x0_in,
                             double dx_in, double x0_out, double dx_out, float2* out,
                             long bstride, hipStream_t s);
hipError_t launch_lct_matrix_split4(const LctMatChunk& c, int nb, int n_in, int n_out,
                                    double x0_in, double dx_in, double x0_out, double dx_out,
                                    uint8_t* out, long batch_bytes, hipStream_t s);
hipError_t launch_lct_chirp(const LctChirpChunk& c, int nb, int n, const float2* in,
                            float2* out, hipStream_t s);
hipError_t launch_pupil_chunk(const PupilChunk& c, int nb, int n, double dx, double x0,
                              double tx, double ty, const float* amp, const float* opd,
                              const float* phase, float scale, const float2* in,
                              float2* out, hipStream_t s);
// split4 producers
hipError_t launch_dft_split4(const DftChunk& c, int nb, int n_in, int n_out, int inverse,
                             uint8_t* out, long batch_bytes, hipStream_t s);
hipError_t launch_pupil_split4(const PupilChunk& c, int nb, int n, double dx, double x0,
                               double tx, double ty, const float* amp, const float* opd,
                               const float* phase, float scale, uint8_t* out, long batch_bytes,
                               hipStream_t s);
hipError_t launch_to_split4(const float2* in, int batch, long rows, int k, long in_batch,
                            uint8_t* out, long out_batch_bytes, hipStream_t s);
hipError_t launch_eval_basis(int nbasis, long npix, int nsets, const float* basis,
                             const float* coeffs, const float* base, float* out,
                             hipStream_t s);
hipError_t launch_power(long npix, int batch, const float2* f, float* out, hipStream_t s);
hipError_t launch_intensity(long npix, int batch, const float2* f, int sum_batch, float* out,
                            hipStream_t s);
hipError_t launch_sum_planes(long npix, int nplanes, const float* planes, int accumulate,
                             float* psf, hipStream_t s);

// Backward reductions (dlx_grad.hip).  partial holds nbasis * nsets * project_parts(npix)
// floats.
hipError_t launch_grad_reduce(const GradChunk& c, int nb, long npix, const float* h,
                              int accumulate, float* g_opd, float* g_phase, hipStream_t s);
int project_parts(long npix);
hipError_t launch_project_basis(int nbasis, long npix, int nsets, const float* basis,
                                const float* g, float* partial, float* out, hipStream_t s);

// Bytes of one batch instance of an R x K complex operand (complex64 or split4).
size_t mat_bytes(long rows, long k, bool split);

// Generate `nb` transfer matrices (optionally scaled by amp[b]) into `out`:
// W (N_out x N_in) or, with `adjoint`, its conjugate transpose W^H (N_in x N_out).
int gen_dft(const dlx_mft_geom& g, int nb, const double* wl, const double* dx,
            const double* pout, const double* shift, const double* amp, void* out, bool split,
            hipStream_t s, bool adjoint = false);

// Device-resident transfer-matrix cache (dlx_cache.hip).  On success *ptr is
// either the cached matrices (ready for work enqueued on s) or NULL when the
// cache was bypassed; cache_release records the entry's last use on s.
int cache_acquire(const dlx_mft_geom& g, int nb, const double* wl, const double* dx,
                  const double* pout, const double* shift, bool split, hipStream_t s,
                  void** ptr, void** handle, bool adjoint = false);
int cache_release(void* handle, hipStream_t s);
// Raw entries keyed by the caller: on a hit (*fresh false) the block is ready
// for work on s; on a miss (*fresh true) the caller fills it on s and then
// calls cache_ready.  *ptr == NULL (and *fresh true) when bypassed (capture,
// budget, memory): the caller generates into its own buffer.
int cache_acquire_raw(const std::vector<double>& key, size_t bytes, hipStream_t s,
                      void** ptr, void** handle, bool* fresh);
int cache_ready(void* handle, hipStream_t s);

// The detector chain (dlx_detector.hip) on one image that is the in-order sum
// of nsum planes sum_stride floats apart (nsum = 1: `batch` images), and its
// workspace size for nsum planes n^2 floats apart.
size_t detector_ws_bytes(int n, int batch, const dlx_detector_op* ops, int nops, int nsum = 1);
int detector_apply_sum(int n, int batch, int nsum, long sum_stride, const float* in,
                       const dlx_detector_op* ops, int nops, float* out, void* workspace,
                       size_t workspace_bytes, hipStream_t s);

// Holds a cache entry from its acquire (pass &lease.h) to its release, and
// releases it on scope exit too, so an error return never leaves an entry
// pinned.  release() records the last use on the stream after the work that
// reads the entry has been enqueued.
struct CacheLease {
  void* h = nullptr;
  hipStream_t s = nullptr;
  explicit CacheLease(hipStream_t st) : s(st) {}
  CacheLease(const CacheLease&) = delete;
  CacheLease& operator=(const CacheLease&) = delete;
  int release() {
    void* x = h;
    h = nullptr;
    return cache_release(x, s);
  }
  ~CacheLease() {
    if (h) (void)cache_release(h, s);
  }
};

}  // namespace dlx
