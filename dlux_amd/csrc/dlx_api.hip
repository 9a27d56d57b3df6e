// extern "C" entry points of libdlux_amd (declared in include/dlux_amd.h):
// argument validation, workspace carving and kernel orchestration.
#include <cstdarg>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "dlx_internal.h"
#include "dlx_split.h"

namespace dlx {

namespace {
thread_local std::string g_err;
}

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}
void clear_error() { g_err.clear(); }

namespace {

// Instances per fused launch of the polychromatic PSF (bounds the workspace).
constexpr int kWlChunk = 64;

// The pupil generated in stage 1's loader (default); DLUX_AMD_PUPIL_PROLOGUE=0
// writes it with the pupil kernel instead (read at every call, for
// comparisons).
bool pupil_prologue_enabled() {
  const char* v = getenv("DLUX_AMD_PUPIL_PROLOGUE");
  return !(v && v[0] == '0');
}

// Optional live timing of the two MFT contraction launches of dlx_psf_poly:
// hipEvents recorded on the caller's stream around each launch.
struct ProfSlot {
  hipEvent_t beg, end;
  int stage;
};
constexpr int kProfSlots = 8192;
struct Profiler {
  int enabled = 0;
  int n = 0;
  std::vector<ProfSlot> slots;
} g_prof;

int prof_begin(int stage, hipStream_t s, int* slot) {
  *slot = -1;
  if (!g_prof.enabled || g_prof.n >= kProfSlots) return DLX_OK;
  if ((int)g_prof.slots.size() <= g_prof.n) {
    ProfSlot ps{};
    DLX_HIP(hipEventCreate(&ps.beg));
    DLX_HIP(hipEventCreate(&ps.end));
    g_prof.slots.push_back(ps);
  }
  *slot = g_prof.n++;
  g_prof.slots[*slot].stage = stage;
  DLX_HIP(hipEventRecord(g_prof.slots[*slot].beg, s));
  return DLX_OK;
}

int prof_end(int slot, hipStream_t s) {
  if (slot < 0) return DLX_OK;
  DLX_HIP(hipEventRecord(g_prof.slots[slot].end, s));
  return DLX_OK;
}

int check_geom(const dlx_mft_geom* g) {
  DLX_CHECK_ARG(g != nullptr, "geom is NULL");
  DLX_CHECK_ARG(g->n_in > 0 && g->n_in <= 16384, "n_in must be in [1, 16384], got %d", g->n_in);
  DLX_CHECK_ARG(g->n_out > 0 && g->n_out <= 16384, "n_out must be in [1, 16384], got %d",
                g->n_out);
  if (g->has_focal)
    DLX_CHECK_ARG(std::isfinite(g->focal_length + g->focal_shift) &&
                      (g->focal_length + g->focal_shift) != 0.0,
                  "focal_length + focal_shift must be finite and non-zero");
  return DLX_OK;
}

}  // namespace

// Bytes of one batch instance of an R x K complex operand.
size_t mat_bytes(long rows, long k, bool split) {
  return split ? (size_t)rows * split4_row_bytes(k) : (size_t)rows * k * sizeof(float2);
}

// Generate `nb` transfer matrices with per-instance (lambda, dx, p_out, shift,
// amplitude) into out[b] -- complex64 or split4 -- in launches of DftChunk::kMax.
// The generator writes exp(sign 2 pi i u_r v_c) over rows r and columns c; the
// adjoint W^H (rows k of the input plane, columns l of the output plane) is the
// same phase with the two coordinate vectors swapped and the sign flipped.
int gen_dft(const dlx_mft_geom& g, int nb, const double* wl, const double* dx,
            const double* pout, const double* shift, const double* amp, void* out, bool split,
            hipStream_t s, bool adjoint) {
  const int rows = adjoint ? g.n_in : g.n_out, cols = adjoint ? g.n_out : g.n_in;
  const int sign_inv = adjoint ? !g.inverse : g.inverse;
  const long stride = (long)rows * cols;
  const long bbytes = (long)mat_bytes(rows, cols, true);
  for (int b0 = 0; b0 < nb; b0 += DftChunk::kMax) {
    const int n = nb - b0 < DftChunk::kMax ? nb - b0 : DftChunk::kMax;
    DftChunk c;
    for (int i = 0; i < n; ++i) {
      const int b = b0 + i;
      DLX_CHECK_ARG(std::isfinite(wl[b]) && wl[b] > 0, "wavelength[%d] must be > 0", b);
      DLX_CHECK_ARG(std::isfinite(dx[b]) && dx[b] != 0, "pixel_scale_in[%d] must be non-zero", b);
      DLX_CHECK_ARG(std::isfinite(pout[b]), "pixel_scale_out[%d] must be finite", b);
      const MftInst p = make_inst(g, wl[b], dx[b], pout[b], shift ? shift[b] : 0.0);
      // columns carry the a (input-plane) vector, rows the b (output-plane) one
      c.a_scale[i] = adjoint ? p.b_scale : p.a_scale;
      c.a_off[i] = adjoint ? p.b_off : p.a_off;
      c.b_scale[i] = adjoint ? p.a_scale : p.b_scale;
      c.b_off[i] = adjoint ? p.a_off : p.b_off;
      c.amp[i] = amp ? amp[b] : 1.0;
    }
    if (split) {
      DLX_HIP(launch_dft_split4(c, n, cols, rows, sign_inv,
                                (uint8_t*)out + (long)b0 * bbytes, bbytes, s));
    } else {
      DLX_HIP(launch_dft_chunk(c, n, cols, rows, sign_inv,
                               (float2*)out + (long)b0 * stride, stride, s));
    }
  }
  return DLX_OK;
}

namespace {

// Transfer matrices for `nb` instances: from the device cache, or generated
// into `fallback` when the cache is bypassed (graph capture, no memory).
int get_matrices(const dlx_mft_geom& g, int nb, const double* wl, const double* dx,
                 const double* pout, const double* shift, bool split, void* fallback,
                 hipStream_t s, void** out, void** handle, bool adjoint = false) {
  if (int st = cache_acquire(g, nb, wl, dx, pout, shift, split, s, out, handle, adjoint))
    return st;
  if (*out) return DLX_OK;
  *out = fallback;
  return gen_dft(g, nb, wl, dx, pout, shift, nullptr, fallback, split, s, adjoint);
}

// FP32 two-stage MFT of `batch` complex64 fields P (N_in x N_in) against
// complex64 transfer matrices Wx / Wy (N_out x N_in):
//   stage 1: U^T = (P Wx^T)^T (complex64, N_out x N_in)
//   stage 2: E = alpha_b Wy U, stored complex64 (EPI_STORE) or |E|^2 (EPI_INTENSITY).
int mft_fp32(const dlx_mft_geom& g, int batch, const float2* P, const float2* Wx,
             const float2* Wy, float2* U, const double* alpha, void* out, int epi,
             hipStream_t s, bool prof) {
  const long pstride = (long)g.n_in * g.n_in;
  const long mstride = (long)g.n_out * g.n_in;
  const long nout2 = (long)g.n_out * g.n_out;
  for (int b0 = 0; b0 < batch; b0 += GemmBatch::kMaxBatch) {
    const int nb = batch - b0 < GemmBatch::kMaxBatch ? batch - b0 : GemmBatch::kMaxBatch;
    GemmBatch g1{};
    g1.A = P + (long)b0 * pstride; g1.lda = g.n_in; g1.sA = pstride;
    g1.B = Wx + (long)b0 * mstride; g1.ldb = g.n_in; g1.sB = mstride;
    g1.C = U + (long)b0 * mstride; g1.ldc = g.n_in; g1.sC = mstride;
    g1.M = g.n_in; g1.N = g.n_out; g1.K = g.n_in; g1.batch = nb;
    g1.epilogue = EPI_STORE_T; g1.precision = DLX_PREC_FP32;
    for (int i = 0; i < nb; ++i) g1.alpha[i] = 1.0f;
    int slot = -1;
    if (prof) { if (int st = prof_begin(1, s, &slot)) return st; }
    DLX_HIP(launch_cgemm_nt(g1, s));
    if (prof) { if (int st = prof_end(slot, s)) return st; }
    GemmBatch g2{};
    g2.A = Wy + (long)b0 * mstride; g2.lda = g.n_in; g2.sA = mstride;
    g2.B = U + (long)b0 * mstride; g2.ldb = g.n_in; g2.sB = mstride;
    g2.C = epi == EPI_STORE ? (void*)((float2*)out + (long)b0 * nout2)
                            : (void*)((float*)out + (long)b0 * nout2);
    g2.ldc = g.n_out; g2.sC = nout2;
    g2.M = g.n_out; g2.N = g.n_out; g2.K = g.n_in; g2.batch = nb;
    g2.epilogue = epi; g2.precision = DLX_PREC_FP32;
    for (int i = 0; i < nb; ++i) g2.alpha[i] = (float)alpha[b0 + i];
    if (prof) { if (int st = prof_begin(2, s, &slot)) return st; }
    DLX_HIP(launch_cgemm_nt(g2, s));
    if (prof) { if (int st = prof_end(slot, s)) return st; }
  }
  return DLX_OK;
}

// Split-BF16x3 two-stage MFT of `batch` split4 pupils P (N_in rows, k = N_in)
// against split4 transfer matrices Wx / Wy (N_out rows, k = N_in):
//   stage 1: U^T = (P Wx^T)^T written split4 (N_out rows, k = N_in)
//   stage 2: E = alpha_b Wy U, stored complex64 (EPI_STORE) or |E|^2 (EPI_INTENSITY).
// Stage 1's launch description for instances [b0, b0 + nb) of a batch.
GemmSplit stage1_split(const dlx_mft_geom& g, int b0, int nb, const uint8_t* P,
                       const uint8_t* Wx, uint8_t* U) {
  const long pbytes = (long)mat_bytes(g.n_in, g.n_in, true);
  const long wbytes = (long)mat_bytes(g.n_out, g.n_in, true);
  const long rb = split4_row_bytes(g.n_in);
  GemmSplit g1{};
  g1.A = P ? P + (long)b0 * pbytes : nullptr; g1.a_row_bytes = rb; g1.a_batch_bytes = pbytes;
  g1.B = Wx + (long)b0 * wbytes; g1.b_row_bytes = rb; g1.b_batch_bytes = wbytes;
  g1.C = U + (long)b0 * wbytes; g1.ldc = rb; g1.c_batch = wbytes;
  g1.M = g.n_in; g1.N = g.n_out; g1.K = g.n_in; g1.batch = nb; g1.epilogue = EPI_SPLIT_T;
  for (int i = 0; i < nb; ++i) g1.alpha[i] = 1.0f;
  return g1;
}

// With `pg` set, stage 1 generates the pupil in its loader (P unused; the
// caller has checked cgemm_pupil_supported for every slice).
int mft_split4(const dlx_mft_geom& g, int batch, const uint8_t* P, const uint8_t* Wx,
               const uint8_t* Wy, uint8_t* U, const double* alpha, void* out, int epi,
               hipStream_t s, int (*pb)(int, hipStream_t, int*), int (*pe)(int, hipStream_t),
               const PupilGen* pg = nullptr) {
  const long wbytes = (long)mat_bytes(g.n_out, g.n_in, true);
  const long rb = split4_row_bytes(g.n_in);
  const long nout2 = (long)g.n_out * g.n_out;
  for (int b0 = 0; b0 < batch; b0 += GemmSplit::kMaxBatch) {
    const int nb = batch - b0 < GemmSplit::kMaxBatch ? batch - b0 : GemmSplit::kMaxBatch;
    const GemmSplit g1 = stage1_split(g, b0, nb, P, Wx, U);
    int slot = -1;
    if (pb) { if (int st = pb(1, s, &slot)) return st; }
    if (pg) {
      PupilGen p = *pg;
      for (int i = 0; i < nb; ++i) p.inv_wl[i] = pg->inv_wl[b0 + i];
      DLX_HIP(launch_cgemm_pupil(g1, p, s));
    } else {
      DLX_HIP(launch_cgemm_split4(g1, s));
    }
    if (pe) { if (int st = pe(slot, s)) return st; }
    GemmSplit g2{};
    g2.A = Wy + (long)b0 * wbytes; g2.a_row_bytes = rb; g2.a_batch_bytes = wbytes;
    g2.B = U + (long)b0 * wbytes; g2.b_row_bytes = rb; g2.b_batch_bytes = wbytes;
    g2.ldc = g.n_out;
    g2.c_batch = nout2;
    g2.C = epi == EPI_STORE ? (void*)((float2*)out + (long)b0 * nout2)
                            : (void*)((float*)out + (long)b0 * nout2);
    g2.M = g.n_out; g2.N = g.n_out; g2.K = g.n_in; g2.batch = nb; g2.epilogue = epi;
    for (int i = 0; i < nb; ++i) g2.alpha[i] = (float)alpha[b0 + i];
    if (pb) { if (int st = pb(2, s, &slot)) return st; }
    DLX_HIP(launch_cgemm_split4(g2, s));
    if (pe) { if (int st = pe(slot, s)) return st; }
  }
  return DLX_OK;
}

// Workspace of the VJP for one chunk of c wavelengths: pupils P, forward and
// adjoint matrices (used when the cache is bypassed), the stage-1 intermediate
// (U^T, later V^T), Ebar, and the per-wavelength gradient planes.
struct VjpLayout {
  size_t P, U, Wf, Wa, T, Eb, H, H2, total;
};
VjpLayout vjp_layout(const dlx_mft_geom& g, int c, bool split) {
  VjpLayout L;
  L.P = align_up((size_t)c * mat_bytes(g.n_in, g.n_in, split));
  L.Wf = align_up((size_t)c * mat_bytes(g.n_out, g.n_in, split));
  L.Wa = align_up((size_t)c * mat_bytes(g.n_in, g.n_out, split));
  const size_t t1 = mat_bytes(g.n_out, g.n_in, split), t2 = mat_bytes(g.n_in, g.n_out, split);
  L.T = align_up((size_t)c * (t1 > t2 ? t1 : t2));
  L.Eb = align_up((size_t)c * mat_bytes(g.n_out, g.n_out, split));
  L.H = align_up((size_t)c * g.n_in * g.n_in * sizeof(float));
  // the amplitude cotangent: phasors U = P / A (P at unit amplitude) and the
  // Re(conj(U) Pbar) planes
  L.U = L.P;
  L.H2 = L.H;
  L.total = L.P + L.U + L.Wf + L.Wa + L.T + L.Eb + L.H + L.H2;
  return L;
}

}  // namespace
}  // namespace dlx

using namespace dlx;

extern "C" {

const char* dlx_version(void) { return "dlux_amd 0.1.0 gfx950"; }

const char* dlx_last_error(void) { return dlx::g_err.c_str(); }

int dlx_transfer_matrix(const dlx_mft_geom* geom, int32_t batch, const double* wavelength,
                        const double* pixel_scale_in, const double* pixel_scale_out,
                        const double* shift, void* out, void* stream) {
  clear_error();
  if (int st = check_geom(geom)) return st;
  DLX_CHECK_ARG(batch >= 0, "batch must be >= 0");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(wavelength && pixel_scale_in && pixel_scale_out && out,
                "null argument to dlx_transfer_matrix");
  return gen_dft(*geom, batch, wavelength, pixel_scale_in, pixel_scale_out, shift, nullptr,
                 out, false, (hipStream_t)stream);
}

size_t dlx_mft_workspace_size(const dlx_mft_geom* geom, int32_t batch, int32_t precision) {
  if (!geom || batch <= 0) return 0;
  const bool split = precision == DLX_PREC_BF16X3;
  const size_t mat = align_up((size_t)batch * mat_bytes(geom->n_out, geom->n_in, split));
  size_t total = 3 * mat;  // W_x, W_y, U^T
  if (split) total += align_up((size_t)batch * mat_bytes(geom->n_in, geom->n_in, true));
  return total;
}

int dlx_mft(const dlx_mft_geom* geom, int32_t batch, const void* in, void* out,
            const double* wavelength, const double* pixel_scale_in,
            const double* pixel_scale_out, const double* shift_x, const double* shift_y,
            const double* field_scale, int32_t precision, void* workspace,
            size_t workspace_bytes, void* stream) {
  clear_error();
  if (int st = check_geom(geom)) return st;
  DLX_CHECK_ARG(batch >= 0, "batch must be >= 0");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(in && out && wavelength && pixel_scale_in && pixel_scale_out,
                "null argument to dlx_mft");
  DLX_CHECK_ARG(precision == DLX_PREC_FP32 || precision == DLX_PREC_BF16X3,
                "unknown precision %d", precision);
  const size_t need = dlx_mft_workspace_size(geom, batch, precision);
  if (!workspace || workspace_bytes < need) {
    set_error("dlx_mft workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  const dlx_mft_geom& g = *geom;
  hipStream_t s = (hipStream_t)stream;
  const bool split = precision == DLX_PREC_BF16X3;
  const size_t mat = align_up((size_t)batch * mat_bytes(g.n_out, g.n_in, split));
  uint8_t* Wx_ws = (uint8_t*)workspace;
  uint8_t* Wy_ws = Wx_ws + mat;
  uint8_t* Ut = Wx_ws + 2 * mat;

  bool same = true;
  for (int b = 0; b < batch && same; ++b) {
    const double sx = shift_x ? shift_x[b] : 0.0, sy = shift_y ? shift_y[b] : 0.0;
    same = (sx == sy);
  }
  // c_b * field_scale_b is applied in the stage-2 epilogue: the matrices are
  // unscaled and therefore cacheable.
  std::vector<double> c(batch);
  for (int b = 0; b < batch; ++b)
    c[b] = mft_norm(g, wavelength[b], pixel_scale_in[b], pixel_scale_out[b]) *
           (field_scale ? field_scale[b] : 1.0);

  void *Wx = nullptr, *Wy = nullptr;
  CacheLease hx(s), hy(s);
  if (int st = get_matrices(g, batch, wavelength, pixel_scale_in, pixel_scale_out, shift_x,
                            split, Wx_ws, s, &Wx, &hx.h))
    return st;
  if (same) {
    Wy = Wx;
  } else if (int st = get_matrices(g, batch, wavelength, pixel_scale_in, pixel_scale_out,
                                   shift_y, split, Wy_ws, s, &Wy, &hy.h)) {
    return st;
  }
  int st;
  if (split) {
    uint8_t* X = Wx_ws + 3 * mat;
    DLX_HIP(launch_to_split4((const float2*)in, batch, g.n_in, g.n_in, (long)g.n_in * g.n_in,
                             X, (long)mat_bytes(g.n_in, g.n_in, true), s));
    st = mft_split4(g, batch, X, (const uint8_t*)Wx, (const uint8_t*)Wy, Ut, c.data(), out,
                    EPI_STORE, s, nullptr, nullptr);
  } else {
    st = mft_fp32(g, batch, (const float2*)in, (const float2*)Wx, (const float2*)Wy,
                  (float2*)Ut, c.data(), out, EPI_STORE, s, false);
  }
  if (st) return st;
  if (int r = hx.release()) return r;
  return hy.release();
}

size_t dlx_lct_workspace_size(int32_t n_in, int32_t n_out, int32_t batch, int32_t precision) {
  if (n_in <= 0 || n_out <= 0 || batch <= 0) return 0;
  const bool split = precision == DLX_PREC_BF16X3;
  size_t total = 2 * align_up((size_t)batch * mat_bytes(n_out, n_in, split));  // K, U^T
  if (split) total += align_up((size_t)batch * mat_bytes(n_in, n_in, true));
  return total;
}

int dlx_lct(int32_t n_in, int32_t n_out, int32_t batch, const double* abcd,
            const double* wavelength, double pixel_scale_in, double center_in,
            double pixel_scale_out, double center_out, const void* in, void* out,
            int32_t precision, void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  DLX_CHECK_ARG(n_in > 0 && n_in <= 16384 && n_out > 0 && n_out <= 16384,
                "n_in and n_out must be in [1, 16384], got %d and %d", n_in, n_out);
  DLX_CHECK_ARG(batch >= 0 && batch <= 65535, "batch must be in [0, 65535], got %d", batch);
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(abcd && wavelength && in && out, "null argument to dlx_lct");
  DLX_CHECK_ARG(precision == DLX_PREC_FP32 || precision == DLX_PREC_BF16X3,
                "unknown precision %d", precision);
  const double A = abcd[0], B = abcd[1], D = abcd[3];
  DLX_CHECK_ARG(std::isfinite(A) && std::isfinite(D) && std::isfinite(B) && B != 0.0,
                "the ABCD matrix needs finite A and D and a finite, non-zero B");
  DLX_CHECK_ARG(std::isfinite(pixel_scale_in) && pixel_scale_in > 0 &&
                    std::isfinite(pixel_scale_out) && pixel_scale_out > 0,
                "pixel scales must be > 0");
  DLX_CHECK_ARG(std::isfinite(center_in) && std::isfinite(center_out), "centres must be finite");
  for (int b = 0; b < batch; ++b)
    DLX_CHECK_ARG(std::isfinite(wavelength[b]) && wavelength[b] > 0,
                  "wavelength[%d] must be > 0", b);
  const size_t need = dlx_lct_workspace_size(n_in, n_out, batch, precision);
  if (!workspace || workspace_bytes < need) {
    set_error("dlx_lct workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  const bool split = precision == DLX_PREC_BF16X3;
  const size_t mat = align_up((size_t)batch * mat_bytes(n_out, n_in, split));
  uint8_t* K_ws = (uint8_t*)workspace;
  uint8_t* Ut = K_ws + mat;
  const long kst = (long)n_out * n_in;
  const long kbytes = (long)mat_bytes(n_out, n_in, split);
  const double x0i = center_in - (n_in - 1) / 2.0 * pixel_scale_in;
  const double x0o = center_out - (n_out - 1) / 2.0 * pixel_scale_out;
  std::vector<double> alpha(batch);
  for (int b = 0; b < batch; ++b)
    alpha[b] = pixel_scale_in * pixel_scale_out / (wavelength[b] * std::fabs(B));
  // the kernels depend only on the grids, the ABCD matrix and lambda: keep them
  // resident in the device cache across calls (generated on a miss)
  std::vector<double> key = {1001.0, (double)n_in, (double)n_out, (double)batch,
                             split ? 1.0 : 0.0, A, B, D, pixel_scale_in, center_in,
                             pixel_scale_out, center_out};
  key.insert(key.end(), wavelength, wavelength + batch);
  void* cached = nullptr;
  CacheLease lease(s);
  bool fresh = true;
  if (int st = cache_acquire_raw(key, (size_t)batch * kbytes, s, &cached, &lease.h, &fresh))
    return st;
  uint8_t* Kp = cached ? (uint8_t*)cached : K_ws;
  if (fresh) {
    for (int b0 = 0; b0 < batch; b0 += LctMatChunk::kMax) {
      const int nb = batch - b0 < LctMatChunk::kMax ? batch - b0 : LctMatChunk::kMax;
      LctMatChunk c;
      for (int i = 0; i < nb; ++i) {
        const double lb = wavelength[b0 + i] * B;
        c.ci2[i] = A / (2.0 * lb);
        c.co2[i] = D / (2.0 * lb);
        c.cx[i] = -1.0 / lb;
        c.c0[i] = B > 0 ? -0.125 : 0.125;   // the phase of 1 / sqrt(i B), per axis
      }
      if (split)
        DLX_HIP(launch_lct_matrix_split4(c, nb, n_in, n_out, x0i, pixel_scale_in, x0o,
                                         pixel_scale_out, Kp + (long)b0 * kbytes, kbytes, s));
      else
        DLX_HIP(launch_lct_matrix(c, nb, n_in, n_out, x0i, pixel_scale_in, x0o,
                                  pixel_scale_out, (float2*)Kp + (long)b0 * kst, kst, s));
    }
    if (int st = cache_ready(lease.h, s)) return st;
  }
  dlx_mft_geom g{};
  g.n_in = n_in;
  g.n_out = n_out;
  int st;
  if (split) {
    uint8_t* X = K_ws + 2 * mat;
    DLX_HIP(launch_to_split4((const float2*)in, batch, n_in, n_in, (long)n_in * n_in, X,
                             (long)mat_bytes(n_in, n_in, true), s));
    st = mft_split4(g, batch, X, Kp, Kp, Ut, alpha.data(), out, EPI_STORE, s, nullptr, nullptr);
  } else {
    st = mft_fp32(g, batch, (const float2*)in, (const float2*)Kp, (const float2*)Kp,
                  (float2*)Ut, alpha.data(), out, EPI_STORE, s, false);
  }
  if (int r = lease.release()) return r;
  return st;
}

size_t dlx_lct_fft_workspace_size(int32_t n, int32_t m, int32_t batch) {
  if (n <= 0 || m <= 0 || batch <= 0) return 0;
  return align_up((size_t)batch * n * n * sizeof(float2)) + dlx_fft2_workspace_size(n, m, batch);
}

int dlx_lct_fft(int32_t n, int32_t m, int32_t batch, const double* abcd,
                const double* wavelength, double pixel_scale_in, double center_in,
                double center_out, const void* in, void* out, double* pixel_scale_out,
                void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  DLX_CHECK_ARG(n > 0 && m >= n && ((m - n) % 2) == 0,
                "need 0 < n <= m with m - n even, got n = %d, m = %d", n, m);
  DLX_CHECK_ARG(batch >= 0 && batch <= 65535, "batch must be in [0, 65535], got %d", batch);
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(abcd && wavelength && in && out, "null argument to dlx_lct_fft");
  const double A = abcd[0], B = abcd[1], D = abcd[3];
  DLX_CHECK_ARG(std::isfinite(A) && std::isfinite(D) && std::isfinite(B) && B != 0.0,
                "the ABCD matrix needs finite A and D and a finite, non-zero B");
  DLX_CHECK_ARG(std::isfinite(pixel_scale_in) && pixel_scale_in > 0, "pixel_scale_in must be > 0");
  DLX_CHECK_ARG(std::isfinite(center_in) && std::isfinite(center_out), "centres must be finite");
  for (int b = 0; b < batch; ++b)
    DLX_CHECK_ARG(std::isfinite(wavelength[b]) && wavelength[b] > 0,
                  "wavelength[%d] must be > 0", b);
  const size_t need = dlx_lct_fft_workspace_size(n, m, batch);
  if (!workspace || workspace_bytes < need) {
    set_error("dlx_lct_fft workspace too small: need %zu bytes, got %zu", need,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  const double sg = B > 0 ? 1.0 : -1.0, dx = pixel_scale_in, M = m;
  // the centred FFT pairs offsets from index m / 2; the grids are symmetric
  // about (m - 1) / 2, de apart
  const double de = (m / 2) - (m - 1) / 2.0;
  float2* C1 = (float2*)workspace;
  uint8_t* fws = (uint8_t*)workspace + align_up((size_t)batch * n * n * sizeof(float2));
  const long nn = (long)n * n, mm = (long)m * m;
  // input chirp (A term), with the linear terms that centre the output on center_out
  for (int b0 = 0; b0 < batch; b0 += LctChirpChunk::kMax) {
    const int nb = batch - b0 < LctChirpChunk::kMax ? batch - b0 : LctChirpChunk::kMax;
    LctChirpChunk c;
    for (int i = 0; i < nb; ++i) {
      const double lb = wavelength[b0 + i] * B;
      c.q2[i] = A / (2.0 * lb);
      c.q1[i] = -center_out / lb - sg * de / (M * dx);
      c.q0[i] = center_out * center_in / lb + sg * de * center_in / (M * dx) + sg * de * de / M;
      c.x0[i] = center_in - (n - 1) / 2.0 * dx;
      c.dx[i] = dx;
    }
    DLX_HIP(launch_lct_chirp(c, nb, n, (const float2*)in + (long)b0 * nn, C1 + (long)b0 * nn,
                             s));
  }
  if (int st = dlx_fft2(n, m, batch, sg < 0, C1, out, fws, dlx_fft2_workspace_size(n, m, batch),
                        stream))
    return st;
  // output chirp (D term), the output-side linear terms and the constant phase
  for (int b0 = 0; b0 < batch; b0 += LctChirpChunk::kMax) {
    const int nb = batch - b0 < LctChirpChunk::kMax ? batch - b0 : LctChirpChunk::kMax;
    LctChirpChunk c;
    for (int i = 0; i < nb; ++i) {
      const double lb = wavelength[b0 + i] * B;
      const double d_out = wavelength[b0 + i] * std::fabs(B) / (M * dx);
      if (pixel_scale_out) pixel_scale_out[b0 + i] = d_out;
      c.q2[i] = D / (2.0 * lb);
      c.q1[i] = -center_in / lb - sg * de / (M * d_out);
      c.q0[i] = sg * de * center_out / (M * d_out) - sg / 8.0;
      c.x0[i] = center_out - (m - 1) / 2.0 * d_out;
      c.dx[i] = d_out;
    }
    DLX_HIP(launch_lct_chirp(c, nb, m, (const float2*)out + (long)b0 * mm,
                             (float2*)out + (long)b0 * mm, s));
  }
  return DLX_OK;
}

int dlx_pupil_field(int32_t n, int32_t batch, const double* wavelength, double pixel_scale,
                    double center, double tilt_x, double tilt_y, const float* amp,
                    const float* opd, const float* phase, double scale, const void* in,
                    void* out, void* stream) {
  clear_error();
  DLX_CHECK_ARG(n > 0, "n must be > 0");
  DLX_CHECK_ARG(batch >= 0, "batch must be >= 0");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(wavelength && out, "null argument to dlx_pupil_field");
  DLX_CHECK_ARG(std::isfinite(pixel_scale) && std::isfinite(center) && std::isfinite(tilt_x) &&
                    std::isfinite(tilt_y) && std::isfinite(scale),
                "non-finite scalar argument to dlx_pupil_field");
  hipStream_t s = (hipStream_t)stream;
  const long npix = (long)n * n;
  const double x0 = center - (n - 1) / 2.0 * pixel_scale;
  for (int b0 = 0; b0 < batch; b0 += PupilChunk::kMax) {
    const int nb = batch - b0 < PupilChunk::kMax ? batch - b0 : PupilChunk::kMax;
    PupilChunk c;
    for (int i = 0; i < nb; ++i) {
      DLX_CHECK_ARG(std::isfinite(wavelength[b0 + i]) && wavelength[b0 + i] > 0,
                    "wavelength[%d] must be > 0", b0 + i);
      c.inv_wl[i] = 1.0 / wavelength[b0 + i];
    }
    DLX_HIP(launch_pupil_chunk(c, nb, n, pixel_scale, x0, tilt_x, tilt_y, amp, opd, phase,
                               (float)scale, in ? (const float2*)in + b0 * npix : nullptr,
                               (float2*)out + b0 * npix, s));
  }
  return DLX_OK;
}

int dlx_eval_basis(int32_t nbasis, int64_t npix, int32_t nsets, const float* basis,
                   const float* coeffs, const float* base, float* out, void* stream) {
  clear_error();
  DLX_CHECK_ARG(nbasis >= 0 && npix >= 0 && nsets >= 0, "negative size to dlx_eval_basis");
  if (npix == 0 || nsets == 0) return DLX_OK;
  DLX_CHECK_ARG(out && (nbasis == 0 || (basis && coeffs)), "null argument to dlx_eval_basis");
  DLX_HIP(launch_eval_basis(nbasis, npix, nsets, basis, coeffs, base, out,
                            (hipStream_t)stream));
  return DLX_OK;
}

int dlx_power(int64_t npix, int32_t batch, const void* field, float* out, void* stream) {
  clear_error();
  DLX_CHECK_ARG(npix >= 0 && batch >= 0, "negative size to dlx_power");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(field && out, "null argument to dlx_power");
  DLX_HIP(launch_power(npix, batch, (const float2*)field, out, (hipStream_t)stream));
  return DLX_OK;
}

int dlx_intensity(int64_t npix, int32_t batch, const void* field, int32_t sum_over_batch,
                  float* out, void* stream) {
  clear_error();
  DLX_CHECK_ARG(npix >= 0 && batch >= 0, "negative size to dlx_intensity");
  if (batch == 0 || npix == 0) return DLX_OK;
  DLX_CHECK_ARG(field && out, "null argument to dlx_intensity");
  DLX_HIP(launch_intensity(npix, batch, (const float2*)field, sum_over_batch, out,
                           (hipStream_t)stream));
  return DLX_OK;
}

}  // extern "C"

namespace dlx {
namespace {

// Bytes of the psf chain's buffers for a chunk of c wavelengths: the pupils P,
// the transfer matrices (used when they are not taken from the cache), the
// stage-1 intermediate U^T and the per-wavelength |E|^2 planes.
size_t psf_poly_bytes(const dlx_mft_geom& g, int c, bool split) {
  const size_t P = align_up((size_t)c * mat_bytes(g.n_in, g.n_in, split));
  const size_t W = align_up((size_t)c * mat_bytes(g.n_out, g.n_in, split));
  const size_t I = align_up((size_t)c * g.n_out * g.n_out * sizeof(float));
  return P + 2 * W + I;
}

int check_psf_poly_args(const dlx_mft_geom* geom, int32_t nwl, const double* wavelength,
                        const double* weight, double pixel_scale_in, double pixel_scale_out,
                        double center, double tilt_x, double tilt_y, double scale, float* psf,
                        int32_t precision) {
  if (int st = check_geom(geom)) return st;
  DLX_CHECK_ARG(nwl > 0, "nwl must be > 0, got %d", nwl);
  DLX_CHECK_ARG(wavelength && weight && psf, "null argument to dlx_psf_poly");
  DLX_CHECK_ARG(precision == DLX_PREC_FP32 || precision == DLX_PREC_BF16X3,
                "unknown precision %d", precision);
  DLX_CHECK_ARG(std::isfinite(pixel_scale_in) && pixel_scale_in != 0.0 &&
                    std::isfinite(pixel_scale_out) && std::isfinite(center) &&
                    std::isfinite(tilt_x) && std::isfinite(tilt_y) && std::isfinite(scale),
                "non-finite scalar argument to dlx_psf_poly");
  for (int l = 0; l < nwl; ++l) {
    DLX_CHECK_ARG(std::isfinite(weight[l]) && weight[l] >= 0, "weight[%d] must be >= 0", l);
    DLX_CHECK_ARG(std::isfinite(wavelength[l]) && wavelength[l] > 0,
                  "wavelength[%d] must be > 0", l);
  }
  return DLX_OK;
}

// The psf chain over chunks of at most c wavelengths in `ws` (psf_poly_bytes(c)
// bytes): pupils, matrices (from the device cache when `use_cache`, else
// generated into ws), the two contractions with the |E|^2 epilogue, and the
// running sum over wavelengths in wavelength order (so the result does not
// depend on c).  With `planes` set, every wavelength's w_l |E_l|^2 plane is
// written there instead (planes + l N_out^2) and nothing is summed.  `prof`
// records the contractions' live timing.
int psf_poly_run(const dlx_mft_geom& g, int32_t nwl, const double* wavelength,
                 const double* weight, double pixel_scale_in, double pixel_scale_out,
                 double center, double tilt_x, double tilt_y, const float* amp,
                 const float* opd, const float* phase, double scale, float* psf, bool split,
                 uint8_t* ws, int c, bool use_cache, bool prof, hipStream_t s,
                 float* planes = nullptr) {
  const size_t Pb = align_up((size_t)c * mat_bytes(g.n_in, g.n_in, split));
  const size_t Wb = align_up((size_t)c * mat_bytes(g.n_out, g.n_in, split));
  uint8_t* P = ws;
  uint8_t* W_ws = P + Pb;
  uint8_t* Ut = P + Pb + Wb;
  float* I = (float*)(P + Pb + 2 * Wb);
  const long nout2 = (long)g.n_out * g.n_out;
  const double x0 = center - (g.n_in - 1) / 2.0 * pixel_scale_in;

  std::vector<double> dx(c, pixel_scale_in), pout(c, pixel_scale_out), alpha(c);
  for (int l0 = 0; l0 < nwl; l0 += c) {
    const int nb = nwl - l0 < c ? nwl - l0 : c;
    float* Ic = planes ? planes + (long)l0 * nout2 : I;
    // |E_l|^2 = (c_l sqrt(w_l))^2 |Y^T P X|^2: the scale is applied in the
    // stage-2 epilogue, so the matrices depend on the geometry only.
    for (int i = 0; i < nb; ++i)
      alpha[i] = mft_norm(g, wavelength[l0 + i], pixel_scale_in, pixel_scale_out) *
                 std::sqrt(weight[l0 + i]);
    void* W = nullptr;
    CacheLease h(s);
    if (use_cache) {
      if (int st = get_matrices(g, nb, wavelength + l0, dx.data(), pout.data(), nullptr, split,
                                W_ws, s, &W, &h.h))
        return st;
    } else {
      W = W_ws;
      if (int st = gen_dft(g, nb, wavelength + l0, dx.data(), pout.data(), nullptr, nullptr,
                           W_ws, split, s))
        return st;
    }
    if (split) {
      // the pupil generated in stage 1's loader from the amplitude and OPD
      // planes where the shape allows (no pupil array in HBM), else written
      // by the pupil kernel
      PupilGen pg{};
      pg.opd = opd;
      pg.amp = amp;
      pg.dx = pixel_scale_in;
      pg.x0 = x0;
      pg.tx = tilt_x;
      pg.ty = tilt_y;
      pg.scale = (float)scale;
      pg.n = g.n_in;
      for (int i = 0; i < nb; ++i) pg.inv_wl[i] = 1.0 / wavelength[l0 + i];
      const bool fuse = !phase && pupil_prologue_enabled() &&
                        cgemm_pupil_supported(stage1_split(g, 0, nb, nullptr,
                                                           (const uint8_t*)W, Ut), pg);
      if (!fuse) {
        PupilChunk pc;
        for (int i = 0; i < nb; ++i) pc.inv_wl[i] = pg.inv_wl[i];
        DLX_HIP(launch_pupil_split4(pc, nb, g.n_in, pixel_scale_in, x0, tilt_x, tilt_y, amp,
                                    opd, phase, (float)scale, P,
                                    (long)mat_bytes(g.n_in, g.n_in, true), s));
      }
      if (int st = mft_split4(g, nb, P, (const uint8_t*)W, (const uint8_t*)W, Ut, alpha.data(),
                              Ic, EPI_INTENSITY, s, prof ? prof_begin : nullptr,
                              prof ? prof_end : nullptr, fuse ? &pg : nullptr))
        return st;
    } else {
      if (int st = dlx_pupil_field(g.n_in, nb, wavelength + l0, pixel_scale_in, center, tilt_x,
                                   tilt_y, amp, opd, phase, scale, nullptr, P, s))
        return st;
      if (int st = mft_fp32(g, nb, (const float2*)P, (const float2*)W, (const float2*)W,
                            (float2*)Ut, alpha.data(), Ic, EPI_INTENSITY, s, prof))
        return st;
    }
    if (int st = h.release()) return st;
    if (!planes) DLX_HIP(launch_sum_planes(nout2, nb, I, l0 > 0, psf, s));
  }
  return DLX_OK;
}

}  // namespace
}  // namespace dlx

extern "C" {

size_t dlx_psf_poly_workspace_size(const dlx_mft_geom* geom, int32_t nwl, int32_t precision) {
  if (!geom || nwl <= 0) return 0;
  return psf_poly_bytes(*geom, nwl < kWlChunk ? nwl : kWlChunk, precision == DLX_PREC_BF16X3);
}

int dlx_psf_poly(const dlx_mft_geom* geom, int32_t nwl, const double* wavelength,
                 const double* weight, double pixel_scale_in, double pixel_scale_out,
                 double center, double tilt_x, double tilt_y, const float* amp,
                 const float* opd, const float* phase, double scale, float* psf,
                 int32_t precision, void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  if (int st = check_psf_poly_args(geom, nwl, wavelength, weight, pixel_scale_in,
                                   pixel_scale_out, center, tilt_x, tilt_y, scale, psf,
                                   precision))
    return st;
  const size_t need = dlx_psf_poly_workspace_size(geom, nwl, precision);
  if (!workspace || workspace_bytes < need) {
    set_error("dlx_psf_poly workspace too small: need %zu bytes, got %zu", need,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  return psf_poly_run(*geom, nwl, wavelength, weight, pixel_scale_in, pixel_scale_out, center,
                      tilt_x, tilt_y, amp, opd, phase, scale, psf,
                      precision == DLX_PREC_BF16X3, (uint8_t*)workspace,
                      nwl < kWlChunk ? nwl : kWlChunk, true, true, (hipStream_t)stream);
}

size_t dlx_psf_poly_detector_workspace_size(const dlx_mft_geom* geom, int32_t nwl,
                                            int32_t precision, const dlx_detector_op* ops,
                                            int32_t nops) {
  if (!geom || nwl <= 0) return 0;
  const int c = nwl < kWlChunk ? nwl : kWlChunk;
  const size_t planes = align_up((size_t)nwl * geom->n_out * geom->n_out * sizeof(float));
  return psf_poly_bytes(*geom, c, precision == DLX_PREC_BF16X3) + planes +
         detector_ws_bytes(geom->n_out, 1, ops, nops, nwl);
}

int dlx_psf_poly_detector(const dlx_mft_geom* geom, int32_t nwl, const double* wavelength,
                          const double* weight, double pixel_scale_in, double pixel_scale_out,
                          double center, double tilt_x, double tilt_y, const float* amp,
                          const float* opd, const float* phase, double scale,
                          const dlx_detector_op* ops, int32_t nops, float* out,
                          int32_t precision, void* workspace, size_t workspace_bytes,
                          void* stream) {
  clear_error();
  if (int st = check_psf_poly_args(geom, nwl, wavelength, weight, pixel_scale_in,
                                   pixel_scale_out, center, tilt_x, tilt_y, scale, out,
                                   precision))
    return st;
  int32_t side = 0;
  if (int st = dlx_detector_out_side(geom->n_out, ops, nops, &side)) return st;
  const size_t need = dlx_psf_poly_detector_workspace_size(geom, nwl, precision, ops, nops);
  if (!workspace || workspace_bytes < need) {
    set_error("dlx_psf_poly_detector workspace too small: need %zu bytes, got %zu", need,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  const bool split = precision == DLX_PREC_BF16X3;
  const int c = nwl < kWlChunk ? nwl : kWlChunk;
  uint8_t* ws = (uint8_t*)workspace;
  const size_t chain = psf_poly_bytes(*geom, c, split);
  float* planes = (float*)(ws + chain);
  const size_t pbytes = align_up((size_t)nwl * geom->n_out * geom->n_out * sizeof(float));
  hipStream_t s = (hipStream_t)stream;
  // every wavelength's w_l |E_l|^2 plane, then the detector chain whose first
  // pass loads the sum over the planes (in wavelength order: the PSF that
  // dlx_psf_poly returns, never written)
  if (int st = psf_poly_run(*geom, nwl, wavelength, weight, pixel_scale_in, pixel_scale_out,
                            center, tilt_x, tilt_y, amp, opd, phase, scale, nullptr, split, ws,
                            c, true, true, s, planes))
    return st;
  return detector_apply_sum(geom->n_out, 1, nwl, (long)geom->n_out * geom->n_out, planes, ops,
                            nops, out, ws + chain + pbytes, workspace_bytes - chain - pbytes, s);
}

size_t dlx_psf_poly_vjp_workspace_size(const dlx_mft_geom* geom, int32_t nwl,
                                       int32_t precision) {
  if (!geom || nwl <= 0) return 0;
  const int c = nwl < kWlChunk ? nwl : kWlChunk;
  return vjp_layout(*geom, c, precision == DLX_PREC_BF16X3).total;
}

int dlx_psf_poly_vjp(const dlx_mft_geom* geom, int32_t nwl, const double* wavelength,
                     const double* weight, double pixel_scale_in, double pixel_scale_out,
                     double center, double tilt_x, double tilt_y, const float* amp,
                     const float* opd, const float* phase, double scale, const float* dpsf,
                     float* grad_opd, float* grad_phase, float* grad_amp, int32_t precision,
                     void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  if (int st = check_geom(geom)) return st;
  DLX_CHECK_ARG(nwl > 0, "nwl must be > 0, got %d", nwl);
  DLX_CHECK_ARG(wavelength && weight && dpsf, "null argument to dlx_psf_poly_vjp");
  DLX_CHECK_ARG(grad_opd || grad_phase || grad_amp,
                "dlx_psf_poly_vjp needs grad_opd, grad_phase or grad_amp");
  DLX_CHECK_ARG(precision == DLX_PREC_FP32 || precision == DLX_PREC_BF16X3,
                "unknown precision %d", precision);
  DLX_CHECK_ARG(std::isfinite(pixel_scale_in) && pixel_scale_in != 0.0 &&
                    std::isfinite(pixel_scale_out) && std::isfinite(center) &&
                    std::isfinite(tilt_x) && std::isfinite(tilt_y) && std::isfinite(scale),
                "non-finite scalar argument to dlx_psf_poly_vjp");
  const size_t need = dlx_psf_poly_vjp_workspace_size(geom, nwl, precision);
  if (!workspace || workspace_bytes < need) {
    set_error("dlx_psf_poly_vjp workspace too small: need %zu bytes, got %zu", need,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  for (int l = 0; l < nwl; ++l) {
    DLX_CHECK_ARG(std::isfinite(weight[l]) && weight[l] >= 0, "weight[%d] must be >= 0", l);
    DLX_CHECK_ARG(std::isfinite(wavelength[l]) && wavelength[l] > 0,
                  "wavelength[%d] must be > 0", l);
  }
  const dlx_mft_geom& g = *geom;
  hipStream_t s = (hipStream_t)stream;
  const bool split = precision == DLX_PREC_BF16X3;
  const int c = nwl < kWlChunk ? nwl : kWlChunk;
  const VjpLayout L = vjp_layout(g, c, split);
  uint8_t* P = (uint8_t*)workspace;
  uint8_t* Wf_ws = P + L.P;
  uint8_t* Wa_ws = Wf_ws + L.Wf;
  uint8_t* T = Wa_ws + L.Wa;
  uint8_t* Eb = T + L.T;
  float* H = (float*)(Eb + L.Eb);
  float* H2 = grad_amp ? (float*)((uint8_t*)H + L.H) : nullptr;
  uint8_t* U = (uint8_t*)H + L.H + L.H2;
  const long nin = g.n_in, nout = g.n_out;
  const double x0 = center - (g.n_in - 1) / 2.0 * pixel_scale_in;

  std::vector<double> dx(c, pixel_scale_in), pout(c, pixel_scale_out);
  for (int l0 = 0; l0 < nwl; l0 += c) {
    const int nb = nwl - l0 < c ? nwl - l0 : c;
    float alpha[kWlChunk];
    GradChunk gc;
    for (int i = 0; i < nb; ++i) {
      alpha[i] = (float)(mft_norm(g, wavelength[l0 + i], pixel_scale_in, pixel_scale_out) *
                         std::sqrt(weight[l0 + i]));
      gc.kw[i] = (float)(2.0 * M_PI / wavelength[l0 + i]);
    }
    void *Wf = nullptr, *Wa = nullptr;
    CacheLease hf(s), ha(s);
    if (int st = get_matrices(g, nb, wavelength + l0, dx.data(), pout.data(), nullptr, split,
                              Wf_ws, s, &Wf, &hf.h))
      return st;
    if (int st = get_matrices(g, nb, wavelength + l0, dx.data(), pout.data(), nullptr, split,
                              Wa_ws, s, &Wa, &ha.h, true))
      return st;
    if (split) {
      const long rb_in = split4_row_bytes(g.n_in), rb_out = split4_row_bytes(g.n_out);
      const long pb = (long)mat_bytes(nin, nin, true), wfb = (long)mat_bytes(nout, nin, true);
      const long wab = (long)mat_bytes(nin, nout, true), eb = (long)mat_bytes(nout, nout, true);
      PupilChunk pc;
      for (int i = 0; i < nb; ++i) pc.inv_wl[i] = 1.0 / wavelength[l0 + i];
      DLX_HIP(launch_pupil_split4(pc, nb, g.n_in, pixel_scale_in, x0, tilt_x, tilt_y, amp, opd,
                                  phase, (float)scale, P, pb, s));
      if (grad_amp)
        DLX_HIP(launch_pupil_split4(pc, nb, g.n_in, pixel_scale_in, x0, tilt_x, tilt_y, nullptr,
                                    opd, phase, (float)scale, U, pb, s));
      // forward stage 1: U^T = (P W^T)^T, split4 rows n_out
      GemmSplit g1{};
      g1.A = P; g1.a_row_bytes = rb_in; g1.a_batch_bytes = pb;
      g1.B = (const uint8_t*)Wf; g1.b_row_bytes = rb_in; g1.b_batch_bytes = wfb;
      g1.C = T; g1.ldc = rb_in; g1.c_batch = wfb;
      g1.M = g.n_in; g1.N = g.n_out; g1.K = g.n_in; g1.batch = nb; g1.epilogue = EPI_SPLIT_T;
      for (int i = 0; i < nb; ++i) g1.alpha[i] = 1.0f;
      DLX_HIP(launch_cgemm_split4(g1, s));
      // forward stage 2 in transposed order, E^T = U^T W^T, so that the split4
      // epilogue writes rows of Ebar = 2 G E (rows n_out, k = n_out)
      GemmSplit g2{};
      g2.A = T; g2.a_row_bytes = rb_in; g2.a_batch_bytes = wfb;
      g2.B = (const uint8_t*)Wf; g2.b_row_bytes = rb_in; g2.b_batch_bytes = wfb;
      g2.C = Eb; g2.ldc = rb_out; g2.c_batch = eb;
      g2.M = g.n_out; g2.N = g.n_out; g2.K = g.n_in; g2.batch = nb;
      g2.epilogue = EPI_VJP_SPLIT_T;
      for (int i = 0; i < nb; ++i) g2.alpha[i] = alpha[i];
      g2.G = dpsf; g2.ldg = g.n_out; g2.g_batch = 0;
      DLX_HIP(launch_cgemm_split4(g2, s));
      // adjoint stage 1: V = Ebar conj(W), stored V^T (rows n_in, k = n_out)
      GemmSplit g3{};
      g3.A = Eb; g3.a_row_bytes = rb_out; g3.a_batch_bytes = eb;
      g3.B = (const uint8_t*)Wa; g3.b_row_bytes = rb_out; g3.b_batch_bytes = wab;
      g3.C = T; g3.ldc = rb_out; g3.c_batch = wab;
      g3.M = g.n_out; g3.N = g.n_in; g3.K = g.n_out; g3.batch = nb; g3.epilogue = EPI_SPLIT_T;
      for (int i = 0; i < nb; ++i) g3.alpha[i] = 1.0f;
      DLX_HIP(launch_cgemm_split4(g3, s));
      // adjoint stage 2: Pbar = alpha W^H V, epilogue Im(conj(P) Pbar)
      GemmSplit g4{};
      g4.A = (const uint8_t*)Wa; g4.a_row_bytes = rb_out; g4.a_batch_bytes = wab;
      g4.B = T; g4.b_row_bytes = rb_out; g4.b_batch_bytes = wab;
      g4.C = H; g4.ldc = nin; g4.c_batch = nin * nin;
      g4.M = g.n_in; g4.N = g.n_in; g4.K = g.n_out; g4.batch = nb; g4.epilogue = EPI_GRAD;
      for (int i = 0; i < nb; ++i) g4.alpha[i] = alpha[i];
      g4.P = P; g4.p_ld = rb_in; g4.p_batch = pb;
      g4.C2 = H2; g4.U = U;
      DLX_HIP(launch_cgemm_split4(g4, s));
    } else {
      const long pst = nin * nin, wst = nout * nin, est = nout * nout;
      if (int st = dlx_pupil_field(g.n_in, nb, wavelength + l0, pixel_scale_in, center, tilt_x,
                                   tilt_y, amp, opd, phase, scale, nullptr, P, stream))
        return st;
      if (grad_amp)
        if (int st = dlx_pupil_field(g.n_in, nb, wavelength + l0, pixel_scale_in, center,
                                     tilt_x, tilt_y, nullptr, opd, phase, scale, nullptr, U,
                                     stream))
          return st;
      GemmBatch g1{};
      g1.A = (const float2*)P; g1.lda = nin; g1.sA = pst;
      g1.B = (const float2*)Wf; g1.ldb = nin; g1.sB = wst;
      g1.C = T; g1.ldc = nin; g1.sC = wst;
      g1.M = g.n_in; g1.N = g.n_out; g1.K = g.n_in; g1.batch = nb;
      g1.epilogue = EPI_STORE_T; g1.precision = DLX_PREC_FP32;
      for (int i = 0; i < nb; ++i) g1.alpha[i] = 1.0f;
      DLX_HIP(launch_cgemm_nt(g1, s));
      GemmBatch g2{};
      g2.A = (const float2*)Wf; g2.lda = nin; g2.sA = wst;
      g2.B = (const float2*)T; g2.ldb = nin; g2.sB = wst;
      g2.C = Eb; g2.ldc = nout; g2.sC = est;
      g2.M = g.n_out; g2.N = g.n_out; g2.K = g.n_in; g2.batch = nb;
      g2.epilogue = EPI_VJP; g2.precision = DLX_PREC_FP32;
      for (int i = 0; i < nb; ++i) g2.alpha[i] = alpha[i];
      g2.G = dpsf; g2.ldg = nout; g2.g_batch = 0;
      DLX_HIP(launch_cgemm_nt(g2, s));
      GemmBatch g3{};
      g3.A = (const float2*)Eb; g3.lda = nout; g3.sA = est;
      g3.B = (const float2*)Wa; g3.ldb = nout; g3.sB = wst;
      g3.C = T; g3.ldc = nout; g3.sC = wst;
      g3.M = g.n_out; g3.N = g.n_in; g3.K = g.n_out; g3.batch = nb;
      g3.epilogue = EPI_STORE_T; g3.precision = DLX_PREC_FP32;
      for (int i = 0; i < nb; ++i) g3.alpha[i] = 1.0f;
      DLX_HIP(launch_cgemm_nt(g3, s));
      GemmBatch g4{};
      g4.A = (const float2*)Wa; g4.lda = nout; g4.sA = wst;
      g4.B = (const float2*)T; g4.ldb = nout; g4.sB = wst;
      g4.C = H; g4.ldc = nin; g4.sC = pst;
      g4.M = g.n_in; g4.N = g.n_in; g4.K = g.n_out; g4.batch = nb;
      g4.epilogue = EPI_GRAD; g4.precision = DLX_PREC_FP32;
      for (int i = 0; i < nb; ++i) g4.alpha[i] = alpha[i];
      g4.P = P; g4.p_ld = nin; g4.p_batch = pst;
      g4.C2 = H2; g4.U = U;
      DLX_HIP(launch_cgemm_nt(g4, s));
    }
    if (int st = hf.release()) return st;
    if (int st = ha.release()) return st;
    if (grad_opd || grad_phase)
      DLX_HIP(launch_grad_reduce(gc, nb, nin * nin, H, l0 > 0, grad_opd, grad_phase, s));
    if (grad_amp)   // sum over wavelengths of Re(conj(U) Pbar) = dL/dA
      DLX_HIP(launch_grad_reduce(gc, nb, nin * nin, H2, l0 > 0, nullptr, grad_amp, s));
  }
  return DLX_OK;
}

size_t dlx_project_basis_workspace_size(int32_t nbasis, int64_t npix, int32_t nsets) {
  if (nbasis <= 0 || npix <= 0 || nsets <= 0) return 0;
  return align_up((size_t)nbasis * nsets * project_parts(npix) * sizeof(float));
}

int dlx_project_basis(int32_t nbasis, int64_t npix, int32_t nsets, const float* basis,
                      const float* grad, float* out, void* workspace, size_t workspace_bytes,
                      void* stream) {
  clear_error();
  DLX_CHECK_ARG(nbasis >= 0 && npix >= 0 && nsets >= 0, "negative size to dlx_project_basis");
  DLX_CHECK_ARG(nbasis <= 65535 && nsets <= 65535, "nbasis and nsets must be <= 65535");
  if (nbasis == 0 || nsets == 0) return DLX_OK;
  DLX_CHECK_ARG(basis && grad && out, "null argument to dlx_project_basis");
  hipStream_t s = (hipStream_t)stream;
  const size_t need = dlx_project_basis_workspace_size(nbasis, npix, nsets);
  if (need && (!workspace || workspace_bytes < need)) {
    set_error("dlx_project_basis workspace too small: need %zu bytes, got %zu", need,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  DLX_HIP(launch_project_basis(nbasis, npix, nsets, basis, grad, (float*)workspace, out, s));
  return DLX_OK;
}

int dlx_profile_enable(int32_t enable) {
  clear_error();
  g_prof.enabled = enable ? 1 : 0;
  g_prof.n = 0;
  return DLX_OK;
}

int dlx_profile_read(int32_t stage, double* total_ms, int32_t* count) {
  clear_error();
  DLX_CHECK_ARG(total_ms && count, "null argument to dlx_profile_read");
  double t = 0.0;
  int c = 0;
  for (int i = 0; i < g_prof.n; ++i) {
    if (g_prof.slots[i].stage != stage) continue;
    DLX_HIP(hipEventSynchronize(g_prof.slots[i].end));
    float ms = 0.f;
    DLX_HIP(hipEventElapsedTime(&ms, g_prof.slots[i].beg, g_prof.slots[i].end));
    t += ms;
    ++c;
  }
  *total_ms = t;
  *count = c;
  return DLX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// dlx_psf_poly_noalloc: the psf chain with every buffer in one fixed device
// array of the library, for callers that cannot allocate (code under graph
// capture, serving loops, harnesses that forbid hipMalloc).  The array is a
// file-scope __device__ symbol: each GPU the library runs on holds its own
// copy (allocated when the code object loads there, kept until the process
// ends).  Its device address is looked up once per device with
// hipGetSymbolAddress and only that address is handed to kernels; the array's
// name in host code is a host-side stand-in, not a device pointer.
// ---------------------------------------------------------------------------
namespace dlx {

constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;   // 64 MiB per GPU
static __device__ __align__(256) uint8_t scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices] = {};

// The current device's copy of `scratch` (looked up on its first use there).
int scratch_address(void** out) {
  *out = nullptr;
  int dev = 0;
  DLX_HIP(hipGetDevice(&dev));
  DLX_CHECK_ARG(dev >= 0 && dev < kMaxScratchDevices, "device %d beyond the %d supported", dev,
                kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      set_error("hipGetSymbolAddress of the scratch array failed on device %d: %s", dev,
                hipGetErrorString(e));
      return DLX_ERR_HIP;
    }
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return DLX_OK;
}

}  // namespace
}  // namespace dlx

extern "C" {

size_t dlx_psf_poly_noalloc_scratch_bytes(void) { return kNoallocScratchBytes; }

int dlx_psf_poly_noalloc(const dlx_mft_geom* geom, int32_t nwl, const double* wavelength,
                         const double* weight, double pixel_scale_in, double pixel_scale_out,
                         double center, double tilt_x, double tilt_y, const float* amp,
                         const float* opd, const float* phase, double scale, float* psf,
                         int32_t precision, void* stream) {
  clear_error();
  if (int st = check_psf_poly_args(geom, nwl, wavelength, weight, pixel_scale_in,
                                   pixel_scale_out, center, tilt_x, tilt_y, scale, psf,
                                   precision))
    return st;
  const bool split = precision == DLX_PREC_BF16X3;
  const size_t one = psf_poly_bytes(*geom, 1, split);
  if (one > kNoallocScratchBytes) {
    set_error("dlx_psf_poly_noalloc: one wavelength of a %d^2 -> %d^2 PSF needs %zu bytes, "
              "more than the %zu-byte scratch array; use dlx_psf_poly with a workspace",
              geom->n_in, geom->n_out, one, kNoallocScratchBytes);
    return DLX_ERR_INVALID;
  }
  // the largest chunk of wavelengths whose buffers fit the array
  int c = nwl < kWlChunk ? nwl : kWlChunk;
  while (c > 1 && psf_poly_bytes(*geom, c, split) > kNoallocScratchBytes) --c;
  void* ws = nullptr;
  if (int st = scratch_address(&ws)) return st;
  // matrices are generated into the array every call: the device cache allocates
  return psf_poly_run(*geom, nwl, wavelength, weight, pixel_scale_in, pixel_scale_out, center,
                      tilt_x, tilt_y, amp, opd, phase, scale, psf, split, (uint8_t*)ws, c,
                      false, false, (hipStream_t)stream);
}

}  // extern "C"
