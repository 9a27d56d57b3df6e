// Detector chain on the PSF (gfx950): the reference's detector layers
// (layers/detector_layers.py:100-293, applied in order by LayeredDetector,
// detectors.py:103-128) as a few fused passes over float32 images.
//
// A chain is planned into passes.  A pass is either a 2-D correlation over an
// LDS tile (convolve, or its adjoint) or a block sum (downsample; a factor of
// 1 is a plain elementwise pass).  Elementwise layers -- pixel response
// multiply, saturation min, constant add -- never get a pass of their own:
// those before the first structural op are applied as each input element is
// loaded, the rest as each output element of the preceding pass is stored.
//
//   convolve  out[y][x] = sum_{ky,kx} in[y + o - ky][x + o - kx] w[ky][kx],
//             o = (K-1)/2, zero outside (jax.scipy.signal.convolve, mode
//             "same", as PSF.convolve calls it, psfs.py:93-110); run as a
//             correlation with the flipped kernel starting at row y + o-(K-1).
//   downsample out[Y][X] = sum of the k x k block (utils/array_ops.py:124-162).
//
// The vector-Jacobian product re-runs the chain keeping each pass's input and
// its pre-store values, then walks the passes backwards: the store-side ops'
// derivatives, then the adjoint of the pass (correlation with the unflipped
// kernel from row y - o, or block replication) times the load-side ops'
// derivatives.  All passes stream the images once; the kernels are HBM-bound.
#include "dlx_internal.h"

namespace dlx {
namespace {

constexpr int kMaxEw = 8;        // elementwise ops fused into one load or store
constexpr int kMaxKernel = 64;   // convolution kernel side limit (LDS tile)
constexpr int kTileX = 64;       // correlation tile: 64 x 16 outputs,
constexpr int kTileY = 16;       // 4 consecutive columns per thread
constexpr int kThreads = 256;

struct EwList {
  int n;
  int kind[kMaxEw];
  float value[kMaxEw];
  const float* map[kMaxEw];
};

// v after the ops of l; p indexes the pixel within its image (maps are per image).
__device__ __forceinline__ float ew_apply(const EwList& l, float v, long p) {
  for (int i = 0; i < l.n; ++i) {
    const int k = l.kind[i];
    if (k == DLX_DET_PIXEL_RESPONSE) {
      v *= l.map[i][p];
    } else if (k == DLX_DET_SATURATE) {
      v = v > l.value[i] ? l.value[i] : v;   // keeps NaN, as jnp.minimum
    } else {
      v += l.value[i];
    }
  }
  return v;
}

// d(ops of l)(v)/dv.  jnp.minimum splits its derivative 1/2, 1/2 at a tie.
__device__ __forceinline__ float ew_deriv(const EwList& l, float v, long p) {
  float d = 1.f;
  for (int i = 0; i < l.n; ++i) {
    const int k = l.kind[i];
    if (k == DLX_DET_PIXEL_RESPONSE) {
      const float m = l.map[i][p];
      d *= m;
      v *= m;
    } else if (k == DLX_DET_SATURATE) {
      const float t = l.value[i];
      d *= v < t ? 1.f : (v == t ? 0.5f : 0.f);
      v = v > t ? t : v;
    } else {
      v += l.value[i];
    }
  }
  return d;
}

// out[y][x] = post( sum_{ky,kx} pre(src[y + lo + ky][x + lo + kx]) * w'[ky][kx] )
//             [* dpre'(aux[y][x])]
// with w' = w flipped in both axes when flip != 0, and zero outside the image.
// One workgroup per 64 x 16 tile and image (blockIdx.y = image).
__global__ void __launch_bounds__(kThreads)
det_corr_kernel(const float* __restrict__ src, int n, int tiles_x, int K, int lo,
                const float* __restrict__ w, int flip, EwList pre, EwList post,
                const float* __restrict__ aux, EwList dpre, float* __restrict__ r_save,
                float* __restrict__ out) {
  extern __shared__ float lds[];
  const int TW = kTileX + K - 1, TH = kTileY + K - 1;
  float* wl = lds;           // K*K weights
  float* tl = lds + K * K;   // TH x TW input tile
  const long img = (long)n * n;
  const long base = (long)blockIdx.y * img;
  const int ty0 = (blockIdx.x / tiles_x) * kTileY;
  const int tx0 = (blockIdx.x % tiles_x) * kTileX;
  const float* s = src + base;
  for (int i = threadIdx.x; i < K * K; i += kThreads) {
    const int ky = i / K, kx = i - ky * K;
    wl[i] = flip ? w[(K - 1 - ky) * K + (K - 1 - kx)] : w[i];
  }
  for (int i = threadIdx.x; i < TH * TW; i += kThreads) {
    const int yy = i / TW, xx = i - yy * TW;
    const int gy = ty0 + lo + yy, gx = tx0 + lo + xx;
    float v = 0.f;
    if (gy >= 0 && gy < n && gx >= 0 && gx < n) {
      const long p = (long)gy * n + gx;
      v = ew_apply(pre, s[p], p);
    }
    tl[i] = v;
  }
  __syncthreads();
  const int row = threadIdx.x / (kTileX / 4);
  const int c0 = (threadIdx.x % (kTileX / 4)) * 4;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int ky = 0; ky < K; ++ky) {
    const float* trow = tl + (row + ky) * TW + c0;   // columns c0 .. c0+K+2 < TW
    const float* wrow = wl + ky * K;
    float a0 = trow[0], a1 = trow[1], a2 = trow[2];
    for (int kx = 0; kx < K; ++kx) {
      const float a3 = trow[kx + 3];
      const float wv = wrow[kx];
      acc[0] += a0 * wv;
      acc[1] += a1 * wv;
      acc[2] += a2 * wv;
      acc[3] += a3 * wv;
      a0 = a1;
      a1 = a2;
      a2 = a3;
    }
  }
  const int y = ty0 + row;
  if (y >= n) return;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int x = tx0 + c0 + j;
    if (x >= n) break;
    const long p = (long)y * n + x;
    const float r = acc[j];
    if (r_save) r_save[base + p] = r;
    float v = ew_apply(post, r, p);
    if (aux) v *= ew_deriv(dpre, aux[base + p], p);
    out[base + p] = v;
  }
}

// out[b][Y][X] = post( sum_{dy,dx < k} pre(src[b][Y k + dy][X k + dx]) ), m = n / k.
// Row sums first, then the sum of the rows (the order of utils/array_ops.py:151-161).
__global__ void __launch_bounds__(256)
det_down_kernel(const float* __restrict__ src, int n, int k, int m, long total, EwList pre,
                EwList post, float* __restrict__ r_save, float* __restrict__ out) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const long mm = (long)m * m;
  const long b = i / mm;
  const long q = i - b * mm;
  const int oy = (int)(q / m), ox = (int)(q - (long)oy * m);
  const float* s = src + b * (long)n * n;
  float acc = 0.f;
  for (int dy = 0; dy < k; ++dy) {
    const long rp = (long)(oy * k + dy) * n + (long)ox * k;
    float rs = 0.f;
    for (int dx = 0; dx < k; ++dx) rs += ew_apply(pre, s[rp + dx], rp + dx);
    acc += rs;
  }
  if (r_save) r_save[i] = acc;
  out[i] = ew_apply(post, acc, q);
}

// Adjoint of the block sum: out[b][y][x] = g[b][y/k][x/k] * dpre'(aux[b][y][x]).
__global__ void __launch_bounds__(256)
det_up_kernel(const float* __restrict__ g, int n, int k, int m, long total,
              const float* __restrict__ aux, EwList dpre, float* __restrict__ out) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const long nn = (long)n * n;
  const long b = i / nn;
  const long p = i - b * nn;
  const int y = (int)(p / n), x = (int)(p - (long)y * n);
  float v = g[b * (long)m * m + (long)(y / k) * m + x / k];
  out[i] = v * ew_deriv(dpre, aux[i], p);
}

// out[i] = g[i] * post'(r[i]), elementwise (in place allowed).
__global__ void __launch_bounds__(256)
det_ew_grad_kernel(const float* g, const float* __restrict__ r, long npix, long total,
                   EwList post, float* out) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  out[i] = g[i] * ew_deriv(post, r[i], i % npix);
}

struct Pass {
  int conv;        // 1: convolve, 0: block sum (k = 1: elementwise only)
  int n_in, n_out;
  int K;           // convolve kernel side
  const float* w;  // convolve kernel (device)
  int k;           // block-sum factor
  EwList pre, post;
};

bool ew_push(EwList& l, const dlx_detector_op& op) {
  if (l.n == kMaxEw) return false;
  l.kind[l.n] = op.kind;
  l.value[l.n] = (float)op.value;
  l.map[l.n] = op.data;
  ++l.n;
  return true;
}

Pass make_pass(int conv, int side) {
  Pass p{};
  p.conv = conv;
  p.n_in = p.n_out = side;
  p.k = 1;
  return p;
}

int plan(int n, const dlx_detector_op* ops, int nops, std::vector<Pass>& passes) {
  DLX_CHECK_ARG(n > 0, "detector image side must be positive, got %d", n);
  DLX_CHECK_ARG(nops >= 0 && (nops == 0 || ops), "bad detector op list");
  passes.clear();
  EwList pending{};
  int cur = n;
  for (int i = 0; i < nops; ++i) {
    const dlx_detector_op& op = ops[i];
    switch (op.kind) {
      case DLX_DET_PIXEL_RESPONSE:
        DLX_CHECK_ARG(op.data, "pixel response %d has no map", i);
        DLX_CHECK_ARG(op.size == cur,
                      "pixel response %d is %dx%d but the image is %dx%d at that point", i,
                      op.size, op.size, cur, cur);
        // fall through
      case DLX_DET_SATURATE:
      case DLX_DET_ADD_CONSTANT: {
        EwList& target = passes.empty() ? pending : passes.back().post;
        if (!ew_push(target, op)) {
          Pass p = make_pass(0, cur);
          if (passes.empty()) {
            p.pre = pending;
            pending = EwList{};
          }
          passes.push_back(p);
          ew_push(passes.back().post, op);
        }
        break;
      }
      case DLX_DET_CONVOLVE: {
        DLX_CHECK_ARG(op.data, "convolution %d has no kernel", i);
        DLX_CHECK_ARG(op.size >= 1 && op.size <= kMaxKernel,
                      "convolution kernel side must be in [1, %d], got %d", kMaxKernel, op.size);
        Pass p = make_pass(1, cur);
        p.K = op.size;
        p.w = op.data;
        if (passes.empty()) {
          p.pre = pending;
          pending = EwList{};
        }
        passes.push_back(p);
        break;
      }
      case DLX_DET_DOWNSAMPLE: {
        DLX_CHECK_ARG(op.size >= 1, "downsample factor must be positive, got %d", op.size);
        DLX_CHECK_ARG(cur % op.size == 0,
                      "Input array has %d pixels, which is not divisible by %d", cur, op.size);
        Pass p = make_pass(0, cur);
        p.k = op.size;
        p.n_out = cur / op.size;
        if (passes.empty()) {
          p.pre = pending;
          pending = EwList{};
        }
        passes.push_back(p);
        cur = p.n_out;
        break;
      }
      default:
        DLX_CHECK_ARG(false, "unknown detector op kind %d at %d", op.kind, i);
    }
  }
  if (passes.empty()) {
    Pass p = make_pass(0, n);
    p.pre = pending;
    passes.push_back(p);
  }
  return DLX_OK;
}

inline size_t img_bytes(int side, int batch) {
  return align_up((size_t)side * side * (size_t)batch * sizeof(float));
}

struct Layout {
  size_t total = 0;
  std::vector<size_t> x_off;   // saved pass inputs (for_vjp; index 0 unused)
  std::vector<size_t> r_off;   // pre-store values (for_vjp)
  size_t g_off[2] = {0, 0};    // ping-pong buffers
};

Layout layout(const std::vector<Pass>& passes, int batch, bool for_vjp) {
  Layout L;
  const size_t np = passes.size();
  int side_max = 0;
  for (const Pass& p : passes) side_max = std::max(side_max, std::max(p.n_in, p.n_out));
  if (for_vjp) {
    L.x_off.assign(np, 0);
    L.r_off.assign(np, 0);
    for (size_t i = 1; i < np; ++i) {
      L.x_off[i] = L.total;
      L.total += img_bytes(passes[i].n_in, batch);
    }
    for (size_t i = 0; i < np; ++i) {
      L.r_off[i] = L.total;
      L.total += img_bytes(passes[i].n_out, batch);
    }
    for (int j = 0; j < 2; ++j) {
      L.g_off[j] = L.total;
      L.total += img_bytes(side_max, batch);
    }
  } else if (np > 1) {
    int inter = 0;
    for (size_t i = 0; i + 1 < np; ++i) inter = std::max(inter, passes[i].n_out);
    for (int j = 0; j < 2; ++j) {
      L.g_off[j] = L.total;
      L.total += img_bytes(inter, batch);
    }
  }
  return L;
}

int check_grid(const std::vector<Pass>& passes, int batch) {
  DLX_CHECK_ARG(batch <= 65535, "detector batch must be <= 65535, got %d", batch);
  for (const Pass& p : passes) {
    const long total = (long)std::max(p.n_in, p.n_out) * std::max(p.n_in, p.n_out) * batch;
    DLX_CHECK_ARG(total / 256 < (1L << 31) - 1, "detector images too large for one launch");
  }
  return DLX_OK;
}

// Forward pass i: src -> dst (r_save may be NULL).
int run_pass(const Pass& p, int batch, const float* src, float* r_save, float* dst,
             hipStream_t st) {
  if (p.conv) {
    const int tiles_x = (p.n_in + kTileX - 1) / kTileX;
    const int tiles_y = (p.n_in + kTileY - 1) / kTileY;
    const int lo = (p.K - 1) / 2 - (p.K - 1);
    const size_t lds = sizeof(float) * ((size_t)p.K * p.K +
                                        (size_t)(kTileY + p.K - 1) * (kTileX + p.K - 1));
    hipLaunchKernelGGL(det_corr_kernel, dim3(tiles_x * tiles_y, batch), dim3(kThreads), lds,
                       st, src, p.n_in, tiles_x, p.K, lo, p.w, 1, p.pre, p.post,
                       (const float*)nullptr, EwList{}, r_save, dst);
    return launch_status("det_corr_kernel");
  }
  const long total = (long)p.n_out * p.n_out * batch;
  hipLaunchKernelGGL(det_down_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     src, p.n_in, p.k, p.n_out, total, p.pre, p.post, r_save, dst);
  return launch_status("det_down_kernel");
}

}  // namespace
}  // namespace dlx

using namespace dlx;

extern "C" {

int dlx_detector_out_side(int32_t n, const dlx_detector_op* ops, int32_t nops,
                          int32_t* out_side) {
  clear_error();
  std::vector<Pass> passes;
  const int st = plan(n, ops, nops, passes);
  if (st) return st;
  if (out_side) *out_side = passes.back().n_out;
  return DLX_OK;
}

size_t dlx_detector_workspace_size(int32_t n, int32_t batch, const dlx_detector_op* ops,
                                   int32_t nops, int32_t for_vjp) {
  std::vector<Pass> passes;
  if (batch < 0 || plan(n, ops, nops, passes)) return 0;
  return layout(passes, batch, for_vjp != 0).total;
}

int dlx_detector_apply(int32_t n, int32_t batch, const float* in, const dlx_detector_op* ops,
                       int32_t nops, float* out, void* workspace, size_t workspace_bytes,
                       void* stream) {
  clear_error();
  std::vector<Pass> passes;
  int st = plan(n, ops, nops, passes);
  if (st) return st;
  DLX_CHECK_ARG(batch >= 0, "negative detector batch");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(in && out, "null image pointer to dlx_detector_apply");
  if ((st = check_grid(passes, batch))) return st;
  const Layout L = layout(passes, batch, false);
  if (workspace_bytes < L.total || (L.total && !workspace)) {
    set_error("dlx_detector_apply workspace too small: need %zu bytes, got %zu", L.total,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  const float* src = in;
  for (size_t i = 0; i < passes.size(); ++i) {
    float* dst = i + 1 == passes.size() ? out : (float*)(ws + L.g_off[i & 1]);
    if ((st = run_pass(passes[i], batch, src, nullptr, dst, s))) return st;
    src = dst;
  }
  return DLX_OK;
}

int dlx_detector_vjp(int32_t n, int32_t batch, const float* in, const dlx_detector_op* ops,
                     int32_t nops, const float* dout, float* din, void* workspace,
                     size_t workspace_bytes, void* stream) {
  clear_error();
  std::vector<Pass> passes;
  int st = plan(n, ops, nops, passes);
  if (st) return st;
  DLX_CHECK_ARG(batch >= 0, "negative detector batch");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(in && dout && din, "null image pointer to dlx_detector_vjp");
  if ((st = check_grid(passes, batch))) return st;
  const Layout L = layout(passes, batch, true);
  if (workspace_bytes < L.total || !workspace) {
    set_error("dlx_detector_vjp workspace too small: need %zu bytes, got %zu", L.total,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  const size_t np = passes.size();
  float* g[2] = {(float*)(ws + L.g_off[0]), (float*)(ws + L.g_off[1])};
  auto x_in = [&](size_t i) -> const float* {
    return i == 0 ? in : (const float*)(ws + L.x_off[i]);
  };
  // forward, keeping every pass input and pre-store value
  for (size_t i = 0; i < np; ++i) {
    float* dst = i + 1 < np ? (float*)(ws + L.x_off[i + 1]) : g[1];
    if ((st = run_pass(passes[i], batch, x_in(i), (float*)(ws + L.r_off[i]), dst, s)))
      return st;
  }
  // backward
  const float* cur = dout;
  int cur_idx = -1;   // -1: cur is the caller's dout
  for (size_t ii = np; ii-- > 0;) {
    const Pass& p = passes[ii];
    const long out_total = (long)p.n_out * p.n_out * batch;
    const float* gr = cur;
    int gr_idx = cur_idx;
    if (p.post.n > 0) {
      gr_idx = cur_idx < 0 ? 0 : cur_idx;
      hipLaunchKernelGGL(det_ew_grad_kernel, dim3((unsigned)((out_total + 255) / 256)),
                         dim3(256), 0, s, cur, (const float*)(ws + L.r_off[ii]),
                         (long)p.n_out * p.n_out, out_total, p.post, g[gr_idx]);
      if ((st = launch_status("det_ew_grad_kernel"))) return st;
      gr = g[gr_idx];
    }
    const int dst_idx = gr_idx == 0 ? 1 : 0;
    float* dst = ii == 0 ? din : g[dst_idx];
    if (p.conv) {
      const int tiles_x = (p.n_in + kTileX - 1) / kTileX;
      const int tiles_y = (p.n_in + kTileY - 1) / kTileY;
      const int lo = -((p.K - 1) / 2);
      const size_t lds = sizeof(float) * ((size_t)p.K * p.K +
                                          (size_t)(kTileY + p.K - 1) * (kTileX + p.K - 1));
      hipLaunchKernelGGL(det_corr_kernel, dim3(tiles_x * tiles_y, batch), dim3(kThreads), lds,
                         s, gr, p.n_in, tiles_x, p.K, lo, p.w, 0, EwList{}, EwList{}, x_in(ii),
                         p.pre, (float*)nullptr, dst);
      if ((st = launch_status("det_corr_kernel"))) return st;
    } else {
      const long in_total = (long)p.n_in * p.n_in * batch;
      hipLaunchKernelGGL(det_up_kernel, dim3((unsigned)((in_total + 255) / 256)), dim3(256), 0,
                         s, gr, p.n_in, p.k, p.n_out, in_total, x_in(ii), p.pre, dst);
      if ((st = launch_status("det_up_kernel"))) return st;
    }
    cur = dst;
    cur_idx = dst_idx;
  }
  return DLX_OK;
}

}  // extern "C"
