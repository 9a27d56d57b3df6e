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
//   A separable kernel w[ky][kx] = wy[ky] wx[kx] (ApplyJitter's Gaussian) runs
//   as a horizontal then a vertical pass through LDS: 2K instead of K^2
//   multiply-adds per pixel, which keeps the pass HBM-bound.
//   A convolve followed by a downsample runs as one pass (det_conv_down_kernel).
//
// The chain's first pass can read the sum of several images instead of one
// (SumIn): the wavelength planes of a polychromatic PSF, summed in plane order
// as each input element is loaded, so the |E|^2 planes go straight into the
// detector without the PSF being written and read back.
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

// Input of a pass: one image, or (n > 1) the sum of n images `stride` floats
// apart, added in order (plane 0 + plane 1 + ...).
struct SumIn {
  int n;
  long stride;
};

// s[idx], or the in-order sum over the planes of si.
__device__ __forceinline__ float ld_sum(const float* __restrict__ s, long idx, const SumIn& si) {
  float v = s[idx];
#pragma unroll 8
  for (int q = 1; q < si.n; ++q) v += s[idx + (long)q * si.stride];
  return v;
}

struct EwList {
  int n;
  int kind[kMaxEw];
  float value[kMaxEw];
  const float* map[kMaxEw];
  int op[kMaxEw];   // index of the op in the caller's list
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

// Whether l multiplies by a pixel-response map.
__device__ __forceinline__ bool has_map(const EwList& l) {
  for (int i = 0; i < l.n; ++i)
    if (l.kind[i] == DLX_DET_PIXEL_RESPONSE) return true;
  return false;
}

// Index of the first pixel-response op of l (-1 if none).
__device__ __forceinline__ int first_map(const EwList& l) {
  for (int i = 0; i < l.n; ++i)
    if (l.kind[i] == DLX_DET_PIXEL_RESPONSE) return i;
  return -1;
}

// ew_apply with the map value of op `pi` already loaded (mv).
__device__ __forceinline__ float ew_apply_pf(const EwList& l, float v, long p, int pi,
                                             float mv) {
  for (int i = 0; i < l.n; ++i) {
    const int k = l.kind[i];
    if (k == DLX_DET_PIXEL_RESPONSE) {
      v *= i == pi ? mv : l.map[i][p];
    } else if (k == DLX_DET_SATURATE) {
      v = v > l.value[i] ? l.value[i] : v;
    } else {
      v += l.value[i];
    }
  }
  return v;
}

// ew_apply on N values at once, for a list with at most one pixel-response
// op, whose map values are mv (unused where the list has none): no memory
// reads but the op list's own, once per op for all N values, so a caller
// streaming its inputs ahead is not made to wait for them here.  Each value
// sees the ops in list order, as in ew_apply.
template <int N>
__device__ __forceinline__ void ew_apply_mv(const EwList& l, float (&v)[N],
                                            const float (&mv)[N]) {
  for (int i = 0; i < l.n; ++i) {
    const int k = l.kind[i];
    const float c = l.value[i];
    if (k == DLX_DET_PIXEL_RESPONSE) {
#pragma unroll
      for (int j = 0; j < N; ++j) v[j] *= mv[j];
    } else if (k == DLX_DET_SATURATE) {
#pragma unroll
      for (int j = 0; j < N; ++j) v[j] = v[j] > c ? c : v[j];
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) v[j] += c;
    }
  }
}

// The four map values of op `pi` at (y, x0..x0+3), clamped into the n x n map;
// one 16-byte load when the four are in the image and aligned.
__device__ __forceinline__ void load_map4(const EwList& l, int pi, int n, int y, int x0,
                                          float (&mv)[4]) {
  if (pi < 0) {
    mv[0] = mv[1] = mv[2] = mv[3] = 1.f;
    return;
  }
  const float* m = l.map[pi];
  y = min(y, n - 1);
  const long row = (long)y * n;
  if (n % 4 == 0 && x0 + 3 < n && ((uintptr_t)m & 15) == 0) {
    const float4 v = *(const float4*)(m + row + x0);
    mv[0] = v.x;
    mv[1] = v.y;
    mv[2] = v.z;
    mv[3] = v.w;
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) mv[j] = m[row + min(x0 + j, n - 1)];
  }
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

// Tile kernels: 256 threads, 64 output columns (4 consecutive per thread)
// and 16 * RPT output rows (RPT rows per thread, 16 apart).
// LDS floats: weights, the (16 RPT + K - 1) x (64 + K - 1) input tile and, for
// separable weights, the (16 RPT + K - 1) x 64 horizontal-pass rows, which
// start at a 16-byte boundary (float4 accesses).
// Tile row stride: at least 64 + K - 1 columns, a multiple of 4, and wide
// enough for every thread's float4 window of K + 3 columns.
__host__ __device__ inline int conv_tile_w(int K) { return kTileX - 4 + 4 * ((K + 6) / 4); }

// Offset of the input tile: after the weights, at a 16-byte boundary.
__host__ __device__ inline int conv_tile_offset(int K, int sep) {
  return ((sep ? 2 * K : K * K) + 3) / 4 * 4;
}

__host__ __device__ inline int conv_hrow_offset(int K, int rpt) {
  return conv_tile_offset(K, 1) + (kTileY * rpt + K - 1) * conv_tile_w(K);
}

__host__ __device__ inline int conv_lds_floats(int K, int sep, int rpt) {
  const int TH = kTileY * rpt + K - 1;
  return sep ? conv_hrow_offset(K, rpt) + TH * kTileX
             : conv_tile_offset(K, 0) + TH * conv_tile_w(K);
}

// Rows per thread of the tile kernels: 2 (32-row tiles, less halo per output)
// when the tile fits the 64 KiB of LDS a workgroup may allocate, else 1.
inline int conv_rpt(int K, int sep) {
  return conv_lds_floats(K, sep, 2) * sizeof(float) <= 65536 ? 2 : 1;
}

// Workgroups are dealt to the 8 XCDs round-robin (hardware id % 8).  Remap
// the (tile, image) pairs so each XCD takes a contiguous run of them.  With
// the tile index fastest (map_major == 0), the halo rows that vertically
// adjacent tiles of an image share stay in one L2.  When the pass reads a
// pixel-response map (map_major != 0) the image index goes fastest instead:
// one tile position is done for every image in a row, so its slice of the map
// is fetched from HBM once per launch rather than once per image.
__device__ __forceinline__ void xcd_tile(int map_major, int& tile, int& image) {
  const int nx = gridDim.x, nb = gridDim.y, total = nx * nb;
  int id = blockIdx.y * nx + blockIdx.x;
  if (total % 8 == 0) id = (id % 8) * (total / 8) + id / 8;
  if (map_major) {
    image = id % nb;
    tile = id / nb;
  } else {
    tile = id % nx;
    image = id / nx;
  }
}

// Kernels of up to kWinK taps read each thread's K + 3 tile values as float4s
// into registers (one 16-byte LDS read per four values) and run the taps
// fully unrolled.
constexpr int kWinK = 13;
constexpr int kWin = kWinK + 3;

__device__ __forceinline__ void load_window(const float* row, int K, float (&a)[kWin]) {
  const float4* r4 = (const float4*)row;   // 16-byte aligned: c0 and the row stride are x4
#pragma unroll
  for (int q = 0; q < kWin / 4; ++q) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (4 * q < K + 3) v = r4[q];          // stays inside the row stride
    a[4 * q] = v.x;
    a[4 * q + 1] = v.y;
    a[4 * q + 2] = v.z;
    a[4 * q + 3] = v.w;
  }
}

// The 4 x RPT correlation outputs of this thread for the tile at (ty0, tx0):
//   acc[u][j] = sum_{ky,kx} pre(src[ty0 + row + 16u + lo + ky][tx0 + c0 + j + lo + kx])
//               * w'[ky][kx]
// (zero outside the image; w' = w flipped in both axes when flip != 0).  With
// sep != 0, w holds K column weights then K row weights (w[ky][kx] = wy[ky] wx[kx]):
// a horizontal pass over every tile row into LDS, then a vertical pass.
// Ends with the input tile no longer read by any thread of the workgroup.
template <int RPT>
__device__ __forceinline__ void conv_tile(const float* __restrict__ s, int n, int ty0, int tx0,
                                          int K, int lo, const float* __restrict__ w, int flip,
                                          int sep, const EwList& pre, const SumIn& si,
                                          float* lds, int row, int c0,
                                          float (&acc)[RPT][4]) {
  const int TW = conv_tile_w(K), TH = kTileY * RPT + K - 1;
  float* wl = lds;
  float* tl = lds + conv_tile_offset(K, sep);
  // Thread t loads tile columns t % 64 and t % 64 + 64 (when < TW) of rows
  // t / 64, t / 64 + 4, ...: twelve rows of loads in flight (every row of a
  // tile up to 48 rows high), addresses clamped into the image (every load
  // valid) and out-of-image elements zeroed after.  The first batch is issued
  // before the weights are staged, so the two latencies overlap.
  constexpr int kRowStep = kThreads / kTileX;   // 4
  constexpr int kRows = 12;
  const int cx = threadIdx.x % kTileX;
  const int r0 = threadIdx.x / kTileX;
  const int gx0 = tx0 + lo + cx, gx1 = gx0 + kTileX;
  const bool has1 = cx + kTileX < TW;
  const bool okx0 = gx0 >= 0 && gx0 < n;
  const bool okx1 = has1 && gx1 >= 0 && gx1 < n;
  const int cx0 = min(max(gx0, 0), n - 1), cx1 = min(max(gx1, 0), n - 1);
  for (int rb = r0; rb < TH; rb += kRows * kRowStep) {
    float v0[kRows], v1[kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      const int gy = min(max(ty0 + lo + rb + u * kRowStep, 0), n - 1);
      v0[u] = s[(long)gy * n + cx0];
      v1[u] = s[(long)gy * n + cx1];
    }
    // summed input: the other planes in order, a whole batch of rows per plane
    for (int q = 1; q < si.n; ++q) {
      const float* sq = s + (long)q * si.stride;
#pragma unroll
      for (int u = 0; u < kRows; ++u) {
        const int gy = min(max(ty0 + lo + rb + u * kRowStep, 0), n - 1);
        v0[u] += sq[(long)gy * n + cx0];
        v1[u] += sq[(long)gy * n + cx1];
      }
    }
    if (rb == r0) {
      if (sep) {
        for (int i = threadIdx.x; i < 2 * K; i += kThreads) {
          const int seg = i / K, j = i - seg * K;
          wl[i] = w[seg * K + (flip ? K - 1 - j : j)];
        }
      } else {
        for (int i = threadIdx.x; i < K * K; i += kThreads) {
          const int ky = i / K, kx = i - ky * K;
          wl[i] = flip ? w[(K - 1 - ky) * K + (K - 1 - kx)] : w[i];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      const int r = rb + u * kRowStep;
      if (r >= TH) break;
      const int gy = ty0 + lo + r;
      const bool oky = gy >= 0 && gy < n;
      const long p = (long)gy * n;
      tl[r * TW + cx] = oky && okx0 ? ew_apply(pre, v0[u], p + gx0) : 0.f;
      if (has1) tl[r * TW + cx + kTileX] = oky && okx1 ? ew_apply(pre, v1[u], p + gx1) : 0.f;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < RPT; ++u) acc[u][0] = acc[u][1] = acc[u][2] = acc[u][3] = 0.f;
  if (sep) {
    const float* wy = wl;
    const float* wx = wl + K;
    float* hl = lds + conv_hrow_offset(K, RPT);   // TH x 64
    if (K <= kWinK) {
      for (int r = row; r < TH; r += kTileY) {
        float a[kWin];
        load_window(tl + r * TW + c0, K, a);
        float h[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kx = 0; kx < kWinK; ++kx) {
          if (kx < K) {
            const float wv = wx[kx];
#pragma unroll
            for (int j = 0; j < 4; ++j) h[j] += a[kx + j] * wv;
          }
        }
        *(float4*)(hl + r * kTileX + c0) = make_float4(h[0], h[1], h[2], h[3]);
      }
    } else {
      for (int r = row; r < TH; r += kTileY) {
        const float* trow = tl + r * TW + c0;   // columns c0 .. c0+K+2 < TW
        float h0 = 0.f, h1 = 0.f, h2 = 0.f, h3 = 0.f;
        float a0 = trow[0], a1 = trow[1], a2 = trow[2];
#pragma unroll 4
        for (int kx = 0; kx < K; ++kx) {
          const float a3 = trow[kx + 3];
          const float wv = wx[kx];
          h0 += a0 * wv;
          h1 += a1 * wv;
          h2 += a2 * wv;
          h3 += a3 * wv;
          a0 = a1;
          a1 = a2;
          a2 = a3;
        }
        *(float4*)(hl + r * kTileX + c0) = make_float4(h0, h1, h2, h3);
      }
    }
    __syncthreads();
    for (int ky = 0; ky < K; ++ky) {
      const float wv = wy[ky];
#pragma unroll
      for (int u = 0; u < RPT; ++u) {
        const float4 h = *(const float4*)(hl + (row + kTileY * u + ky) * kTileX + c0);
        acc[u][0] += h.x * wv;
        acc[u][1] += h.y * wv;
        acc[u][2] += h.z * wv;
        acc[u][3] += h.w * wv;
      }
    }
  } else if (K <= kWinK) {
    for (int ky = 0; ky < K; ++ky) {
      const float* wrow = wl + ky * K;
#pragma unroll
      for (int u = 0; u < RPT; ++u) {
        float a[kWin];
        load_window(tl + (row + kTileY * u + ky) * TW + c0, K, a);
#pragma unroll
        for (int kx = 0; kx < kWinK; ++kx) {
          if (kx < K) {
            const float wv = wrow[kx];
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[u][j] += a[kx + j] * wv;
          }
        }
      }
    }
  } else {
    for (int ky = 0; ky < K; ++ky) {
      const float* wrow = wl + ky * K;
#pragma unroll
      for (int u = 0; u < RPT; ++u) {
        const float* trow = tl + (row + kTileY * u + ky) * TW + c0;   // cols < c0+K+3 <= TW
        float a0 = trow[0], a1 = trow[1], a2 = trow[2];
#pragma unroll 4
        for (int kx = 0; kx < K; ++kx) {
          const float a3 = trow[kx + 3];
          const float wv = wrow[kx];
          acc[u][0] += a0 * wv;
          acc[u][1] += a1 * wv;
          acc[u][2] += a2 * wv;
          acc[u][3] += a3 * wv;
          a0 = a1;
          a1 = a2;
          a2 = a3;
        }
      }
    }
  }
  __syncthreads();
}

// out[y][x] = post( correlation of pre(src) with w' at (y, x) )  [* dpre'(aux[y][x])]
// One workgroup per 64 x 16 RPT tile and image (blockIdx.y = image).
template <int RPT>
__global__ void __launch_bounds__(kThreads)
det_corr_kernel(const float* __restrict__ src, int n, int tiles_x, int K, int lo,
                const float* __restrict__ w, int flip, int sep, EwList pre, EwList post,
                const float* __restrict__ aux, EwList dpre, float* __restrict__ r_save,
                float* __restrict__ out, SumIn si) {
  extern __shared__ float lds[];
  int tile, image;
  xcd_tile(has_map(pre) || has_map(post) || has_map(dpre), tile, image);
  const long base = (long)image * n * n;
  const int ty0 = (tile / tiles_x) * (kTileY * RPT);
  const int tx0 = (tile % tiles_x) * kTileX;
  const int row = threadIdx.x / (kTileX / 4);
  const int c0 = (threadIdx.x % (kTileX / 4)) * 4;
  // the store-side response map is loaded before the tile, so its latency
  // overlaps the tile loads
  const int pi = first_map(post);
  float mv[RPT][4];
#pragma unroll
  for (int u = 0; u < RPT; ++u) load_map4(post, pi, n, ty0 + row + kTileY * u, tx0 + c0, mv[u]);
  float acc[RPT][4];
  conv_tile<RPT>(src + base, n, ty0, tx0, K, lo, w, flip, sep, pre, si, lds, row, c0, acc);
  // four consecutive outputs go out as one 16-byte store where they can
  const bool vec = n % 4 == 0 && (((uintptr_t)out | (uintptr_t)r_save) & 15) == 0;
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    const int y = ty0 + row + kTileY * u;
    if (y >= n) break;
    if (vec && tx0 + c0 + 3 < n) {
      const long p = (long)y * n + tx0 + c0;
      float4 v;
      float* vv = &v.x;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        vv[j] = ew_apply_pf(post, acc[u][j], p + j, pi, mv[u][j]);
        if (aux) vv[j] *= ew_deriv(dpre, aux[base + p + j], p + j);
      }
      if (r_save)
        *(float4*)(r_save + base + p) = make_float4(acc[u][0], acc[u][1], acc[u][2], acc[u][3]);
      *(float4*)(out + base + p) = v;
      continue;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int x = tx0 + c0 + j;
      if (x >= n) break;
      const long p = (long)y * n + x;
      const float r = acc[u][j];
      if (r_save) r_save[base + p] = r;
      float v = ew_apply_pf(post, r, p, pi, mv[u][j]);
      if (aux) v *= ew_deriv(dpre, aux[base + p], p);
      out[base + p] = v;
    }
  }
}

// Convolve followed by a block sum, in one pass: the convolution over a
// th x tw tile (multiples of the block factor k, th <= 16 RPT, tw <= 64) with
// the convolve pass's store-side ops, then the k x k block sums of that tile
// out of LDS with the block-sum pass's store-side ops.  The intermediate image
// is never written.  Blocks never straddle tiles (k divides th, tw and n).
template <int RPT>
__global__ void __launch_bounds__(kThreads)
det_conv_down_kernel(const float* __restrict__ src, int n, int th, int tw, int tiles_x, int K,
                     int lo, const float* __restrict__ w, int sep, EwList pre, EwList post,
                     int k, int m, EwList post2, float* __restrict__ out, SumIn si) {
  extern __shared__ float lds[];
  int tile, image;
  xcd_tile(has_map(pre) || has_map(post) || has_map(post2), tile, image);
  const long base = (long)image * n * n;
  const int ty0 = (tile / tiles_x) * th;
  const int tx0 = (tile % tiles_x) * tw;
  const int row = threadIdx.x / (kTileX / 4);
  const int c0 = (threadIdx.x % (kTileX / 4)) * 4;
  const int pi = first_map(post);   // prefetched as in det_corr_kernel
  float mv[RPT][4];
#pragma unroll
  for (int u = 0; u < RPT; ++u) load_map4(post, pi, n, ty0 + row + kTileY * u, tx0 + c0, mv[u]);
  float acc[RPT][4];
  conv_tile<RPT>(src + base, n, ty0, tx0, K, lo, w, 1, sep, pre, si, lds, row, c0, acc);
  float* ot = lds;   // (16 RPT) x 64 convolve outputs (the tile is no longer read)
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    const int rr = row + kTileY * u;
    const int y = ty0 + rr;
    float4 v;
    float* vv = &v.x;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int x = tx0 + c0 + j;
      vv[j] = 0.f;
      if (rr < th && c0 + j < tw && y < n && x < n)
        vv[j] = ew_apply_pf(post, acc[u][j], (long)y * n + x, pi, mv[u][j]);
    }
    *(float4*)(ot + rr * kTileX + c0) = v;
  }
  __syncthreads();
  const int bw = tw / k, bh = th / k;
  const long obase = (long)image * m * m;
  for (int o = threadIdx.x; o < bw * bh; o += kThreads) {
    const int oy = o / bw, ox = o - oy * bw;
    const int gy = ty0 / k + oy, gx = tx0 / k + ox;
    if (gy >= m || gx >= m) continue;
    float a = 0.f;
    for (int dy = 0; dy < k; ++dy) {
      const float* orow = ot + (oy * k + dy) * kTileX + ox * k;
      float rs = 0.f;
      if (k % 4 == 0) {   // 16-byte aligned groups of four
        for (int dx = 0; dx < k; dx += 4) {
          const float4 q = *(const float4*)(orow + dx);
          rs += ((q.x + q.y) + q.z) + q.w;
        }
      } else {
        for (int dx = 0; dx < k; ++dx) rs += orow[dx];
      }
      a += rs;
    }
    const long q = (long)gy * m + gx;
    out[obase + q] = ew_apply(post2, a, q);
  }
}

// ---------------------------------------------------------------------------
// Column-walking separable convolution (odd K <= 13), optionally followed by a
// KD x KD block sum (KD = 2 or 4; KD = 1: the convolution alone).  A workgroup
// is one wave; its 64 lanes own 4 adjacent columns each (a 256-column band of
// one image) and walk down a segment of rows with no barriers.  Input rows
// (the band plus its K - 1 halo columns) and the rows of the map that the ops
// after the convolution multiply by are streamed global -> LDS by the wave's
// own direct loads (buffer loads into LDS) into a ring of kColRing slots, kColRing - 1
// rows ahead, without holding registers.  Per row a lane reads its 4 + K - 1
// inputs from the slot and forms the 4 horizontal sums; each output row is the
// vertical sum over the last K of those.  Rows are processed in unrolled
// blocks of kColU, the horizontal sums of a block and of the K - 1 rows before
// it held in one register window that moves by kColU rows per block.  Block
// sums of KD rows accumulate in registers as the rows go by.  Pieces of rows
// or columns outside the image load as zeros (buffer range checks), so no
// lane masks anything.  Needs n % 4 == 0, 16-byte aligned images and map, and no
// op before the convolution.  The arithmetic, order included, is that of
// conv_tile with det_corr_kernel / det_conv_down_kernel.
// ---------------------------------------------------------------------------
constexpr int kColBand = 256;              // columns per wave
constexpr int kColRing = 4;                // LDS row slots per wave
constexpr int kColU = 8;                   // rows per unrolled block (a multiple of kColRing)
constexpr int kSlotIn = kColBand + 16;     // a slot's input row: the band + 16 halo columns
constexpr int kSlotMap = kSlotIn + kColBand;   // + the map row
static_assert(kColU % kColRing == 0, "slot indices are constant in a block");

// Word 3 of a raw buffer descriptor (32-bit data format, no swizzle, no
// stride): loads range-checked against the descriptor's size in bytes.
constexpr int kRawBuffer = 0x00020000;
constexpr int kColOutside = 0x40000000;   // a byte offset past any row

typedef __attribute__((address_space(3))) void col_lds_t;

template <int K>
struct ColPieces {
  static constexpr int LO = (K - 1) / 2 - (K - 1);       // first tap's offset
  static constexpr int LO4 = -((-LO + 3) / 4) * 4;       // LO rounded down to x4
  static constexpr int OFF = LO - LO4;                   // window start in the pieces
  static constexpr int NV = (OFF + 4 + K - 1 + 3) / 4;   // 16-byte pieces per row
};

// s_waitcnt immediate: at most n vector-memory operations outstanding
// (gfx9: vmcnt in bits 3:0 and 15:14; expcnt and lgkmcnt left unconstrained).
constexpr int vmcnt_imm(int n) { return (n & 15) | (7 << 4) | (15 << 8) | (((n >> 4) & 3) << 14); }

typedef __attribute__((address_space(3))) float lds_float;

typedef float lds_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 lds_ld4(const lds_float* p) {
  const lds_f4 v = *(const __attribute__((address_space(3))) lds_f4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}

// Per-wave state of det_sep_col_kernel: where its rows come from and go to.
template <int KD, int K, bool MAP>
struct ColWave {
  static constexpr int LO = ColPieces<K>::LO, OFF = ColPieces<K>::OFF, NV = ColPieces<K>::NV;
  static constexpr int NB = KD == 1 ? 1 : 4 / KD;
  static constexpr int kLoads = MAP ? 3 : 2;   // direct loads per row
  lds_float* ring[kColRing];
  const float* s;      // this image
  const float* mapp;   // the streamed map (MAP)
  float* out;          // this image's output
  int n, m, y0, total, lane, c0, vl, vh, vm;
  float wy[K], wx[K];
  const EwList* post;
  const EwList* post2;

  // H row i (input row y0 + LO + i) and the map row of the output row it
  // completes (y0 + i - (K - 1)) into a slot
  __device__ __forceinline__ void issue(int i, lds_float* slot) const {
    const int r = y0 + LO + i;
    const bool okr = r >= 0 && r < n;
    const __amdgpu_buffer_rsrc_t row = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(s + (long)(okr ? r : 0) * n), (short)0, okr ? 4 * n : 0, kRawBuffer);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(row, (col_lds_t*)slot, 16, vl, 0, 0, 0);
    if (lane < 4)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(row, (col_lds_t*)(slot + kColBand), 16, vh, 0, 0,
                                               0);
    if (MAP) {
      const int ry = min(max(y0 + i - (K - 1), 0), n - 1);
      const __amdgpu_buffer_rsrc_t mrow = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(mapp + (long)ry * n), (short)0, 4 * n, kRawBuffer);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(mrow, (col_lds_t*)(slot + kSlotIn), 16, vm, 0, 0,
                                               0);
    }
  }

  // One block of kColU rows from row i0.  CHECK: the block may run past the
  // last row or be one of the first K - 1 rows (checked per row); otherwise
  // every row of the block, and every row it issues, exists, and every row
  // completes an output row.  hw[t]: the horizontal sums of H row
  // i0 - (K - 1) + t.
  template <bool CHECK>
  __device__ __forceinline__ void block(int i0, float (&hw)[K - 1 + kColU][4],
                                        float (&blk)[NB]) const {
#pragma unroll
    for (int u = 0; u < kColU; ++u) {
      const int i = i0 + u;
      if (CHECK && i >= total) break;
      // row i's loads have landed once at most those of the kColRing - 2 rows
      // issued after it remain outstanding (loads complete in order; stores
      // issued since only make the wait stricter); near the end, fewer rows
      // were issued: wait for all
      if (!CHECK || i + kColRing - 2 < total)
        __builtin_amdgcn_s_waitcnt(vmcnt_imm(kLoads * (kColRing - 2)));
      else
        __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
      const lds_float* slot = ring[u % kColRing];
      float a[4 * NV], mv[4];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const float4 t = lds_ld4(slot + 4 * lane + 4 * v);
        a[4 * v] = t.x; a[4 * v + 1] = t.y; a[4 * v + 2] = t.z; a[4 * v + 3] = t.w;
      }
      if (MAP) {
        const float4 t = lds_ld4(slot + kSlotIn + 4 * lane);
        mv[0] = t.x; mv[1] = t.y; mv[2] = t.z; mv[3] = t.w;
      } else {
        mv[0] = mv[1] = mv[2] = mv[3] = 1.f;
      }
      // refill the slot row i - 1 used with row i + kColRing - 1
      if (!CHECK || i + kColRing - 1 < total)
        issue(i + kColRing - 1, ring[(u + kColRing - 1) % kColRing]);
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        float h = 0.f;
#pragma unroll
        for (int t = 0; t < K; ++t) h += a[OFF + jj + t] * wx[t];
        hw[K - 1 + u][jj] = h;
      }
      if (CHECK && i < K - 1) continue;
      const int Y = y0 + i - (K - 1);
      float o[4];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        float acc = 0.f;
#pragma unroll
        for (int t = 0; t < K; ++t) acc += hw[u + t][jj] * wy[t];
        o[jj] = acc;
      }
      ew_apply_mv(*post, o, mv);
      if (KD == 1) {
        if (c0 < n) *(float4*)(out + (long)Y * n + c0) = make_float4(o[0], o[1], o[2], o[3]);
        continue;
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (KD == 4) blk[b] += ((o[0] + o[1]) + o[2]) + o[3];
        else blk[b] += o[2 * b] + o[2 * b + 1];
      }
      if (Y % KD == KD - 1) {
        float one[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) one[b] = 1.f;
        ew_apply_mv(*post2, blk, one);
        if (c0 < n) {
          float* dst = out + (long)(Y / KD) * m + c0 / KD;
#pragma unroll
          for (int b = 0; b < NB; ++b) dst[b] = blk[b];
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) blk[b] = 0.f;
      }
    }
    // move the window to the next block
#pragma unroll
    for (int t = 0; t < K - 1; ++t)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) hw[t][jj] = hw[t + kColU][jj];
  }
};

// MAP: the ops after the convolution hold one pixel-response op, whose map
// rows are streamed with the input rows.
template <int KD, int K, bool MAP>
__global__ void __launch_bounds__(64)
det_sep_col_kernel(const float* __restrict__ src, int n, int bands, int seg_h,
                   const float* __restrict__ w, EwList post, int m, EwList post2,
                   float* __restrict__ out) {
  using W = ColWave<KD, K, MAP>;
  constexpr int LO4 = ColPieces<K>::LO4;
  static_assert(4 * 63 + 4 * W::NV <= kSlotIn, "window inside the slot");
  // one LDS object per slot, so every slot index below is a constant and the
  // compiler orders each slot's reads after that slot's own loads only
  constexpr int SL = MAP ? kSlotMap : kSlotIn;   // floats per slot
  __shared__ __attribute__((aligned(16))) float sl0[SL], sl1[SL], sl2[SL], sl3[SL];
  float* const sl[kColRing] = {sl0, sl1, sl2, sl3};
  static_assert(kColRing == 4, "one LDS object per slot");
  W wv;
#pragma unroll
  for (int q = 0; q < kColRing; ++q) wv.ring[q] = (lds_float*)sl[q];
  int tile, image;
  xcd_tile(has_map(post), tile, image);
  const int band = tile % bands, seg = tile / bands;
  wv.lane = threadIdx.x;
  const int cb = band * kColBand;          // the wave's first column
  wv.c0 = cb + 4 * wv.lane;                // this lane's
  wv.n = n;
  wv.m = m;
  wv.s = src + (long)image * n * n;
  wv.out = out + (long)image * (KD == 1 ? n : m) * (KD == 1 ? n : m);
  wv.y0 = seg * seg_h;
  wv.total = (min(wv.y0 + seg_h, n) - wv.y0) + K - 1;   // H rows y0 + LO + i, i < total
  // the convolution as a correlation with the reversed factors (w: K column
  // then K row weights)
#pragma unroll
  for (int t = 0; t < K; ++t) {
    wv.wy[t] = w[K - 1 - t];
    wv.wx[t] = w[2 * K - 1 - t];
  }
  // The loads are raw buffer loads with one descriptor per row, 16 bytes a
  // lane, written to the slot at 16 bytes x lane: lane l brings columns
  // cb + LO4 + 4 l, lanes 0..3 also the 16 past the band.  A piece is wholly
  // inside or wholly outside the image (n and every offset are multiples of
  // 4); a lane whose byte offset is at or past the descriptor's size reads
  // zeros, which is how pieces outside the image (columns, or a whole row)
  // come in.
  const int xl = cb + LO4 + 4 * wv.lane;
  const int xh = cb + LO4 + kColBand + 4 * (wv.lane & 3);
  wv.vl = xl >= 0 && xl < n ? 4 * xl : kColOutside;
  wv.vh = xh >= 0 && xh < n ? 4 * xh : kColOutside;
  wv.vm = 4 * min(wv.c0, n - 4);   // lanes past the image store nothing
  wv.mapp = MAP ? post.map[first_map(post)] : src;
  wv.post = &post;
  wv.post2 = &post2;
  float hw[K - 1 + kColU][4];
  float blk[W::NB];
#pragma unroll
  for (int b = 0; b < W::NB; ++b) blk[b] = 0.f;
#pragma unroll
  for (int u = 0; u + 1 < kColRing; ++u)
    if (u < wv.total) wv.issue(u, wv.ring[u]);
  // the first blocks hold the K - 1 rows that complete no output; blocks
  // whose rows and refills all exist run unchecked
  constexpr int kFirst = (K - 1 + kColU - 1) / kColU * kColU;
  int i0 = 0;
  for (; i0 < kFirst && i0 < wv.total; i0 += kColU) wv.template block<true>(i0, hw, blk);
  for (; i0 + kColU + kColRing - 1 <= wv.total; i0 += kColU) wv.template block<false>(i0, hw, blk);
  for (; i0 < wv.total; i0 += kColU) wv.template block<true>(i0, hw, blk);
}

// ---------------------------------------------------------------------------
// Adjoint of the column-walking chain (image cotangent only): for
//   out = post2( blocksum_KD( post( conv_K(x) ) ) )
// with post holding only pixel-response (at most one) and add ops and post2
// no map,
//   din = corr_K( d1 * up_KD( dout * post2'(b) ) )
// where b are the block sums entering post2 (saved by a forward recompute with
// post2 left out), post2' is ew_deriv of post2 at b, d1 = post's derivative
// (the map value, or 1), up_KD replicates a block's value over its KD x KD
// pixels and corr_K is the correlation with the unflipped factors (the
// convolution's adjoint; zero outside the image).  Same walk as
// det_sep_col_kernel: one wave per 256-column band and segment of rows, the
// rows the cotangent image h = d1 * up(...) needs (map row, dout and b block
// rows) streamed into a 4-slot LDS ring by buffer loads three rows ahead; each
// lane forms h on its own 16-byte piece of the band (the lanes 0..3 also on
// the 16 columns past it) into a wave-private LDS row, then reads its window
// of that row as the forward reads its input.  The products are those of
// det_ew_grad_kernel / det_up_kernel / det_corr_kernel (flip 0), in the same
// order, so the result equals the multi-pass adjoint to the bit.  Needs
// n % (4 KD) == 0 (block pieces wholly inside or outside the image).
// ---------------------------------------------------------------------------
template <int KD>
struct ColBlk {
  // block columns a slot holds, from a multiple of 4 at or below the first
  // block column of the band's window; lanes of 16 bytes that load them
  static constexpr int kCols = KD == 1 ? 272 : KD == 2 ? 144 : 72;
  static constexpr int kLanes = kCols / 4;
};

// d(ops of l)(v)/dv at N values, for a list with no map (each value as
// ew_deriv; the op list read once per op for all values).
template <int N>
__device__ __forceinline__ void ew_deriv_nomap(const EwList& l, float (&v)[N], float (&d)[N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) d[j] = 1.f;
  for (int i = 0; i < l.n; ++i) {
    const int k = l.kind[i];
    const float t = l.value[i];
    if (k == DLX_DET_SATURATE) {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        d[j] *= v[j] < t ? 1.f : (v[j] == t ? 0.5f : 0.f);
        v[j] = v[j] > t ? t : v[j];
      }
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) v[j] += t;
    }
  }
}

template <int KD, int K, bool MAP>
struct ColVjpWave {
  static constexpr int LO = ColPieces<K>::LO, LO4 = ColPieces<K>::LO4;
  static constexpr int OFF = ColPieces<K>::OFF, NV = ColPieces<K>::NV;
  static constexpr int BC = ColBlk<KD>::kCols, BL = ColBlk<KD>::kLanes;
  // slot layout (floats): [map row: kSlotIn] [dout block row: BC] [b block row: BC]
  static constexpr int kMapOff = 0;
  static constexpr int kGOff = MAP ? kSlotIn : 0;
  static constexpr int kBOff = kGOff + BC;
  static constexpr int kSlotF = KD == 1 ? kBOff : kBOff + BC;
  // direct loads per row: map (band + halo), dout block row (KD = 1: band +
  // halo), b block row (KD > 1)
  static constexpr int kLoads = (MAP ? 2 : 0) + (KD == 1 ? 2 : 2);
  lds_float* ring[kColRing];
  lds_float* hrow;          // the cotangent row h (kSlotIn floats)
  const float* g;           // dout of this image (m x m)
  const float* bs;          // saved block sums of this image (m x m; KD > 1)
  const float* mapp;        // the map (MAP)
  float* out;               // din of this image
  int n, m, y0, total, lane, c0, cb, bc0, vl, vh, vbl, vbh;
  float wy[K], wx[K];
  const EwList* post2;

  __device__ __forceinline__ void issue(int i, lds_float* slot) const {
    const int r = y0 + LO + i;
    const bool okr = r >= 0 && r < n;
    if (MAP) {
      const __amdgpu_buffer_rsrc_t row = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(mapp + (long)(okr ? r : 0) * n), (short)0, okr ? 4 * n : 0, kRawBuffer);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(row, (col_lds_t*)(slot + kMapOff), 16, vl, 0, 0,
                                               0);
      if (lane < 4)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(row, (col_lds_t*)(slot + kMapOff + kColBand),
                                                 16, vh, 0, 0, 0);
    }
    const int br = okr ? r / KD : 0;
    const __amdgpu_buffer_rsrc_t grow = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(g + (long)br * m), (short)0, okr ? 4 * m : 0, kRawBuffer);
    if (KD == 1) {
      __builtin_amdgcn_raw_ptr_buffer_load_lds(grow, (col_lds_t*)(slot + kGOff), 16, vbl, 0, 0,
                                               0);
      if (lane < 4)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(grow, (col_lds_t*)(slot + kGOff + kColBand), 16,
                                                 vbh, 0, 0, 0);
    } else {
      const __amdgpu_buffer_rsrc_t brow = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(bs + (long)br * m), (short)0, okr ? 4 * m : 0, kRawBuffer);
      if (lane < BL) {
        __builtin_amdgcn_raw_ptr_buffer_load_lds(grow, (col_lds_t*)(slot + kGOff), 16, vbl, 0,
                                                 0, 0);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(brow, (col_lds_t*)(slot + kBOff), 16, vbl, 0,
                                                 0, 0);
      }
    }
  }

  // The block values of piece q of the window (columns x0 .. x0 + 3, x0 =
  // cb + LO4 + 4 q): dout * post2'(b) of each column's block (KD = 4: one
  // block, KD = 2: two).  Block column floor(x / KD) - bc0 >= 0 for every
  // column of the window; the slot holds zeros for block columns outside the
  // image (and the map zeros for columns outside it), so h is 0 there.
  __device__ __forceinline__ void block_g(const lds_float* slot, int q, float (&gd)[2]) const {
    const int x0 = cb + LO4 + 4 * q;
    if (KD == 2) {
      const int bc = (x0 >> 1) - bc0;
      float b[2] = {slot[kBOff + bc], slot[kBOff + bc + 1]}, d[2];
      ew_deriv_nomap(*post2, b, d);
      gd[0] = slot[kGOff + bc] * d[0];
      gd[1] = slot[kGOff + bc + 1] * d[1];
    } else if (KD == 4) {
      const int bc = (x0 >> 2) - bc0;
      float b[1] = {slot[kBOff + bc]}, d[1];
      ew_deriv_nomap(*post2, b, d);
      gd[0] = gd[1] = slot[kGOff + bc] * d[0];
    }
  }

  // h on piece q from its block values (KD > 1) or the dout row (KD = 1),
  // times the map value where post holds one
  __device__ __forceinline__ float4 h_piece(const lds_float* slot, int q,
                                            const float (&gd)[2]) const {
    float v[4];
    if (KD == 1) {
      const float4 t = lds_ld4(slot + kGOff + 4 * q);
      v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
      v[0] = v[1] = gd[0];
      v[2] = v[3] = gd[1];
    }
    if (MAP) {
      const float4 mv = lds_ld4(slot + kMapOff + 4 * q);
      v[0] *= mv.x; v[1] *= mv.y; v[2] *= mv.z; v[3] *= mv.w;
    }
    return make_float4(v[0], v[1], v[2], v[3]);
  }

  // gm, gh: the block values of this lane's piece and of its halo piece,
  // kept while the rows stay in one block row (a new block row starts where
  // y0 + LO + i is a multiple of KD: y0 and i0 are multiples of KD, so at fixed
  // u in every block).
  template <bool CHECK>
  __device__ __forceinline__ void block(int i0, float (&hw)[K - 1 + kColU][4], float (&gm)[2],
                                        float (&gh)[2]) const {
#pragma unroll
    for (int u = 0; u < kColU; ++u) {
      const int i = i0 + u;
      if (CHECK && i >= total) break;
      if (!CHECK || i + kColRing - 2 < total)
        __builtin_amdgcn_s_waitcnt(vmcnt_imm(kLoads * (kColRing - 2)));
      else
        __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
      const lds_float* slot = ring[u % kColRing];
      const bool new_block = ((LO % KD) + KD + u) % KD == 0;   // constant once unrolled
      if (KD > 1 && (new_block || (CHECK && i == 0))) {
        block_g(slot, lane, gm);
        block_g(slot, 64 + (lane & 3), gh);
      }
      // h on this lane's piece (and the halo lanes' second piece) into the row
      const float4 hp = h_piece(slot, lane, gm);
      const float4 hh = h_piece(slot, 64 + (lane & 3), gh);   // stored by lanes 0..3
      if (!CHECK || i + kColRing - 1 < total)
        issue(i + kColRing - 1, ring[(u + kColRing - 1) % kColRing]);
      *(__attribute__((address_space(3))) lds_f4*)(hrow + 4 * lane) =
          lds_f4{hp.x, hp.y, hp.z, hp.w};
      if (lane < 4)
        *(__attribute__((address_space(3))) lds_f4*)(hrow + kColBand + 4 * lane) =
            lds_f4{hh.x, hh.y, hh.z, hh.w};
      // (a wave's LDS accesses complete in order: the window reads below see
      // the writes above, and the next row's writes follow these reads)
      float a[4 * NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const float4 t = lds_ld4(hrow + 4 * lane + 4 * v);
        a[4 * v] = t.x; a[4 * v + 1] = t.y; a[4 * v + 2] = t.z; a[4 * v + 3] = t.w;
      }
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        float h = 0.f;
#pragma unroll
        for (int t = 0; t < K; ++t) h += a[OFF + jj + t] * wx[t];
        hw[K - 1 + u][jj] = h;
      }
      if (CHECK && i < K - 1) continue;
      const int Y = y0 + i - (K - 1);
      float o[4];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        float acc = 0.f;
#pragma unroll
        for (int t = 0; t < K; ++t) acc += hw[u + t][jj] * wy[t];
        o[jj] = acc;
      }
      if (c0 < n) *(float4*)(out + (long)Y * n + c0) = make_float4(o[0], o[1], o[2], o[3]);
    }
#pragma unroll
    for (int t = 0; t < K - 1; ++t)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) hw[t][jj] = hw[t + kColU][jj];
  }
};

template <int KD, int K, bool MAP>
__global__ void __launch_bounds__(64)
det_sep_col_vjp_kernel(const float* __restrict__ dout, const float* __restrict__ bsave, int n,
                       int bands, int seg_h, const float* __restrict__ w, EwList post, int m,
                       EwList post2, float* __restrict__ din) {
  using W = ColVjpWave<KD, K, MAP>;
  constexpr int LO4 = ColPieces<K>::LO4;
  constexpr int SL = W::kSlotF;
  static_assert(4 * 63 + 4 * W::NV <= kSlotIn, "window inside the row");
  __shared__ __attribute__((aligned(16))) float sl0[SL], sl1[SL], sl2[SL], sl3[SL];
  __shared__ __attribute__((aligned(16))) float hr[kSlotIn];
  float* const sl[kColRing] = {sl0, sl1, sl2, sl3};
  static_assert(kColRing == 4, "one LDS object per slot");
  W wv;
#pragma unroll
  for (int q = 0; q < kColRing; ++q) wv.ring[q] = (lds_float*)sl[q];
  wv.hrow = (lds_float*)hr;
  int tile, image;
  xcd_tile(has_map(post), tile, image);
  const int band = tile % bands, seg = tile / bands;
  wv.lane = threadIdx.x;
  wv.cb = band * kColBand;
  wv.c0 = wv.cb + 4 * wv.lane;
  wv.n = n;
  wv.m = m;
  wv.g = dout + (long)image * m * m;
  wv.bs = bsave + (long)image * m * m;
  wv.out = din + (long)image * n * n;
  wv.y0 = seg * seg_h;
  wv.total = (min(wv.y0 + seg_h, n) - wv.y0) + K - 1;
  // the adjoint correlates with the unflipped factors
#pragma unroll
  for (int t = 0; t < K; ++t) {
    wv.wy[t] = w[t];
    wv.wx[t] = w[K + t];
  }
  const int xl = wv.cb + LO4 + 4 * wv.lane;
  const int xh = wv.cb + LO4 + kColBand + 4 * (wv.lane & 3);
  wv.vl = xl >= 0 && xl < n ? 4 * xl : kColOutside;
  wv.vh = xh >= 0 && xh < n ? 4 * xh : kColOutside;
  // block columns from bc0 (a multiple of 4; m % 4 == 0, so a piece is wholly
  // inside or outside the image)
  constexpr int kLog2KD = KD == 4 ? 2 : KD == 2 ? 1 : 0;
  wv.bc0 = ((wv.cb + LO4) >> kLog2KD) & ~3;   // floor, then down to a multiple of 4
  const int bl = wv.bc0 + 4 * wv.lane;
  wv.vbl = bl >= 0 && bl < m ? 4 * bl : kColOutside;
  const int bh = wv.bc0 + kColBand + 4 * (wv.lane & 3);   // KD = 1 only
  wv.vbh = bh >= 0 && bh < m ? 4 * bh : kColOutside;
  wv.mapp = MAP ? post.map[first_map(post)] : dout;
  wv.post2 = &post2;
  float hw[K - 1 + kColU][4];
  float gm[2] = {0.f, 0.f}, gh[2] = {0.f, 0.f};
#pragma unroll
  for (int u = 0; u + 1 < kColRing; ++u)
    if (u < wv.total) wv.issue(u, wv.ring[u]);
  constexpr int kFirst = (K - 1 + kColU - 1) / kColU * kColU;
  int i0 = 0;
  for (; i0 < kFirst && i0 < wv.total; i0 += kColU) wv.template block<true>(i0, hw, gm, gh);
  for (; i0 + kColU + kColRing - 1 <= wv.total; i0 += kColU)
    wv.template block<false>(i0, hw, gm, gh);
  for (; i0 < wv.total; i0 += kColU) wv.template block<true>(i0, hw, gm, gh);
}

// out[b][Y][X] = post( sum_{dy,dx < k} pre(src[b][Y k + dy][X k + dx]) ), m = n / k.
// Row sums first, then the sum of the rows (the order of utils/array_ops.py:151-161).
// Grid: (pixel blocks of the output image, image); in-image indices are 32-bit.
__global__ void __launch_bounds__(256)
det_down_kernel(const float* __restrict__ src, int n, int k, int m, EwList pre, EwList post,
                float* __restrict__ r_save, float* __restrict__ out, SumIn si) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= m * m) return;
  const float* s = src + (long)blockIdx.y * n * n;
  const long o = (long)blockIdx.y * m * m + q;
  float acc;
  if (k == 1) {
    acc = ew_apply(pre, ld_sum(s, q, si), q);
  } else {
    const int oy = q / m, ox = q - oy * m;
    acc = 0.f;
    for (int dy = 0; dy < k; ++dy) {
      const int rp = (oy * k + dy) * n + ox * k;
      float rs = 0.f;
      for (int dx = 0; dx < k; ++dx) rs += ew_apply(pre, ld_sum(s, rp + dx, si), rp + dx);
      acc += rs;
    }
  }
  if (r_save) r_save[o] = acc;
  out[o] = ew_apply(post, acc, q);
}

// Elementwise pass, four pixels per thread (npix % 4 == 0: float4 aligned):
// out = post(pre(src)), r_save = pre(src).
__global__ void __launch_bounds__(256)
det_ew_kernel(const float* __restrict__ src, int npix, EwList pre, EwList post,
              float* __restrict__ r_save, float* __restrict__ out, SumIn si) {
  const int q = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (q >= npix) return;
  const long o = (long)blockIdx.y * npix + q;
  float4 v = *(const float4*)(src + o);
#pragma unroll 8
  for (int pl = 1; pl < si.n; ++pl) {
    const float4 a = *(const float4*)(src + o + (long)pl * si.stride);
    v.x += a.x;
    v.y += a.y;
    v.z += a.z;
    v.w += a.w;
  }
  float4 r;
  r.x = ew_apply(pre, v.x, q);
  r.y = ew_apply(pre, v.y, q + 1);
  r.z = ew_apply(pre, v.z, q + 2);
  r.w = ew_apply(pre, v.w, q + 3);
  if (r_save) *(float4*)(r_save + o) = r;
  float4 w;
  w.x = ew_apply(post, r.x, q);
  w.y = ew_apply(post, r.y, q + 1);
  w.z = ew_apply(post, r.z, q + 2);
  w.w = ew_apply(post, r.w, q + 3);
  *(float4*)(out + o) = w;
}

// Adjoint of the block sum: out[b][y][x] = g[b][y/k][x/k] * dpre'(aux[b][y][x]).
__global__ void __launch_bounds__(256)
det_up_kernel(const float* __restrict__ g, int n, int k, int m, const float* __restrict__ aux,
              EwList dpre, float* __restrict__ out) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n * n) return;
  const long i = (long)blockIdx.y * n * n + p;
  const long gb = (long)blockIdx.y * m * m;
  float v;
  if (k == 1) {
    v = g[gb + p];
  } else {
    const int y = p / n, x = p - y * n;
    v = g[gb + (y / k) * m + x / k];
  }
  out[i] = v * ew_deriv(dpre, aux[i], p);
}

// out = g * post'(r), elementwise over (pixel block, image) (in place allowed).
__global__ void __launch_bounds__(256)
det_ew_grad_kernel(const float* g, const float* __restrict__ r, int npix, EwList post,
                   float* out) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  const long i = (long)blockIdx.y * npix + p;
  out[i] = g[i] * ew_deriv(post, r[i], p);
}

// Parameter cotangents of one elementwise list.  Per pixel, the values entering
// each op are recomputed from v0 (the value entering the list) and the
// cotangent gout (after the list) is walked back through the ops:
//   map m:        dL/dm[p] += v * g,          g *= m[p]
//   saturate t:   dL/dt    += g [v > t] (1/2 at v == t, jnp.minimum's tie rule)
//   add c:        dL/dc    += g
// Maps sum over the batch in the thread (fixed order); scalars leave one
// partial per block and slot for det_param_final_kernel.  With din set, the
// cotangent entering the list is stored too.
struct ParamOut {
  float* map[kMaxEw];
};

__global__ void __launch_bounds__(256)
det_param_grad_kernel(EwList l, const float* __restrict__ v0, const float* __restrict__ gout,
                      int npix, int batch, float* __restrict__ din, ParamOut po,
                      double* __restrict__ part) {
  __shared__ double red[kMaxEw][4];
  const int p = blockIdx.x * 256 + threadIdx.x;
  double mg[kMaxEw], sc[kMaxEw];
#pragma unroll
  for (int j = 0; j < kMaxEw; ++j) { mg[j] = 0.0; sc[j] = 0.0; }
  if (p < npix) {
    for (int b = 0; b < batch; ++b) {
      const long i = (long)b * npix + p;
      float v[kMaxEw + 1];
      v[0] = v0[i];
#pragma unroll
      for (int j = 0; j < kMaxEw; ++j) {
        if (j >= l.n) break;
        const int k = l.kind[j];
        const float x = v[j];
        v[j + 1] = k == DLX_DET_PIXEL_RESPONSE ? x * l.map[j][p]
                   : k == DLX_DET_SATURATE     ? (x > l.value[j] ? l.value[j] : x)
                                               : x + l.value[j];
      }
      float g = gout[i];
#pragma unroll
      for (int j = kMaxEw - 1; j >= 0; --j) {
        if (j >= l.n) continue;
        const int k = l.kind[j];
        const float x = v[j];
        if (k == DLX_DET_PIXEL_RESPONSE) {
          mg[j] += (double)x * g;
          g *= l.map[j][p];
        } else if (k == DLX_DET_SATURATE) {
          const float t = l.value[j];
          sc[j] += (double)g * (x > t ? 1.0 : (x == t ? 0.5 : 0.0));
          g *= x < t ? 1.f : (x == t ? 0.5f : 0.f);
        } else {
          sc[j] += (double)g;
        }
      }
      if (din) din[i] = g;
    }
#pragma unroll
    for (int j = 0; j < kMaxEw; ++j)
      if (j < l.n && po.map[j]) po.map[j][p] = (float)mg[j];
  }
  // block sums of the scalar slots: wave shuffles, then the 4 waves in order
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kMaxEw; ++j) {
    double x = sc[j];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[j][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x < kMaxEw) {
    const int j = threadIdx.x;
    part[(long)blockIdx.x * kMaxEw + j] = red[j][0] + red[j][1] + red[j][2] + red[j][3];
  }
}

struct ScalarOut {
  float* out[kMaxEw];
};

// Cotangent of a convolution kernel: dL/dK[a][b] = sum over outputs i of
// g[i] * x[i + c - (a, b)] (x as the convolution saw it: load-side ops
// applied, zero outside the image; c = (K - 1) / 2 per axis).  One workgroup
// per 64 x 16 output tile and image stages the tile's x window and g in LDS
// and leaves one FP64 partial per tap; det_kgrad_final_kernel sums them.
__global__ void __launch_bounds__(256)
det_kgrad_kernel(const float* __restrict__ x, const float* __restrict__ g, int n, int tiles_x,
                 int K, EwList pre, double* __restrict__ part) {
  extern __shared__ float sm[];
  const int W = kTileX + K - 1, H = kTileY + K - 1;
  float* xs = sm;
  float* gs = sm + H * W;
  const int tile = blockIdx.x;
  const int tx0 = (tile % tiles_x) * kTileX, ty0 = (tile / tiles_x) * kTileY;
  const int c = (K - 1) / 2;
  const long img = (long)blockIdx.y * n * n;
  for (int e = threadIdx.x; e < H * W; e += 256) {
    const int r = e / W, q = e - r * W;
    const int y = ty0 + c - (K - 1) + r, xx = tx0 + c - (K - 1) + q;
    float v = 0.f;
    if (y >= 0 && y < n && xx >= 0 && xx < n) {
      const long p = (long)y * n + xx;
      v = ew_apply(pre, x[img + p], p);
    }
    xs[e] = v;
  }
  for (int e = threadIdx.x; e < kTileY * kTileX; e += 256) {
    const int r = e / kTileX, q = e - r * kTileX;
    const int y = ty0 + r, xx = tx0 + q;
    gs[e] = (y < n && xx < n) ? g[img + (long)y * n + xx] : 0.f;
  }
  __syncthreads();
  const long base = ((long)blockIdx.y * gridDim.x + tile) * K * K;
  for (int t = threadIdx.x; t < K * K; t += 256) {
    const int a = t / K, bb = t - a * K;
    double acc = 0.0;
    for (int iy = 0; iy < kTileY; ++iy) {
      const float* xr = xs + (iy + K - 1 - a) * W + (K - 1 - bb);
      const float* gr = gs + iy * kTileX;
      float sum = 0.f;
      for (int ix = 0; ix < kTileX; ++ix) sum += gr[ix] * xr[ix];
      acc += sum;
    }
    part[base + t] = acc;
  }
}

// One workgroup per tap: the fixed-order sum of its nparts partials.
__global__ void __launch_bounds__(256)
det_kgrad_final_kernel(const double* __restrict__ part, long nparts, int K2,
                       float* __restrict__ out) {
  __shared__ double red[256];
  const int t = blockIdx.x;
  double acc = 0.0;
  for (long w = threadIdx.x; w < nparts; w += 256) acc += part[w * K2 + t];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[t] = (float)red[0];
}

// Separable kernels K[a][b] = wy[a] wx[b]: dL/dwy[a] = sum_b dK[a][b] wx[b],
// dL/dwx[b] = sum_a dK[a][b] wy[a] (out: the K column then the K row values).
__global__ void __launch_bounds__(128)
det_kgrad_sep_kernel(const float* __restrict__ dK, int K, const float* __restrict__ w,
                     float* __restrict__ out) {
  for (int t = threadIdx.x; t < 2 * K; t += 128) {
    double acc = 0.0;
    if (t < K) {
      for (int b = 0; b < K; ++b) acc += (double)dK[t * K + b] * w[K + b];
    } else {
      const int b = t - K;
      for (int a = 0; a < K; ++a) acc += (double)dK[a * K + b] * w[a];
    }
    out[t] = (float)acc;
  }
}

__global__ void __launch_bounds__(64)
det_param_final_kernel(const double* __restrict__ part, int nblk, ScalarOut so) {
  const int j = threadIdx.x;
  if (j >= kMaxEw || !so.out[j]) return;
  double acc = 0.0;
  for (int b = 0; b < nblk; ++b) acc += part[(long)b * kMaxEw + j];
  so.out[j][0] = (float)acc;
}

struct Pass {
  int op_index;    // the convolve op's index in the caller's list
  int conv;        // 1: convolve, 0: block sum (k = 1: elementwise only)
  int n_in, n_out;
  int K;           // convolve kernel side
  int sep;         // 1: w holds K column then K row weights
  const float* w;  // convolve kernel (device)
  int k;           // block-sum factor
  EwList pre, post;
};

bool ew_push(EwList& l, const dlx_detector_op& op, int index) {
  if (l.n == kMaxEw) return false;
  l.kind[l.n] = op.kind;
  l.value[l.n] = (float)op.value;
  l.map[l.n] = op.data;
  l.op[l.n] = index;
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
        if (!ew_push(target, op, i)) {
          Pass p = make_pass(0, cur);
          if (passes.empty()) {
            p.pre = pending;
            pending = EwList{};
          }
          passes.push_back(p);
          ew_push(passes.back().post, op, i);
        }
        break;
      }
      case DLX_DET_CONVOLVE:
      case DLX_DET_CONVOLVE_SEPARABLE: {
        DLX_CHECK_ARG(op.data, "convolution %d has no kernel", i);
        DLX_CHECK_ARG(op.size >= 1 && op.size <= kMaxKernel,
                      "convolution kernel side must be in [1, %d], got %d", kMaxKernel, op.size);
        Pass p = make_pass(1, cur);
        p.op_index = i;
        p.K = op.size;
        p.sep = op.kind == DLX_DET_CONVOLVE_SEPARABLE;
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
  // in-image pixel indices are 32-bit
  DLX_CHECK_ARG(passes.front().n_in <= 46340, "detector image side must be <= 46340, got %d",
                passes.front().n_in);
  return DLX_OK;
}

// Forward pass i: src -> dst (r_save may be NULL).
// The column-walking kernel takes a separable forward convolution of odd K <=
// 13 on a single (not summed) input whose side is a multiple of 4, alone or
// with a 2 x 2 / 4 x 4 block sum, when the batch is large enough: at least 512
// (256-column band, segment of rows, image) waves.  A wave walks its rows one
// after the other, so on a small batch (one 256^2 frame: 8 waves) the tile
// kernels, which cover the image with many short-lived workgroups, win.
bool streams(const Pass& c, int k, const EwList& post2, SumIn si, int batch) {
  const long bands = (c.n_in + kColBand - 1) / kColBand;
  const long segs = std::max(1, c.n_in / 32);
  // the kernel reads no map in its loop but the one it streams: no op before
  // the convolution, no map after the block sum, at most one map after the
  // convolution (16-byte aligned: it is read 16 bytes a lane)
  int maps = 0;
  bool map16 = true;
  for (int i = 0; i < c.post.n; ++i)
    if (c.post.kind[i] == DLX_DET_PIXEL_RESPONSE) {
      ++maps;
      map16 = (reinterpret_cast<uintptr_t>(c.post.map[i]) & 15) == 0;
    }
  bool others = c.pre.n > 0;
  for (int i = 0; i < post2.n; ++i) others = others || post2.kind[i] == DLX_DET_PIXEL_RESPONSE;
  return c.conv && c.sep && c.K <= 13 && c.K % 2 == 1 && c.n_in % 4 == 0 &&
         (k == 1 || k == 2 || k == 4) && si.n == 1 && maps <= 1 && map16 && !others &&
         bands * segs * batch >= 512;
}

// Bands of kColBand columns; segments of at least 32 rows (a multiple of 4),
// about kColWaves waves in all: one resident round at 16 one-wave workgroups
// per CU.
constexpr long kColWaves = 4096;

void stream_grid(int n, int batch, int& bands, int& seg_h, int& nseg) {
  bands = (n + kColBand - 1) / kColBand;
  const long per = (long)bands * batch;
  long want = (kColWaves + per - 1) / per;
  want = std::min<long>(want, std::max(1, n / 32));
  seg_h = (int)(((n + want - 1) / want + 3) / 4 * 4);
  nseg = (n + seg_h - 1) / seg_h;
}

template <int KD, int K>
void launch_col(const dim3& grid, hipStream_t st, const Pass& c, int bands, int seg_h, int m,
                const EwList& post2, const float* src, float* dst) {
  bool map = false;
  for (int i = 0; i < c.post.n; ++i) map = map || c.post.kind[i] == DLX_DET_PIXEL_RESPONSE;
  if (map)
    hipLaunchKernelGGL((det_sep_col_kernel<KD, K, true>), grid, dim3(64), 0, st, src, c.n_in,
                       bands, seg_h, c.w, c.post, m, post2, dst);
  else
    hipLaunchKernelGGL((det_sep_col_kernel<KD, K, false>), grid, dim3(64), 0, st, src, c.n_in,
                       bands, seg_h, c.w, c.post, m, post2, dst);
}

template <int KD>
void launch_col_k(const dim3& grid, hipStream_t st, const Pass& c, int bands, int seg_h, int m,
                  const EwList& post2, const float* src, float* dst) {
  switch (c.K) {
    case 1: launch_col<KD, 1>(grid, st, c, bands, seg_h, m, post2, src, dst); break;
    case 3: launch_col<KD, 3>(grid, st, c, bands, seg_h, m, post2, src, dst); break;
    case 5: launch_col<KD, 5>(grid, st, c, bands, seg_h, m, post2, src, dst); break;
    case 7: launch_col<KD, 7>(grid, st, c, bands, seg_h, m, post2, src, dst); break;
    case 9: launch_col<KD, 9>(grid, st, c, bands, seg_h, m, post2, src, dst); break;
    case 11: launch_col<KD, 11>(grid, st, c, bands, seg_h, m, post2, src, dst); break;
    default: launch_col<KD, 13>(grid, st, c, bands, seg_h, m, post2, src, dst); break;
  }
}

int run_stream(const Pass& c, int k, int m, const EwList& post2, int batch, const float* src,
               float* dst, hipStream_t st) {
  int bands, seg_h, nseg;
  stream_grid(c.n_in, batch, bands, seg_h, nseg);
  const dim3 grid(bands * nseg, batch);
  if (k == 4) launch_col_k<4>(grid, st, c, bands, seg_h, m, post2, src, dst);
  else if (k == 2) launch_col_k<2>(grid, st, c, bands, seg_h, m, post2, src, dst);
  else launch_col_k<1>(grid, st, c, bands, seg_h, c.n_in, EwList{}, src, dst);
  return launch_status("det_sep_col_kernel");
}

template <int KD, int K>
void launch_col_vjp(const dim3& grid, hipStream_t st, const Pass& c, int bands, int seg_h, int m,
                    const EwList& post2, const float* dout, const float* bsave, float* din) {
  bool map = false;
  for (int i = 0; i < c.post.n; ++i) map = map || c.post.kind[i] == DLX_DET_PIXEL_RESPONSE;
  if (map)
    hipLaunchKernelGGL((det_sep_col_vjp_kernel<KD, K, true>), grid, dim3(64), 0, st, dout, bsave,
                       c.n_in, bands, seg_h, c.w, c.post, m, post2, din);
  else
    hipLaunchKernelGGL((det_sep_col_vjp_kernel<KD, K, false>), grid, dim3(64), 0, st, dout,
                       bsave, c.n_in, bands, seg_h, c.w, c.post, m, post2, din);
}

template <int KD>
void launch_col_vjp_k(const dim3& grid, hipStream_t st, const Pass& c, int bands, int seg_h,
                      int m, const EwList& post2, const float* dout, const float* bsave,
                      float* din) {
  switch (c.K) {
    case 1: launch_col_vjp<KD, 1>(grid, st, c, bands, seg_h, m, post2, dout, bsave, din); break;
    case 3: launch_col_vjp<KD, 3>(grid, st, c, bands, seg_h, m, post2, dout, bsave, din); break;
    case 5: launch_col_vjp<KD, 5>(grid, st, c, bands, seg_h, m, post2, dout, bsave, din); break;
    case 7: launch_col_vjp<KD, 7>(grid, st, c, bands, seg_h, m, post2, dout, bsave, din); break;
    case 9: launch_col_vjp<KD, 9>(grid, st, c, bands, seg_h, m, post2, dout, bsave, din); break;
    case 11: launch_col_vjp<KD, 11>(grid, st, c, bands, seg_h, m, post2, dout, bsave, din); break;
    default: launch_col_vjp<KD, 13>(grid, st, c, bands, seg_h, m, post2, dout, bsave, din); break;
  }
}

bool aligned16(const void* a, const void* b) {
  return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
}

// Rows per thread of a tile launch: two where the LDS allows, one when two
// would leave fewer than 256 workgroups (one small frame: 64 x 16 tiles halve
// the latency of one 64 x 32 tile each) and the block height k still fits.
int tile_rpt(const Pass& c, int k, int batch) {
  const int rpt = conv_rpt(c.K, c.sep);
  if (rpt == 1 || k > kTileY) return rpt;
  const int th = k > 1 ? (kTileY * 2 / k) * k : kTileY * 2, tw = k > 1 ? (kTileX / k) * k : kTileX;
  const long tiles = (long)((c.n_in + tw - 1) / tw) * ((c.n_in + th - 1) / th);
  return tiles * batch < 256 ? 1 : 2;
}

int run_pass(const Pass& p, int batch, const float* src, float* r_save, float* dst,
             hipStream_t st, SumIn si = SumIn{1, 0}) {
  if (p.conv && !r_save && streams(p, 1, EwList{}, si, batch) && aligned16(src, dst))
    return run_stream(p, 1, p.n_in, EwList{}, batch, src, dst, st);
  if (p.conv) {
    const int rpt = tile_rpt(p, 1, batch);
    const int tiles_x = (p.n_in + kTileX - 1) / kTileX;
    const int tiles_y = (p.n_in + kTileY * rpt - 1) / (kTileY * rpt);
    const int lo = (p.K - 1) / 2 - (p.K - 1);
    const size_t lds = sizeof(float) * (size_t)conv_lds_floats(p.K, p.sep, rpt);
    const dim3 grid(tiles_x * tiles_y, batch);
    if (rpt == 2)
      hipLaunchKernelGGL(det_corr_kernel<2>, grid, dim3(kThreads), lds, st, src, p.n_in, tiles_x,
                         p.K, lo, p.w, 1, p.sep, p.pre, p.post, (const float*)nullptr, EwList{},
                         r_save, dst, si);
    else
      hipLaunchKernelGGL(det_corr_kernel<1>, grid, dim3(kThreads), lds, st, src, p.n_in, tiles_x,
                         p.K, lo, p.w, 1, p.sep, p.pre, p.post, (const float*)nullptr, EwList{},
                         r_save, dst, si);
    return launch_status("det_corr_kernel");
  }
  const int npix = p.n_out * p.n_out;
  const bool aligned = ((uintptr_t)src | (uintptr_t)dst | (uintptr_t)r_save) % 16 == 0 &&
                       (si.n == 1 || si.stride % 4 == 0);
  if (p.k == 1 && npix % 4 == 0 && aligned) {
    hipLaunchKernelGGL(det_ew_kernel, dim3((npix / 4 + 255) / 256, batch), dim3(256), 0, st, src,
                       npix, p.pre, p.post, r_save, dst, si);
    return launch_status("det_ew_kernel");
  }
  hipLaunchKernelGGL(det_down_kernel, dim3((npix + 255) / 256, batch), dim3(256), 0, st, src,
                     p.n_in, p.k, p.n_out, p.pre, p.post, r_save, dst, si);
  return launch_status("det_down_kernel");
}

// A convolve pass followed by a block-sum pass whose factor fits the tile
// height runs as one det_conv_down_kernel launch.
bool can_fuse(const std::vector<Pass>& passes, size_t i) {
  return passes[i].conv && i + 1 < passes.size() && !passes[i + 1].conv &&
         passes[i + 1].k <= kTileY * conv_rpt(passes[i].K, passes[i].sep) &&
         passes[i + 1].pre.n == 0;
}

int run_conv_down(const Pass& c, const Pass& d, int batch, const float* src, float* dst,
                  hipStream_t st, SumIn si = SumIn{1, 0}) {
  if (streams(c, d.k, d.post, si, batch) && aligned16(src, dst))
    return run_stream(c, d.k, d.n_out, d.post, batch, src, dst, st);
  const int rpt = tile_rpt(c, d.k, batch);
  const int k = d.k;
  const int th = (kTileY * rpt / k) * k, tw = (kTileX / k) * k;
  const int tiles_x = (c.n_in + tw - 1) / tw;
  const int tiles_y = (c.n_in + th - 1) / th;
  const int lo = (c.K - 1) / 2 - (c.K - 1);
  const size_t lds = sizeof(float) * (size_t)conv_lds_floats(c.K, c.sep, rpt);
  const dim3 grid(tiles_x * tiles_y, batch);
  if (rpt == 2)
    hipLaunchKernelGGL(det_conv_down_kernel<2>, grid, dim3(kThreads), lds, st, src, c.n_in, th,
                       tw, tiles_x, c.K, lo, c.w, c.sep, c.pre, c.post, k, d.n_out, d.post, dst,
                       si);
  else
    hipLaunchKernelGGL(det_conv_down_kernel<1>, grid, dim3(kThreads), lds, st, src, c.n_in, th,
                       tw, tiles_x, c.K, lo, c.w, c.sep, c.pre, c.post, k, d.n_out, d.post, dst,
                       si);
  return launch_status("det_conv_down_kernel");
}

}  // namespace

// Whether the chain's first launch (pass 0, fused with pass 1 where
// run_conv_down takes both) is the column-walking kernel, given aligned images.
bool first_streams(const std::vector<Pass>& passes, int batch) {
  const bool fuse = can_fuse(passes, 0);
  return streams(passes[0], fuse ? passes[1].k : 1, fuse ? passes[1].post : EwList{},
                 SumIn{1, 0}, batch);
}

// Workgroups of the chain's first launch.
long first_launch_groups(const std::vector<Pass>& passes, int batch) {
  const Pass& p = passes[0];
  if (first_streams(passes, batch)) {
    int bands, seg_h, nseg;
    stream_grid(p.n_in, batch, bands, seg_h, nseg);
    return (long)bands * nseg * batch;
  }
  if (p.conv) {
    const int rpt = tile_rpt(p, can_fuse(passes, 0) ? passes[1].k : 1, batch);
    long tiles;
    if (can_fuse(passes, 0)) {
      const int k = passes[1].k;
      const int th = (kTileY * rpt / k) * k, tw = (kTileX / k) * k;
      tiles = (long)((p.n_in + tw - 1) / tw) * ((p.n_in + th - 1) / th);
    } else {
      tiles = (long)((p.n_in + kTileX - 1) / kTileX) * ((p.n_in + kTileY * rpt - 1) / (kTileY * rpt));
    }
    return tiles * batch;
  }
  const long npix = (long)p.n_out * p.n_out;
  return (p.k == 1 ? (npix / 4 + 255) / 256 : (npix + 255) / 256) * batch;
}

// A summed input (nsum planes) is loaded by the first pass, summing as it
// loads, when that launch fills the GPU and is a tile kernel; otherwise it is
// summed first by the pixel-parallel plane sum (same order, same bits) into
// one more image of workspace: a small first launch's few workgroups would
// each add up nsum planes serially, and the streaming kernel loads one image.
constexpr long kSumInLoaderGroups = 512;

bool presum(const std::vector<Pass>& passes, int n, int batch, int nsum, long sum_stride) {
  return nsum > 1 && sum_stride == (long)n * n &&
         (first_launch_groups(passes, batch) < kSumInLoaderGroups ||
          first_streams(passes, batch));
}

size_t detector_ws_bytes(int n, int batch, const dlx_detector_op* ops, int nops, int nsum) {
  std::vector<Pass> passes;
  if (batch < 0 || plan(n, ops, nops, passes)) return 0;
  const size_t extra =
      presum(passes, n, batch, nsum, (long)n * n) ? align_up((size_t)n * n * sizeof(float)) : 0;
  return layout(passes, batch, false).total + extra;
}

int detector_apply_sum(int n, int batch, int nsum, long sum_stride, const float* in,
                       const dlx_detector_op* ops, int nops, float* out, void* workspace,
                       size_t workspace_bytes, hipStream_t s) {
  std::vector<Pass> passes;
  int st = plan(n, ops, nops, passes);
  if (st) return st;
  DLX_CHECK_ARG(batch >= 0, "negative detector batch");
  DLX_CHECK_ARG(nsum >= 1 && (nsum == 1 || batch == 1),
                "a summed detector input needs one image and nsum >= 1 (got batch %d, nsum %d)",
                batch, nsum);
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(in && out, "null image pointer to the detector chain");
  if ((st = check_grid(passes, batch))) return st;
  const Layout L = layout(passes, batch, false);
  const bool pre = presum(passes, n, batch, nsum, sum_stride);
  const size_t need = L.total + (pre ? align_up((size_t)n * n * sizeof(float)) : 0);
  if (workspace_bytes < need || (need && !workspace)) {
    set_error("detector chain workspace too small: need %zu bytes, got %zu", need,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  char* ws = (char*)workspace;
  const float* src = in;
  if (pre) {
    float* sum = (float*)(ws + L.total);
    DLX_HIP(launch_sum_planes((long)n * n, nsum, in, 0, sum, s));
    src = sum;
    nsum = 1;
    sum_stride = 0;
  }
  const size_t np = passes.size();
  int buf = 0;
  for (size_t i = 0; i < np;) {
    const bool fuse = can_fuse(passes, i);
    const size_t next = i + (fuse ? 2 : 1);
    float* dst = next == np ? out : (float*)(ws + L.g_off[buf]);
    buf ^= 1;
    const SumIn si = i == 0 ? SumIn{nsum, sum_stride} : SumIn{1, 0};
    st = fuse ? run_conv_down(passes[i], passes[i + 1], batch, src, dst, s, si)
              : run_pass(passes[i], batch, src, nullptr, dst, s, si);
    if (st) return st;
    src = dst;
    i = next;
  }
  return DLX_OK;
}

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
  return detector_apply_sum(n, batch, 1, 0, in, ops, nops, out, workspace, workspace_bytes,
                            (hipStream_t)stream);
}

}  // extern "C"

namespace {

// Whether any op of l has a parameter cotangent requested.
bool wants(const EwList& l, float* const* grads) {
  if (!grads) return false;
  for (int j = 0; j < l.n; ++j)
    if (grads[l.op[j]]) return true;
  return false;
}

size_t kgrad_parts(const Pass& p, int batch) {
  const size_t tiles = (size_t)((p.n_in + kTileX - 1) / kTileX) * ((p.n_in + kTileY - 1) / kTileY);
  return tiles * batch * p.K * p.K;
}

size_t kgrad_bytes(const std::vector<Pass>& passes, int batch) {
  size_t most = 0;
  for (const Pass& p : passes)
    if (p.conv)
      most = std::max(most, align_up(kgrad_parts(p, batch) * sizeof(double)) +
                                align_up((size_t)p.K * p.K * sizeof(float)));
  return most;
}

size_t param_extra(const std::vector<Pass>& passes, int batch) {
  int side_max = 0;
  for (const Pass& p : passes) side_max = std::max(side_max, std::max(p.n_in, p.n_out));
  const size_t nblk = ((size_t)side_max * side_max + 255) / 256;
  return img_bytes(passes[0].n_in, batch) + align_up(nblk * kMaxEw * sizeof(double)) +
         kgrad_bytes(passes, batch);
}

// dL/d(kernel) of convolve pass p (K x K) or of its separable factors (2K).
int kernel_grad(const Pass& p, const float* x, const float* g, int batch, float* out,
                uint8_t* ws, hipStream_t s) {
  const int tiles_x = (p.n_in + kTileX - 1) / kTileX;
  const int tiles_y = (p.n_in + kTileY - 1) / kTileY;
  const long nparts = (long)tiles_x * tiles_y * batch;
  double* part = (double*)ws;
  float* dK = (float*)(ws + align_up(kgrad_parts(p, batch) * sizeof(double)));
  const size_t lds = sizeof(float) * ((size_t)(kTileX + p.K - 1) * (kTileY + p.K - 1) +
                                      kTileX * kTileY);
  hipLaunchKernelGGL(det_kgrad_kernel, dim3(tiles_x * tiles_y, batch), dim3(256), lds, s, x, g,
                     p.n_in, tiles_x, p.K, p.pre, part);
  if (int st = launch_status("det_kgrad_kernel")) return st;
  const int K2 = p.K * p.K;
  hipLaunchKernelGGL(det_kgrad_final_kernel, dim3(K2), dim3(256), 0, s, (const double*)part,
                     nparts, K2, p.sep ? dK : out);
  if (int st = launch_status("det_kgrad_final_kernel")) return st;
  if (p.sep) {
    hipLaunchKernelGGL(det_kgrad_sep_kernel, dim3(1), dim3(128), 0, s, (const float*)dK, p.K,
                       p.w, out);
    if (int st = launch_status("det_kgrad_sep_kernel")) return st;
  }
  return DLX_OK;
}

int param_grads(const EwList& l, float* const* grads, const float* v0, const float* gout,
                int side, int batch, float* din, double* part, hipStream_t s) {
  ParamOut po{};
  ScalarOut so{};
  for (int j = 0; j < l.n; ++j) {
    float* gp = grads[l.op[j]];
    if (l.kind[j] == DLX_DET_PIXEL_RESPONSE) po.map[j] = gp;
    else so.out[j] = gp;
  }
  const int npix = side * side;
  const int nblk = (npix + 255) / 256;
  hipLaunchKernelGGL(det_param_grad_kernel, dim3(nblk), dim3(256), 0, s, l, v0, gout, npix,
                     batch, din, po, part);
  if (int st = launch_status("det_param_grad_kernel")) return st;
  hipLaunchKernelGGL(det_param_final_kernel, dim3(1), dim3(64), 0, s, (const double*)part, nblk,
                     so);
  return launch_status("det_param_final_kernel");
}

// The image cotangent of a chain the column-walking kernel runs forward, by
// its adjoint kernel: a convolve pass (no op before it; after it only adds
// and at most one pixel response, whose derivatives need no values) alone or
// fused with a block sum of 2 or 4 (whose ops hold no map), on sides that are
// multiples of 4 KD.
bool vjp_streams(const std::vector<Pass>& passes, int batch) {
  const Pass& c = passes[0];
  const bool fuse = passes.size() == 2 && can_fuse(passes, 0);
  if (passes.size() != 1 && !fuse) return false;
  const int k = fuse ? passes[1].k : 1;
  if (k != 1 && k != 2 && k != 4) return false;
  for (int i = 0; i < c.post.n; ++i)
    if (c.post.kind[i] != DLX_DET_PIXEL_RESPONSE && c.post.kind[i] != DLX_DET_ADD_CONSTANT)
      return false;
  return c.n_in % (4 * k) == 0 &&
         streams(c, k, fuse ? passes[1].post : EwList{}, SumIn{1, 0}, batch);
}

// Forward recompute of the block sums entering the last ops (KD > 1) into bs,
// then the adjoint walk into din.
int run_stream_vjp(const std::vector<Pass>& passes, int batch, const float* in,
                   const float* dout, float* bs, float* din, hipStream_t st) {
  const Pass& c = passes[0];
  const bool fuse = passes.size() == 2;
  const int k = fuse ? passes[1].k : 1;
  const int m = c.n_in / k;
  const EwList post2 = fuse ? passes[1].post : EwList{};
  int res;
  if (k > 1 && (res = run_stream(c, k, m, EwList{}, batch, in, bs, st))) return res;
  int bands, seg_h, nseg;
  stream_grid(c.n_in, batch, bands, seg_h, nseg);
  const dim3 grid(bands * nseg, batch);
  if (k == 4) launch_col_vjp_k<4>(grid, st, c, bands, seg_h, m, post2, dout, bs, din);
  else if (k == 2) launch_col_vjp_k<2>(grid, st, c, bands, seg_h, m, post2, dout, bs, din);
  else launch_col_vjp_k<1>(grid, st, c, bands, seg_h, m, EwList{}, dout, dout, din);
  return launch_status("det_sep_col_vjp_kernel");
}

int detector_vjp_impl(int32_t n, int32_t batch, const float* in, const dlx_detector_op* ops,
                      int32_t nops, const float* dout, float* din, float* const* grads,
                      void* workspace, size_t workspace_bytes, void* stream, const char* name) {
  std::vector<Pass> passes;
  int st = plan(n, ops, nops, passes);
  if (st) return st;
  DLX_CHECK_ARG(batch >= 0, "negative detector batch");
  if (grads)
    for (int k = 0; k < nops; ++k)
      DLX_CHECK_ARG(!grads[k] || ops[k].kind != DLX_DET_DOWNSAMPLE,
                    "op %d: a downsample factor has no gradient", k);
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(in && dout && din, "null image pointer to %s", name);
  if ((st = check_grid(passes, batch))) return st;
  const Layout L = layout(passes, batch, true);
  const size_t need = L.total + (grads ? param_extra(passes, batch) : 0);
  if (workspace_bytes < need || !workspace) {
    set_error("%s workspace too small: need %zu bytes, got %zu", name, need, workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  const size_t np = passes.size();
  float* g[2] = {(float*)(ws + L.g_off[0]), (float*)(ws + L.g_off[1])};
  float* T = grads ? (float*)(ws + L.total) : nullptr;
  double* part = grads ? (double*)(ws + L.total + img_bytes(passes[0].n_in, batch)) : nullptr;
  uint8_t* kws = nullptr;
  if (grads) {
    int side_max = 0;
    for (const Pass& q : passes) side_max = std::max(side_max, std::max(q.n_in, q.n_out));
    const size_t nblk = ((size_t)side_max * side_max + 255) / 256;
    kws = (uint8_t*)part + align_up(nblk * kMaxEw * sizeof(double));
  }
  if (!grads && vjp_streams(passes, batch) && aligned16(in, din) && aligned16(dout, ws))
    return run_stream_vjp(passes, batch, in, dout, (float*)ws, din, s);
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
    const float* gr = cur;
    int gr_idx = cur_idx;
    if (p.post.n > 0) {
      if (wants(p.post, grads) &&
          (st = param_grads(p.post, grads, (const float*)(ws + L.r_off[ii]), cur, p.n_out,
                            batch, nullptr, part, s)))
        return st;
      gr_idx = cur_idx < 0 ? 0 : cur_idx;
      const int npix = p.n_out * p.n_out;
      hipLaunchKernelGGL(det_ew_grad_kernel, dim3((npix + 255) / 256, batch), dim3(256), 0, s,
                         cur, (const float*)(ws + L.r_off[ii]), npix, p.post, g[gr_idx]);
      if ((st = launch_status("det_ew_grad_kernel"))) return st;
      gr = g[gr_idx];
    }
    const int dst_idx = gr_idx == 0 ? 1 : 0;
    float* dst = ii == 0 ? din : g[dst_idx];
    if (p.conv && grads && grads[p.op_index] &&
        (st = kernel_grad(p, x_in(ii), gr, batch, grads[p.op_index], kws, s)))
      return st;
    // a pre list with parameter cotangents: the adjoint without it into T,
    // then one pass for the list's parameters and the cotangent entering it
    const bool pre_params = ii == 0 && wants(p.pre, grads);
    const EwList pre = pre_params ? EwList{} : p.pre;
    float* adst = pre_params ? T : dst;
    if (p.conv) {
      const int rpt = conv_rpt(p.K, p.sep);
      const int tiles_x = (p.n_in + kTileX - 1) / kTileX;
      const int tiles_y = (p.n_in + kTileY * rpt - 1) / (kTileY * rpt);
      const int lo = -((p.K - 1) / 2);
      const size_t lds = sizeof(float) * (size_t)conv_lds_floats(p.K, p.sep, rpt);
      const dim3 grid(tiles_x * tiles_y, batch);
      if (rpt == 2)
        hipLaunchKernelGGL(det_corr_kernel<2>, grid, dim3(kThreads), lds, s, gr, p.n_in, tiles_x,
                           p.K, lo, p.w, 0, p.sep, EwList{}, EwList{}, x_in(ii), pre,
                           (float*)nullptr, adst, SumIn{1, 0});
      else
        hipLaunchKernelGGL(det_corr_kernel<1>, grid, dim3(kThreads), lds, s, gr, p.n_in, tiles_x,
                           p.K, lo, p.w, 0, p.sep, EwList{}, EwList{}, x_in(ii), pre,
                           (float*)nullptr, adst, SumIn{1, 0});
      if ((st = launch_status("det_corr_kernel"))) return st;
    } else {
      const int npix = p.n_in * p.n_in;
      hipLaunchKernelGGL(det_up_kernel, dim3((npix + 255) / 256, batch), dim3(256), 0, s, gr,
                         p.n_in, p.k, p.n_out, x_in(ii), pre, adst);
      if ((st = launch_status("det_up_kernel"))) return st;
    }
    if (pre_params &&
        (st = param_grads(p.pre, grads, in, T, p.n_in, batch, dst, part, s)))
      return st;
    cur = dst;
    cur_idx = dst_idx;
  }
  return DLX_OK;
}

}  // namespace

extern "C" {

int dlx_detector_vjp(int32_t n, int32_t batch, const float* in, const dlx_detector_op* ops,
                     int32_t nops, const float* dout, float* din, void* workspace,
                     size_t workspace_bytes, void* stream) {
  clear_error();
  return detector_vjp_impl(n, batch, in, ops, nops, dout, din, nullptr, workspace,
                           workspace_bytes, stream, "dlx_detector_vjp");
}

size_t dlx_detector_vjp_params_workspace_size(int32_t n, int32_t batch,
                                              const dlx_detector_op* ops, int32_t nops) {
  std::vector<Pass> passes;
  if (batch < 0 || plan(n, ops, nops, passes)) return 0;
  return layout(passes, batch, true).total + param_extra(passes, batch);
}

int dlx_detector_vjp_params(int32_t n, int32_t batch, const float* in,
                            const dlx_detector_op* ops, int32_t nops, const float* dout,
                            float* din, float* const* grads, void* workspace,
                            size_t workspace_bytes, void* stream) {
  clear_error();
  DLX_CHECK_ARG(grads, "null parameter-gradient list to dlx_detector_vjp_params");
  return detector_vjp_impl(n, batch, in, ops, nops, dout, din, grads, workspace,
                           workspace_bytes, stream, "dlx_detector_vjp_params");
}

}  // extern "C"
