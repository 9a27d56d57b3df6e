// Device builders for the wavelength-independent inputs of the hot path:
// Noll-indexed Zernike bases and static (oversampled, box-averaged) circular
// apertures with a secondary obscuration and spider vanes.  One lane per
// output pixel; FP64 arithmetic (these run once per optical system and feed
// resident device buffers).
#include <algorithm>
#include <cmath>
#include <vector>

#include "dlx_internal.h"

namespace dlx {
namespace {

constexpr int kMaxTerms = 24;   // radial terms per mode (n <= 2*24-2 for m = 0)
constexpr int kModesPerLaunch = 8;

struct ZernChunk {
  int n[kModesPerLaunch], m[kModesPerLaunch], nterms[kModesPerLaunch];
  double c[kModesPerLaunch][kMaxTerms];
  double norm[kModesPerLaunch];
  int kmax[kModesPerLaunch];      // terms with a finite coefficient: k <= (n - |m|)/2
  int nan0[kModesPerLaunch];      // a zero-coefficient term has a negative power of rho
};

// Z_q(rho, theta) with cos(theta) = ct, sin(theta) = st: the radial sum as
// rho^|m| times a Horner polynomial in rho^2 over the finite-coefficient terms
// (the others have coefficient 0 from 1/Gamma(<= 0); at rho == 0 a negative
// power makes them 0 * inf = NaN, kept via nan0), and cos/sin(|m| theta) by
// complex powers of (ct + i st).
__device__ __forceinline__ double zern_eval(const ZernChunk& z, int q, double rho, double ct,
                                            double st) {
  const int am = z.m[q] < 0 ? -z.m[q] : z.m[q];
  const double u = rho * rho;
  double acc = z.c[q][0];
  for (int k = 1; k <= z.kmax[q]; ++k) acc = acc * u + z.c[q][k];
  double rp = 1.0, cr = 1.0, ci = 0.0;
  for (int k = 0; k < am; ++k) {
    rp *= rho;
    const double t = cr * ct - ci * st;
    ci = cr * st + ci * ct;
    cr = t;
  }
  double radial = acc * rp;
  if (z.nan0[q] && rho == 0.0) radial = NAN;
  const double trig = z.m[q] >= 0 ? cr : ci;
  return radial * (z.norm[q] * trig);
}

// Z_j = [rho <= 1] * sum_k c_k rho^(n - 2k) * norm * trig(|m| theta)
// (utils/zernikes.py:153-204, :227-257).  Terms whose coefficient is zero
// because a factorial argument is negative are kept, so rho^(negative) at
// rho == 0 yields NaN exactly where the reference does.
__global__ void __launch_bounds__(256)
zernike_kernel(ZernChunk z, int nmodes, long npix, const float* __restrict__ xs,
               const float* __restrict__ ys, double inv_r, float* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const double x = (double)xs[p] * inv_r, y = (double)ys[p] * inv_r;
  const double rho = hypot(x, y);
  const double ct = rho > 0.0 ? x / rho : 1.0, st = rho > 0.0 ? y / rho : 0.0;   // atan2(0, 0) = 0
  const double ap = rho <= 1.0 ? 1.0 : 0.0;
  for (int q = 0; q < nmodes; ++q)
    out[(long)q * npix + p] = (float)(ap * zern_eval(z, q, rho, ct, st));
}

// Static circular aperture (utils/apertures.py:153-242): oversampled hard
// circle (r < D/2), optional secondary (r > Ds/2) and one-sided spider vanes
// (blocked where max(|x'| - w/2, y') < 0, geometry.py:195-218, :463-485),
// multiplied and box-averaged over os x os sub-pixels.  support = any
// sub-pixel inside the primary.
struct SpiderSet {
  static constexpr int kMax = 16;
  int n;
  double cosa[kMax], sina[kMax];
};

__global__ void __launch_bounds__(256)
aperture_kernel(int npix, int os, double d_sub, double r_prim, double r_sec, double half_w,
                SpiderSet sp, float* __restrict__ trans, float* __restrict__ support) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)npix * npix) return;
  const int i = (int)(p / npix), j = (int)(p - (long)i * npix);
  const int nsub = npix * os;
  const double c0 = (nsub - 1) / 2.0;
  double acc = 0.0;
  int any = 0;
  for (int si = 0; si < os; ++si) {
    const double y = ((double)(i * os + si) - c0) * d_sub;
    for (int sj = 0; sj < os; ++sj) {
      const double x = ((double)(j * os + sj) - c0) * d_sub;
      const double r = hypot(x, y);
      const int in_prim = (r - r_prim) < 0.0;
      any |= in_prim;
      double t = in_prim ? 1.0 : 0.0;
      if (r_sec > 0.0) t *= (r - r_sec) > 0.0 ? 1.0 : 0.0;
      if (sp.n > 0) {
        int blocked = 0;
        for (int a = 0; a < sp.n; ++a) {
          const double xr = sp.cosa[a] * x - sp.sina[a] * y;
          const double yr = sp.sina[a] * x + sp.cosa[a] * y;
          blocked |= fmax(fabs(xr) - half_w, yr) < 0.0;
        }
        t *= blocked ? 0.0 : 1.0;
      }
      acc += t;
    }
  }
  trans[p] = (float)(acc / (double)(os * os));
  if (support) support[p] = any ? 1.f : 0.f;
}

// ---------------------------------------------------------------------------
// Dynamic soft apertures (layers/apertures.py:155-641, utils/geometry.py)
// ---------------------------------------------------------------------------

constexpr int kMaxApTransforms = 3;
constexpr int kMaxApEdges = 64;
constexpr int kMaxAps = 16;

struct TfDev {
  int flags;
  double tx, ty, sx, sy, cx, cy, rc, rs;   // rc, rs: cos(-r), sin(-r)
};

struct TfChain {
  int n;
  TfDev t[kMaxApTransforms];
};

// Pixel (i, j) of the grid through the chain.  Shear reads the transposed
// plane, so the chain is carried for (i, j) and (j, i) together: P and Q.
__device__ __forceinline__ void tf_point(const TfChain& ch, int i, int j, double c, double d,
                                         double h, double& x, double& y) {
  double px = c + (j - h) * d, py = c + (i - h) * d;   // pixel_coords, 'xy' order
  double qx = c + (i - h) * d, qy = c + (j - h) * d;   // the same at (j, i)
  for (int k = 0; k < ch.n; ++k) {
    const TfDev& t = ch.t[k];
    if (t.flags & DLX_TF_TRANSLATION) {
      px -= t.tx; py -= t.ty;
      qx -= t.tx; qy -= t.ty;
    }
    if (t.flags & DLX_TF_SHEAR) {
      const double npx = px + qx * t.sx, npy = py + qy * t.sy;
      const double nqx = qx + px * t.sx, nqy = qy + py * t.sy;
      px = npx; py = npy; qx = nqx; qy = nqy;
    }
    if (t.flags & DLX_TF_COMPRESSION) {
      px *= t.cx; py *= t.cy;
      qx *= t.cx; qy *= t.cy;
    }
    if (t.flags & DLX_TF_ROTATION) {
      const double npx = t.rc * px + t.rs * py, npy = -t.rs * px + t.rc * py;
      const double nqx = t.rc * qx + t.rs * qy, nqy = -t.rs * qx + t.rc * qy;
      px = npx; py = npy; qx = nqx; qy = nqy;
    }
  }
  x = px;
  y = py;
}

struct ApDev {
  int kind, occulting, nedges, nang;
  const double* xy;   // explicit coordinate planes (x then y), or NULL: the chain
  double s0, s1, clip;
  double ca[16], sa[16];                             // spider arm rotations
  double ea[kMaxApEdges], eb[kMaxApEdges], ec[kMaxApEdges];   // polygon edges a x + b y + c
  TfChain ch;
};

// Clipped soft distance (positive inside, negated when occulting) and the
// per-block minimum and maximum of it.
__global__ void __launch_bounds__(256)
soft_distance_kernel(ApDev ap, int npix, double dx, double c, double* __restrict__ dist,
                     double2* __restrict__ part) {
  __shared__ double smin[256], smax[256];
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long total = (long)npix * npix;
  double v = 0.0;
  const bool in = p < total;
  if (in) {
    const int i = (int)(p / npix), j = (int)(p - (long)i * npix);
    double x, y;
    if (ap.xy) {
      x = ap.xy[p];
      y = ap.xy[total + p];
    } else {
      tf_point(ap.ch, i, j, c, dx, (npix - 1) / 2.0, x, y);
    }
    double dist_edge;   // the reference's *_distance: negative inside
    if (ap.kind == DLX_AP_CIRCLE) {
      dist_edge = hypot(x, y) - ap.s0;
    } else if (ap.kind == DLX_AP_SQUARE) {
      dist_edge = fmax(fabs(x), fabs(y)) - ap.s0 / 2.0;
    } else if (ap.kind == DLX_AP_RECTANGLE) {
      dist_edge = fmax(fabs(x) - ap.s0 / 2.0, fabs(y) - ap.s1 / 2.0);
    } else if (ap.kind == DLX_AP_REG_POLYGON) {
      dist_edge = -INFINITY;
      for (int e = 0; e < ap.nedges; ++e)
        dist_edge = fmax(dist_edge, ap.ea[e] * x + ap.eb[e] * y + ap.ec[e]);
    } else {   // spider: -(min over arms of the arm distance) is positive on an arm
      double m = INFINITY;
      for (int a = 0; a < ap.nang; ++a) {
        const double xr = ap.ca[a] * x + ap.sa[a] * y;
        const double yr = -ap.sa[a] * x + ap.ca[a] * y;
        m = fmin(m, fmax(fabs(xr) - ap.s0 / 2.0, yr));
      }
      dist_edge = m;
    }
    v = -dist_edge;
    if (ap.occulting) v = -v;
    v = fmin(fmax(v, -ap.clip), ap.clip);
    dist[p] = v;
  }
  smin[threadIdx.x] = in ? v : INFINITY;
  smax[threadIdx.x] = in ? v : -INFINITY;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(smin[0], smax[0]);
}

struct ApGroups {
  int nap, nblk;
  int group[kMaxAps];
};

// soften (utils/geometry.py:49-92) of every aperture, combined as
// sum over groups of products.  Each block first reduces the per-block
// extrema of every aperture (a fixed order: deterministic).
__global__ void __launch_bounds__(256)
soft_combine_kernel(ApGroups g, long total, const double* __restrict__ dist,
                    const double2* __restrict__ part, float* __restrict__ out,
                    double* __restrict__ tplanes, double2* __restrict__ ext) {
  __shared__ double smin[256], smax[256];
  __shared__ double amin[kMaxAps], amax[kMaxAps];
  for (int a = 0; a < g.nap; ++a) {
    double mn = INFINITY, mx = -INFINITY;
    for (int b = threadIdx.x; b < g.nblk; b += 256) {
      const double2 q = part[(long)a * g.nblk + b];
      mn = fmin(mn, q.x);
      mx = fmax(mx, q.y);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
        smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      amin[a] = smin[0];
      amax[a] = smax[0];
      if (ext && blockIdx.x == 0) ext[a] = make_double2(smin[0], smax[0]);
    }
    __syncthreads();
  }
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= total) return;
  double sum = 0.0, prod = 1.0;
  for (int a = 0; a < g.nap; ++a) {
    const double d = dist[(long)a * total + p];
    const double mn = amin[a], mx = amax[a];
    const double t = mx == mn ? (d > 0.0 ? 1.0 : 0.0) : (d - mn) / (mx - mn);
    if (tplanes) tplanes[(long)a * total + p] = t;
    prod *= t;
    if (a + 1 == g.nap || g.group[a + 1] != g.group[a]) {
      sum += prod;
      prod = 1.0;
    }
  }
  if (out) out[p] = (float)sum;
}

// Zernike (nsides == 0) or n-sided polike basis on transformed coordinates
// divided by the aperture extent (utils/zernikes.py:257-294, :352-397).
__global__ void __launch_bounds__(256)
aperture_basis_kernel(ZernChunk z, int nmodes, int npix, double dx, double c, TfChain ch,
                      const double* __restrict__ xy, double inv_extent, int nsides,
                      float* __restrict__ out) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long total = (long)npix * npix;
  if (p >= total) return;
  const int i = (int)(p / npix), j = (int)(p - (long)i * npix);
  double x, y;
  if (xy) {
    x = xy[p];
    y = xy[total + p];
  } else {
    tf_point(ch, i, j, c, dx, (npix - 1) / 2.0, x, y);
  }
  x *= inv_extent;
  y *= inv_extent;
  double scale = 1.0;
  if (nsides >= 3) {   // 1/r_alpha * Z(coords / r_alpha)
    const double alpha = M_PI / nsides;
    const double phi = atan2(y, x) + alpha;
    const double wedge = floor((phi + alpha) / (2.0 * alpha));
    const double u = phi - wedge * (2.0 * alpha);
    const double r_alpha = cos(alpha) / cos(u);
    x /= r_alpha;
    y /= r_alpha;
    scale = 1.0 / r_alpha;
  }
  const double rho = hypot(x, y);
  const double ct = rho > 0.0 ? x / rho : 1.0, st = rho > 0.0 ? y / rho : 0.0;
  const double ap = rho <= 1.0 ? 1.0 : 0.0;
  for (int q = 0; q < nmodes; ++q)
    out[(long)q * total + p] = (float)(scale * ap * zern_eval(z, q, rho, ct, st));
}

// Resampling through a coordinate transformation: output pixel (i, j) of an
// n_out grid (spacing d_out, centre c_out) is sent through the transform chain
// and sampled from each n_in x n_in input image on its knot grid (spacing
// d_in, centre c_in): bilinear (method 1) or nearest knot (method 0); points
// outside the knots' span take `fill` (interpax.interp2d with extrap = fill,
// utils/interpolation.py:13-50).
__global__ void __launch_bounds__(256)
interp_kernel(TfChain ch, int n_out, double d_out, double c_out, int n_in, double d_in,
              double c_in, int method, float fill, int batch, const float* __restrict__ in,
              float* __restrict__ out) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long total = (long)n_out * n_out;
  if (p >= total) return;
  const int i = (int)(p / n_out), j = (int)(p - (long)i * n_out);
  double x, y;
  tf_point(ch, i, j, c_out, d_out, (n_out - 1) / 2.0, x, y);
  const double h = (n_in - 1) / 2.0;
  const double u = (x - c_in) / d_in + h, v = (y - c_in) / d_in + h;   // column, row
  const bool inside = u >= 0.0 && u <= n_in - 1 && v >= 0.0 && v <= n_in - 1;
  const long nn = (long)n_in * n_in;
  if (!inside) {
    for (int b = 0; b < batch; ++b) out[(long)b * total + p] = fill;
    return;
  }
  if (method == 0) {
    const int c = min((int)floor(u + 0.5), n_in - 1), r = min((int)floor(v + 0.5), n_in - 1);
    for (int b = 0; b < batch; ++b) out[(long)b * total + p] = in[(long)b * nn + (long)r * n_in + c];
    return;
  }
  const int c0 = min((int)floor(u), n_in > 1 ? n_in - 2 : 0);
  const int r0 = min((int)floor(v), n_in > 1 ? n_in - 2 : 0);
  const int c1 = min(c0 + 1, n_in - 1), r1 = min(r0 + 1, n_in - 1);
  const double tu = u - c0, tv = v - r0;
  for (int b = 0; b < batch; ++b) {
    const float* f = in + (long)b * nn;
    const double f00 = f[(long)r0 * n_in + c0], f01 = f[(long)r0 * n_in + c1];
    const double f10 = f[(long)r1 * n_in + c0], f11 = f[(long)r1 * n_in + c1];
    const double top = f00 + (f01 - f00) * tu, bot = f10 + (f11 - f10) * tu;
    out[(long)b * total + p] = (float)(top + (bot - top) * tv);
  }
}

// The same sampling at explicit points (qx, qy: npts float64 coordinates).
__global__ void __launch_bounds__(256)
interp_points_kernel(long npts, const double* __restrict__ qx, const double* __restrict__ qy,
                     int n_in, double d_in, double c_in, int method, float fill, int batch,
                     const float* __restrict__ in, float* __restrict__ out) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= npts) return;
  const double h = (n_in - 1) / 2.0;
  const double u = (qx[p] - c_in) / d_in + h, v = (qy[p] - c_in) / d_in + h;
  const bool inside = u >= 0.0 && u <= n_in - 1 && v >= 0.0 && v <= n_in - 1;
  const long nn = (long)n_in * n_in;
  for (int b = 0; b < batch; ++b) {
    const float* f = in + (long)b * nn;
    float r;
    if (!inside) {
      r = fill;
    } else if (method == 0) {
      r = f[(long)min((int)floor(v + 0.5), n_in - 1) * n_in + min((int)floor(u + 0.5), n_in - 1)];
    } else {
      const int c0 = min((int)floor(u), n_in > 1 ? n_in - 2 : 0);
      const int r0 = min((int)floor(v), n_in > 1 ? n_in - 2 : 0);
      const int c1 = min(c0 + 1, n_in - 1), r1 = min(r0 + 1, n_in - 1);
      const double tu = u - c0, tv = v - r0;
      const double f00 = f[(long)r0 * n_in + c0], f01 = f[(long)r0 * n_in + c1];
      const double f10 = f[(long)r1 * n_in + c0], f11 = f[(long)r1 * n_in + c1];
      const double top = f00 + (f01 - f00) * tu, bot = f10 + (f11 - f10) * tu;
      r = (float)(top + (bot - top) * tv);
    }
    out[(long)b * npts + p] = r;
  }
}

// ---- adjoint of the resampling (the images' cotangent) --------------------
// Each output point adds its cotangent times its interpolation weights into
// the knots it read: the transpose of the sampling matrix.  The sums go into
// an FP64 workspace by atomics and are rounded to float32 once, so the result
// does not depend on the order of the additions beyond FP64 rounding.  Points
// outside the knots' span read no knot (their value is the constant fill).
__global__ void __launch_bounds__(256) zero_f64_kernel(long n, double* __restrict__ a) {
  for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long)gridDim.x * 256)
    a[k] = 0.0;
}

__global__ void __launch_bounds__(256)
f64_to_f32_kernel(long n, const double* __restrict__ a, float* __restrict__ o) {
  for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long)gridDim.x * 256)
    o[k] = (float)a[k];
}

// qx == NULL: point p is output pixel p of the n_out grid through the chain.
__global__ void __launch_bounds__(256)
interp_vjp_kernel(TfChain ch, int n_out, double d_out, double c_out, const double* __restrict__ qx,
                  const double* __restrict__ qy, long npts, int n_in, double d_in, double c_in,
                  int method, int batch, const float* __restrict__ dout,
                  double* __restrict__ acc) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= npts) return;
  double x, y;
  if (qx) {
    x = qx[p];
    y = qy[p];
  } else {
    const int i = (int)(p / n_out), j = (int)(p - (long)i * n_out);
    tf_point(ch, i, j, c_out, d_out, (n_out - 1) / 2.0, x, y);
  }
  const double h = (n_in - 1) / 2.0;
  const double u = (x - c_in) / d_in + h, v = (y - c_in) / d_in + h;
  if (!(u >= 0.0 && u <= n_in - 1 && v >= 0.0 && v <= n_in - 1)) return;
  const long nn = (long)n_in * n_in;
  if (method == 0) {
    const long k = (long)min((int)floor(v + 0.5), n_in - 1) * n_in + min((int)floor(u + 0.5), n_in - 1);
    for (int b = 0; b < batch; ++b) atomicAdd(acc + (long)b * nn + k, (double)dout[(long)b * npts + p]);
    return;
  }
  const int c0 = min((int)floor(u), n_in > 1 ? n_in - 2 : 0);
  const int r0 = min((int)floor(v), n_in > 1 ? n_in - 2 : 0);
  const int c1 = min(c0 + 1, n_in - 1), r1 = min(r0 + 1, n_in - 1);
  const double tu = u - c0, tv = v - r0;
  const double w00 = (1.0 - tu) * (1.0 - tv), w01 = tu * (1.0 - tv);
  const double w10 = (1.0 - tu) * tv, w11 = tu * tv;
  for (int b = 0; b < batch; ++b) {
    const double g = dout[(long)b * npts + p];
    double* a = acc + (long)b * nn;
    atomicAdd(a + (long)r0 * n_in + c0, g * w00);
    atomicAdd(a + (long)r0 * n_in + c1, g * w01);
    atomicAdd(a + (long)r1 * n_in + c0, g * w10);
    atomicAdd(a + (long)r1 * n_in + c1, g * w11);
  }
}

// ---- geometry gradients of the soft apertures ----------------------------
// Parameters per aperture, in this order: size[0], size[1], then the last
// transform of its chain: translation x, y, shear x, y, compression x, y,
// rotation.  The per-pixel derivative of the clipped soft distance dc is
// carried forward through that transform and the edge distance; soften's
// extrema enter through JAX's rule for min/max (the derivative averaged
// over the tied pixels), so
//   dT/dth = (ddc - dmn)/R - (dc - mn)(dmx - dmn)/R^2,  R = mx - mn.
constexpr int kGeoParams = 9;
constexpr int kVjpSums = 3 * kGeoParams + 4;   // S1[9], Smn[9], Smx[9], S0, S2, cmn, cmx

// dist and its gradient for one pixel: returns the clipped soft distance dc
// and d(dc)/d(theta) in dd[9] (zero where clipped).
__device__ void soft_point_grad(const ApDev& ap, int i, int j, int npix, double dx, double c,
                                double& dc, double (&dd)[kGeoParams]) {
  const double h = (npix - 1) / 2.0;
  for (int k = 0; k < kGeoParams; ++k) dd[k] = 0.0;
  // every transform but the last, without derivatives
  TfChain pre = ap.ch;
  const int own = pre.n > 0 ? pre.n - 1 : -1;
  pre.n = own > 0 ? own : 0;
  double px = c + (j - h) * dx, py = c + (i - h) * dx;
  double qx = c + (i - h) * dx, qy = c + (j - h) * dx;
  {
    double x, y;
    tf_point(pre, i, j, c, dx, h, x, y);
    double xq, yq;
    tf_point(pre, j, i, c, dx, h, xq, yq);
    px = x; py = y; qx = xq; qy = yq;
  }
  // Jacobians of (x, y) with respect to the 9 parameters (sizes: 0)
  double jx[kGeoParams] = {0}, jy[kGeoParams] = {0};
  double jqx[kGeoParams] = {0}, jqy[kGeoParams] = {0};
  if (ap.xy) {   // explicit coordinates: no transform slots
    const long total = (long)npix * npix, p = (long)i * npix + j;
    px = ap.xy[p];
    py = ap.xy[total + p];
  } else if (own >= 0) {
    const TfDev& t = ap.ch.t[own];
    if (t.flags & DLX_TF_TRANSLATION) {
      px -= t.tx; py -= t.ty; qx -= t.tx; qy -= t.ty;
      jx[2] = -1.0; jy[3] = -1.0; jqx[2] = -1.0; jqy[3] = -1.0;
    }
    if (t.flags & DLX_TF_SHEAR) {
      const double npx = px + qx * t.sx, npy = py + qy * t.sy;
      const double nqx = qx + px * t.sx, nqy = qy + py * t.sy;
      double njx[kGeoParams], njy[kGeoParams], njqx[kGeoParams], njqy[kGeoParams];
      for (int k = 0; k < kGeoParams; ++k) {
        njx[k] = jx[k] + jqx[k] * t.sx;
        njy[k] = jy[k] + jqy[k] * t.sy;
        njqx[k] = jqx[k] + jx[k] * t.sx;
        njqy[k] = jqy[k] + jy[k] * t.sy;
      }
      njx[4] += qx; njy[5] += qy; njqx[4] += px; njqy[5] += py;
      for (int k = 0; k < kGeoParams; ++k) {
        jx[k] = njx[k]; jy[k] = njy[k]; jqx[k] = njqx[k]; jqy[k] = njqy[k];
      }
      px = npx; py = npy; qx = nqx; qy = nqy;
    }
    if (t.flags & DLX_TF_COMPRESSION) {
      for (int k = 0; k < kGeoParams; ++k) {
        jx[k] *= t.cx; jy[k] *= t.cy;
      }
      jx[6] += px; jy[7] += py;
      px *= t.cx; py *= t.cy;
    }
    if (t.flags & DLX_TF_ROTATION) {
      for (int k = 0; k < kGeoParams; ++k) {
        const double a = jx[k], b = jy[k];
        jx[k] = t.rc * a + t.rs * b;
        jy[k] = -t.rs * a + t.rc * b;
      }
      jx[8] += t.rs * px - t.rc * py;
      jy[8] += t.rc * px + t.rs * py;
      const double npx = t.rc * px + t.rs * py, npy = -t.rs * px + t.rc * py;
      px = npx; py = npy;
    }
  }
  const double x = px, y = py;
  // edge distance (negative inside) and its partials
  double dist, gx = 0.0, gy = 0.0, g0 = 0.0, g1 = 0.0;
  if (ap.kind == DLX_AP_CIRCLE) {
    const double r = hypot(x, y);
    dist = r - ap.s0;
    if (r > 0.0) { gx = x / r; gy = y / r; }
    g0 = -1.0;
  } else if (ap.kind == DLX_AP_SQUARE) {
    if (fabs(x) >= fabs(y)) { dist = fabs(x) - ap.s0 / 2.0; gx = x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }
    else { dist = fabs(y) - ap.s0 / 2.0; gy = y > 0 ? 1.0 : (y < 0 ? -1.0 : 0.0); }
    g0 = -0.5;
  } else if (ap.kind == DLX_AP_RECTANGLE) {
    const double ax = fabs(x) - ap.s0 / 2.0, ay = fabs(y) - ap.s1 / 2.0;
    if (ax >= ay) { dist = ax; gx = x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); g0 = -0.5; }
    else { dist = ay; gy = y > 0 ? 1.0 : (y < 0 ? -1.0 : 0.0); g1 = -0.5; }
  } else if (ap.kind == DLX_AP_REG_POLYGON) {
    int best = 0;
    dist = -INFINITY;
    for (int e = 0; e < ap.nedges; ++e) {
      const double v = ap.ea[e] * x + ap.eb[e] * y + ap.ec[e];
      if (v > dist) { dist = v; best = e; }
    }
    gx = ap.ea[best]; gy = ap.eb[best];
    g0 = ap.s0 != 0.0 ? ap.ec[best] / ap.s0 : 0.0;   // c scales with rmax
  } else {   // spider: min over arms of max(|x'| - w/2, y')
    dist = INFINITY;
    for (int a = 0; a < ap.nang; ++a) {
      const double xr = ap.ca[a] * x + ap.sa[a] * y, yr = -ap.sa[a] * x + ap.ca[a] * y;
      const double bx = fabs(xr) - ap.s0 / 2.0;
      const double v = fmax(bx, yr);
      if (v < dist) {
        dist = v;
        if (bx >= yr) {
          const double sg = xr > 0 ? 1.0 : (xr < 0 ? -1.0 : 0.0);
          gx = sg * ap.ca[a]; gy = sg * ap.sa[a]; g0 = -0.5;
        } else {
          gx = -ap.sa[a]; gy = ap.ca[a]; g0 = 0.0;
        }
      }
    }
  }
  const double sgn = ap.occulting ? 1.0 : -1.0;   // d = -dist, or +dist when occulting
  const double d = sgn * dist;
  dc = fmin(fmax(d, -ap.clip), ap.clip);
  if (d > -ap.clip && d < ap.clip) {
    dd[0] = sgn * g0;
    dd[1] = sgn * g1;
    for (int k = 2; k < kGeoParams; ++k) dd[k] = sgn * (gx * jx[k] + gy * jy[k]);
  }
}

// Per-block partial sums for one aperture a: with gT = dL/dT_a = dL/dout times
// the other members of its group,
//   S1[k] = sum gT ddc_k,  Smn[k] / Smx[k] = sum of ddc_k over pixels at the
//   minimum / maximum, S0 = sum gT, S2 = sum gT (dc - mn), cmn / cmx = counts.
__global__ void __launch_bounds__(256)
soft_vjp_kernel(ApDev ap, int a, ApGroups g, int npix, double dx, double c,
                const float* __restrict__ dout, const double* __restrict__ tplanes,
                const double2* __restrict__ ext, double* __restrict__ part) {
  __shared__ double red[kVjpSums][4];
  const long total = (long)npix * npix;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  double v[kVjpSums];
  for (int k = 0; k < kVjpSums; ++k) v[k] = 0.0;
  if (p < total) {
    const int i = (int)(p / npix), j = (int)(p - (long)i * npix);
    double dc, dd[kGeoParams];
    soft_point_grad(ap, i, j, npix, dx, c, dc, dd);
    const double mn = ext[a].x, mx = ext[a].y;
    double gt = dout[p];
    for (int b = 0; b < g.nap; ++b)
      if (b != a && g.group[b] == g.group[a]) gt *= tplanes[(long)b * total + p];
    for (int k = 0; k < kGeoParams; ++k) v[k] = gt * dd[k];
    if (dc == mn) {
      for (int k = 0; k < kGeoParams; ++k) v[kGeoParams + k] = dd[k];
      v[3 * kGeoParams + 2] = 1.0;
    }
    if (dc == mx) {
      for (int k = 0; k < kGeoParams; ++k) v[2 * kGeoParams + k] = dd[k];
      v[3 * kGeoParams + 3] = 1.0;
    }
    v[3 * kGeoParams] = gt;
    v[3 * kGeoParams + 1] = gt * (dc - mn);
  }
  // wave sums, then the 4 waves of the block
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  for (int k = 0; k < kVjpSums; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[k][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x < kVjpSums) {
    const int k = threadIdx.x;
    part[(long)blockIdx.x * kVjpSums + k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
  }
}

// One block per aperture: fixed-order sum of the block partials, then
//   dL/dth_k = S1_k / R - dmn_k S0 / R - (dmx_k - dmn_k) S2 / R^2
// with dmn_k = Smn_k / cmn, dmx_k = Smx_k / cmx (0 when the field is constant).
__global__ void __launch_bounds__(64)
soft_vjp_final_kernel(int nblk, const double* __restrict__ part, const double2* __restrict__ ext,
                      float* __restrict__ out) {
  const int a = blockIdx.x;
  const int k = threadIdx.x;
  const double* pa = part + (long)a * nblk * kVjpSums;
  __shared__ double tot[kVjpSums];
  if (k < kVjpSums) {
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += pa[(long)b * kVjpSums + k];
    tot[k] = acc;
  }
  __syncthreads();
  if (k >= kGeoParams) return;
  const double mn = ext[a].x, mx = ext[a].y, R = mx - mn;
  double gk = 0.0;
  if (R != 0.0) {
    const double cmn = tot[3 * kGeoParams + 2], cmx = tot[3 * kGeoParams + 3];
    const double dmn = cmn > 0 ? tot[kGeoParams + k] / cmn : 0.0;
    const double dmx = cmx > 0 ? tot[2 * kGeoParams + k] / cmx : 0.0;
    const double S0 = tot[3 * kGeoParams], S2 = tot[3 * kGeoParams + 1];
    gk = tot[k] / R - dmn * S0 / R - (dmx - dmn) * S2 / (R * R);
  }
  out[a * kGeoParams + k] = (float)gk;
}

void noll(int j, int& n, int& m) {
  n = (int)std::ceil(-0.5 + std::sqrt(1.0 + 8.0 * j) / 2.0) - 1;
  const double smallest = n * (n + 1) / 2.0 + 1.0;
  const double shifts = std::floor((j - smallest + (~(n & 1)) + 2) / 2.0);
  const int sign = -(j & 1) + (~(j & 1)) + 2;
  const int base = n & 1;
  m = (int)(sign * (base + shifts * 2));
}

// factorial via lgamma; +inf at negative integers (utils/math.py:137-157).
double fact(double x) {
  if (x == 0.0) return 1.0;
  if (x < 0 && x == std::floor(x)) return INFINITY;
  return std::exp(std::lgamma(x + 1.0));
}

// Radial coefficients, orders and normalisations of nq Noll modes
// (zernike_factors / eval_azimuthal, utils/zernikes.py:122-204).
int zern_chunk(const int32_t* js, int nq, ZernChunk& z) {
  for (int q = 0; q < nq; ++q) {
    const int j = js[q];
    DLX_CHECK_ARG(j >= 1, "The Zernike index must be greater than 0.");
    int n, m;
    noll(j, n, m);
    const int nt = (n - m) / 2 + 1;  // k = arange((n - m) // 2 + 1), signed m
    DLX_CHECK_ARG(nt <= kMaxTerms, "Zernike index %d is beyond the supported order", j);
    z.n[q] = n;
    z.m[q] = m;
    z.nterms[q] = nt;
    z.nan0[q] = 0;
    for (int k = 0; k < nt; ++k) {
      const double sign = (k & 1) ? -1.0 : 1.0;
      z.c[q][k] = sign * fact(std::fabs((double)(n - k))) / fact(k) /
                  fact((double)((n + m) / 2 - k)) / fact((double)((n - m) / 2 - k));
      if (z.c[q][k] == 0.0 && n - 2 * k < 0) z.nan0[q] = 1;
    }
    // the finite-coefficient terms are k = 0 .. (n - |m|)/2, powers rho^(n - 2k) >= rho^|m|
    z.kmax[q] = (n - (m < 0 ? -m : m)) / 2;
    for (int k = 0; k <= z.kmax[q]; ++k)
      DLX_CHECK_ARG(std::isfinite(z.c[q][k]) && z.c[q][k] != 0.0,
                    "Zernike index %d: unexpected radial coefficient", j);
    z.norm[q] = std::sqrt((double)n + 1.0) * (m != 0 ? std::sqrt(2.0) : 1.0);
  }
  return DLX_OK;
}

}  // namespace
}  // namespace dlx

using namespace dlx;

extern "C" int dlx_noll_indices(int32_t j, int32_t* n, int32_t* m) {
  clear_error();
  DLX_CHECK_ARG(j >= 1, "The Zernike index must be greater than 0.");
  DLX_CHECK_ARG(n && m, "null argument to dlx_noll_indices");
  int nn, mm;
  noll(j, nn, mm);
  *n = nn;
  *m = mm;
  return DLX_OK;
}

extern "C" int dlx_zernike_basis(int32_t nmodes, const int32_t* js, int64_t npix,
                                 const float* x, const float* y, double diameter, float* out,
                                 void* stream) {
  clear_error();
  DLX_CHECK_ARG(nmodes >= 0 && npix >= 0, "negative size to dlx_zernike_basis");
  if (nmodes == 0 || npix == 0) return DLX_OK;
  DLX_CHECK_ARG(js && x && y && out, "null argument to dlx_zernike_basis");
  DLX_CHECK_ARG(std::isfinite(diameter) && diameter != 0.0, "diameter must be non-zero");
  hipStream_t s = (hipStream_t)stream;
  for (int q0 = 0; q0 < nmodes; q0 += kModesPerLaunch) {
    const int nq = nmodes - q0 < kModesPerLaunch ? nmodes - q0 : kModesPerLaunch;
    ZernChunk z{};
    const int st = zern_chunk(js + q0, nq, z);
    if (st) return st;
    hipLaunchKernelGGL(zernike_kernel, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, z,
                       nq, npix, x, y, 2.0 / diameter, out + (long)q0 * npix);
    DLX_HIP(hipGetLastError());
  }
  return DLX_OK;
}

extern "C" int dlx_circular_aperture(int32_t npix, double diameter, int32_t oversample,
                                     double secondary_diameter, double array_diameter,
                                     double spider_width, int32_t n_spiders,
                                     const double* spider_angles_deg, float* transmission,
                                     float* support, void* stream) {
  clear_error();
  DLX_CHECK_ARG(npix > 0 && oversample > 0, "npix and oversample must be > 0");
  DLX_CHECK_ARG(transmission, "null transmission");
  DLX_CHECK_ARG(n_spiders >= 0 && n_spiders <= SpiderSet::kMax, "at most %d spider vanes",
                SpiderSet::kMax);
  DLX_CHECK_ARG(n_spiders == 0 || spider_angles_deg, "null spider angles");
  const double coord_diam = array_diameter > 0 ? array_diameter : diameter;
  SpiderSet sp{};
  sp.n = spider_width > 0 ? n_spiders : 0;
  // x' = cos(r) x - sin(r) y,  y' = sin(r) x + cos(r) y  (rotate_coords with r = deg2rad(a))
  for (int a = 0; a < sp.n; ++a) {
    const double rad = spider_angles_deg[a] * (M_PI / 180.0);
    sp.cosa[a] = std::cos(rad);
    sp.sina[a] = std::sin(rad);
  }
  const long total = (long)npix * npix;
  hipLaunchKernelGGL(aperture_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, npix, oversample,
                     coord_diam / ((double)npix * oversample), diameter / 2.0,
                     secondary_diameter > 0 ? secondary_diameter / 2.0 : 0.0, spider_width / 2.0,
                     sp, transmission, support);
  DLX_HIP(hipGetLastError());
  return DLX_OK;
}

namespace dlx {
namespace {

int build_chain(int n, const dlx_coord_transform* t, TfChain& ch) {
  DLX_CHECK_ARG(n >= 0 && n <= kMaxApTransforms, "at most %d coordinate transforms, got %d",
                kMaxApTransforms, n);
  DLX_CHECK_ARG(n == 0 || t, "null transforms");
  ch.n = n;
  for (int k = 0; k < n; ++k) {
    TfDev& d = ch.t[k];
    d.flags = t[k].flags;
    d.tx = t[k].translation[0];
    d.ty = t[k].translation[1];
    d.sx = t[k].shear[0];
    d.sy = t[k].shear[1];
    d.cx = t[k].compression[0];
    d.cy = t[k].compression[1];
    d.rc = std::cos(-t[k].rotation);
    d.rs = std::sin(-t[k].rotation);
  }
  return DLX_OK;
}

// The vertex angles of linspace(-pi, pi, n, endpoint=False): k * step + start,
// multiplied and added as two roundings.
double polygon_angle(int k, int n) {
  const double step = (M_PI - (-M_PI)) / n;
  volatile double t = k * step;
  return t + (-M_PI);
}

// Edges of the regular polygon as lines a x + b y + c (reg_polygon_edges and
// line_distance, utils/geometry.py:487-579): vertices at those angles on
// radius rmax, slope m = nandiv(dy, dx) (+inf where dx == 0, nandiv's
// default fill, utils/math.py:228-250): the vertical rule for m = inf, the
// horizontal rule for m == 0, the general line otherwise.
void polygon_edges(int n, double rmax, ApDev& a) {
  a.nedges = n;
  for (int e = 0; e < n; ++e) {
    const int f = (e + 1) % n;
    const double ti = polygon_angle(e, n), tj = polygon_angle(f, n);
    const double xi = rmax * std::cos(ti), yi = rmax * std::sin(ti);
    const double xj = rmax * std::cos(tj), yj = rmax * std::sin(tj);
    const double den = xj - xi;
    auto sgn = [](double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); };
    if (den == 0.0) {   // -sign(xi) * (xi - x)
      const double s = sgn(xi);
      a.ea[e] = s;
      a.eb[e] = 0.0;
      a.ec[e] = -s * xi;
      continue;
    }
    const double m = (yj - yi) / den;
    if (m == 0.0) {   // -sign(yi) * (yi - y)
      const double s = sgn(yi);
      a.ea[e] = 0.0;
      a.eb[e] = s;
      a.ec[e] = -s * yi;
    } else {          // -sign(yi - m xi) * (m x - y + yi - m xi) / sqrt(m^2 + 1)
      const double s = sgn(yi - m * xi), r = std::sqrt(m * m + 1.0);
      a.ea[e] = -s * m / r;
      a.eb[e] = s / r;
      a.ec[e] = -s * (yi - m * xi) / r;
    }
  }
}

}  // namespace
}  // namespace dlx

namespace dlx {
namespace {

// Device descriptors and groups of n soft apertures (validated).
int build_apertures(int npix, double pixel_scale, int n, const dlx_soft_aperture* aps,
                    std::vector<ApDev>& dev, ApGroups& g) {
  DLX_CHECK_ARG(npix > 0 && npix <= 46340, "npix must be in [1, 46340], got %d", npix);
  DLX_CHECK_ARG(n >= 1 && n <= kMaxAps, "between 1 and %d apertures, got %d", kMaxAps, n);
  DLX_CHECK_ARG(aps, "null aperture descriptors");
  DLX_CHECK_ARG(std::isfinite(pixel_scale) && pixel_scale > 0, "pixel_scale must be > 0");
  g = ApGroups{};
  g.nap = n;
  dev.assign(n, ApDev{});
  for (int a = 0; a < n; ++a) {
    const dlx_soft_aperture& s = aps[a];
    ApDev& d = dev[a];
    d.kind = s.kind;
    d.occulting = s.occulting != 0;
    d.s0 = s.size[0];
    d.s1 = s.size[1];
    d.xy = s.coords;
    DLX_CHECK_ARG(s.softening > 0, "softening must be greater than 0.");
    d.clip = pixel_scale * s.softening / 2.0;
    int st = build_chain(s.n_transforms, s.transforms, d.ch);
    if (st) return st;
    switch (s.kind) {
      case DLX_AP_CIRCLE:
      case DLX_AP_SQUARE:
      case DLX_AP_RECTANGLE:
        break;
      case DLX_AP_REG_POLYGON:
        DLX_CHECK_ARG(s.nsides >= 3 && s.nsides <= kMaxApEdges,
                      "a regular polygon needs 3 to %d sides, got %d", kMaxApEdges, s.nsides);
        polygon_edges(s.nsides, s.size[0], d);
        break;
      case DLX_AP_SPIDER:
        DLX_CHECK_ARG(s.n_angles >= 1 && s.n_angles <= 16, "a spider needs 1 to 16 arms, got %d",
                      s.n_angles);
        d.nang = s.n_angles;
        for (int k = 0; k < s.n_angles; ++k) {   // rotate_coords by deg2rad(angle)
          const double r = s.angles_deg[k] * (M_PI / 180.0);
          d.ca[k] = std::cos(-r);
          d.sa[k] = std::sin(-r);
        }
        break;
      default:
        DLX_CHECK_ARG(false, "unknown aperture kind %d", s.kind);
    }
    g.group[a] = s.group;
    // groups consecutive: a group index, once left, does not come back
    for (int b = 0; b + 1 < a; ++b)
      DLX_CHECK_ARG(!(aps[b].group == s.group && aps[a - 1].group != s.group),
                    "apertures of one group must be consecutive");
  }
  return DLX_OK;
}

struct SoftLayout {
  size_t dist, part, tplanes, ext, vpart, total;
};

SoftLayout soft_layout(int npix, int n, bool vjp) {
  SoftLayout L{};
  const long total = (long)npix * npix;
  const long nblk = (total + 255) / 256;
  L.dist = 0;
  L.part = align_up((size_t)n * total * sizeof(double));
  L.tplanes = L.part + align_up((size_t)n * nblk * sizeof(double2));
  L.total = L.tplanes;
  if (vjp) {
    L.ext = L.tplanes + align_up((size_t)n * total * sizeof(double));
    L.vpart = L.ext + align_up((size_t)n * sizeof(double2));
    L.total = L.vpart + align_up((size_t)n * nblk * kVjpSums * sizeof(double));
  }
  return L;
}

// The distance passes of every aperture, then the combination (and, for the
// VJP, the per-aperture transmissions and extrema).
int soft_forward(const std::vector<ApDev>& dev, const ApGroups& g, int npix, double pixel_scale,
                 double center, float* out, uint8_t* ws, const SoftLayout& L, bool vjp,
                 hipStream_t st) {
  const long total = (long)npix * npix;
  const int nblk = (int)((total + 255) / 256);
  double* dist = (double*)(ws + L.dist);
  double2* part = (double2*)(ws + L.part);
  for (int a = 0; a < g.nap; ++a) {
    hipLaunchKernelGGL(soft_distance_kernel, dim3(nblk), dim3(256), 0, st, dev[a], npix,
                       pixel_scale, center, dist + (long)a * total, part + (long)a * nblk);
    DLX_HIP(hipGetLastError());
  }
  ApGroups gg = g;
  gg.nblk = nblk;
  hipLaunchKernelGGL(soft_combine_kernel, dim3(nblk), dim3(256), 0, st, gg, total, dist, part,
                     out, vjp ? (double*)(ws + L.tplanes) : (double*)nullptr,
                     vjp ? (double2*)(ws + L.ext) : (double2*)nullptr);
  DLX_HIP(hipGetLastError());
  return DLX_OK;
}

}  // namespace
}  // namespace dlx

extern "C" size_t dlx_soft_apertures_workspace_size(int32_t npix, int32_t n) {
  if (npix <= 0 || n <= 0) return 0;
  return soft_layout(npix, n, false).total;
}

extern "C" int dlx_soft_apertures(int32_t npix, double pixel_scale, double center, int32_t n,
                                 const dlx_soft_aperture* aps, float* out, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  clear_error();
  std::vector<ApDev> dev;
  ApGroups g;
  int st = build_apertures(npix, pixel_scale, n, aps, dev, g);
  if (st) return st;
  DLX_CHECK_ARG(out, "null argument to dlx_soft_apertures");
  const SoftLayout L = soft_layout(npix, n, false);
  if (workspace_bytes < L.total || !workspace) {
    set_error("dlx_soft_apertures workspace too small: need %zu bytes, got %zu", L.total,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  return soft_forward(dev, g, npix, pixel_scale, center, out, (uint8_t*)workspace, L, false,
                      (hipStream_t)stream);
}

extern "C" size_t dlx_soft_apertures_vjp_workspace_size(int32_t npix, int32_t n) {
  if (npix <= 0 || n <= 0) return 0;
  return soft_layout(npix, n, true).total;
}

extern "C" int dlx_soft_apertures_vjp(int32_t npix, double pixel_scale, double center,
                                     int32_t n, const dlx_soft_aperture* aps, const float* dout,
                                     float* grads, void* workspace, size_t workspace_bytes,
                                     void* stream) {
  clear_error();
  std::vector<ApDev> dev;
  ApGroups g;
  int st = build_apertures(npix, pixel_scale, n, aps, dev, g);
  if (st) return st;
  DLX_CHECK_ARG(dout && grads, "null argument to dlx_soft_apertures_vjp");
  const SoftLayout L = soft_layout(npix, n, true);
  if (workspace_bytes < L.total || !workspace) {
    set_error("dlx_soft_apertures_vjp workspace too small: need %zu bytes, got %zu", L.total,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  uint8_t* ws = (uint8_t*)workspace;
  hipStream_t s = (hipStream_t)stream;
  if ((st = soft_forward(dev, g, npix, pixel_scale, center, nullptr, ws, L, true, s))) return st;
  const long total = (long)npix * npix;
  const int nblk = (int)((total + 255) / 256);
  ApGroups gg = g;
  gg.nblk = nblk;
  double* vpart = (double*)(ws + L.vpart);
  for (int a = 0; a < n; ++a) {
    hipLaunchKernelGGL(soft_vjp_kernel, dim3(nblk), dim3(256), 0, s, dev[a], a, gg, npix,
                       pixel_scale, center, dout, (const double*)(ws + L.tplanes),
                       (const double2*)(ws + L.ext), vpart + (long)a * nblk * kVjpSums);
    DLX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(soft_vjp_final_kernel, dim3(n), dim3(64), 0, s, nblk, vpart,
                     (const double2*)(ws + L.ext), grads);
  DLX_HIP(hipGetLastError());
  return DLX_OK;
}

extern "C" int dlx_aperture_basis(int32_t npix, double pixel_scale, double center,
                                  int32_t n_transforms, const dlx_coord_transform* transforms,
                                  const double* coords, double extent, int32_t nsides,
                                  int32_t nmodes, const int32_t* js, float* out,
                                  void* stream) {
  clear_error();
  DLX_CHECK_ARG(npix > 0 && npix <= 46340, "npix must be in [1, 46340], got %d", npix);
  DLX_CHECK_ARG(nmodes >= 0, "negative mode count");
  if (nmodes == 0) return DLX_OK;
  DLX_CHECK_ARG(js && out, "null argument to dlx_aperture_basis");
  DLX_CHECK_ARG(std::isfinite(extent) && extent != 0.0, "extent must be non-zero");
  DLX_CHECK_ARG(nsides == 0 || nsides >= 3, "nsides must be zero or greater than two.");
  TfChain ch{};
  int st = build_chain(n_transforms, transforms, ch);
  if (st) return st;
  const long total = (long)npix * npix;
  hipStream_t s = (hipStream_t)stream;
  for (int q0 = 0; q0 < nmodes; q0 += kModesPerLaunch) {
    const int nq = nmodes - q0 < kModesPerLaunch ? nmodes - q0 : kModesPerLaunch;
    ZernChunk z{};
    st = zern_chunk(js + q0, nq, z);
    if (st) return st;
    hipLaunchKernelGGL(aperture_basis_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       s, z, nq, npix, pixel_scale, center, ch, coords, 1.0 / extent, nsides,
                       out + (long)q0 * total);
    DLX_HIP(hipGetLastError());
  }
  return DLX_OK;
}

extern "C" int dlx_interp_transform(int32_t n_in, double pixel_scale_in, double center_in,
                                    int32_t n_out, double pixel_scale_out, double center_out,
                                    int32_t n_transforms, const dlx_coord_transform* transforms,
                                    int32_t method, float fill, int32_t batch, const float* in,
                                    float* out, void* stream) {
  clear_error();
  DLX_CHECK_ARG(n_in > 0 && n_in <= 46340 && n_out > 0 && n_out <= 46340,
                "n_in and n_out must be in [1, 46340], got %d and %d", n_in, n_out);
  DLX_CHECK_ARG(batch >= 0 && batch <= 65535, "batch must be in [0, 65535], got %d", batch);
  DLX_CHECK_ARG(method == 0 || method == 1, "method must be 0 (nearest) or 1 (linear), got %d",
                method);
  DLX_CHECK_ARG(std::isfinite(pixel_scale_in) && pixel_scale_in > 0 &&
                    std::isfinite(pixel_scale_out) && pixel_scale_out > 0,
                "pixel scales must be > 0");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(in && out, "null argument to dlx_interp_transform");
  TfChain ch{};
  if (int st = build_chain(n_transforms, transforms, ch)) return st;
  const long total = (long)n_out * n_out;
  hipLaunchKernelGGL(interp_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, ch, n_out, pixel_scale_out, center_out, n_in,
                     pixel_scale_in, center_in, method, fill, batch, in, out);
  DLX_HIP(hipGetLastError());
  return DLX_OK;
}

extern "C" int dlx_interp_points(int32_t n_in, double pixel_scale_in, double center_in,
                                 int64_t npts, const double* qx, const double* qy,
                                 int32_t method, float fill, int32_t batch, const float* in,
                                 float* out, void* stream) {
  clear_error();
  DLX_CHECK_ARG(n_in > 0 && n_in <= 46340, "n_in must be in [1, 46340], got %d", n_in);
  DLX_CHECK_ARG(npts >= 0 && batch >= 0 && batch <= 65535, "bad point count or batch");
  DLX_CHECK_ARG(method == 0 || method == 1, "method must be 0 (nearest) or 1 (linear), got %d",
                method);
  DLX_CHECK_ARG(std::isfinite(pixel_scale_in) && pixel_scale_in > 0, "pixel scale must be > 0");
  if (npts == 0 || batch == 0) return DLX_OK;
  DLX_CHECK_ARG(qx && qy && in && out, "null argument to dlx_interp_points");
  hipLaunchKernelGGL(interp_points_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (long)npts, qx, qy, n_in, pixel_scale_in, center_in,
                     method, fill, batch, in, out);
  DLX_HIP(hipGetLastError());
  return DLX_OK;
}

namespace dlx {
namespace {
int interp_vjp_impl(const TfChain& ch, int n_out, double d_out, double c_out, const double* qx,
                    const double* qy, long npts, int n_in, double d_in, double c_in, int method,
                    int batch, const float* dout, float* din, double* ws, hipStream_t s) {
  const long n = (long)batch * n_in * n_in;
  const unsigned gz = (unsigned)std::min<long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(zero_f64_kernel, dim3(gz), dim3(256), 0, s, n, ws);
  DLX_HIP(hipGetLastError());
  if (npts > 0) {
    hipLaunchKernelGGL(interp_vjp_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s,
                       ch, n_out, d_out, c_out, qx, qy, npts, n_in, d_in, c_in, method, batch,
                       dout, ws);
    DLX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3(gz), dim3(256), 0, s, n, (const double*)ws, din);
  DLX_HIP(hipGetLastError());
  return DLX_OK;
}
}  // namespace
}  // namespace dlx

extern "C" int dlx_interp_transform_vjp(int32_t n_in, double pixel_scale_in, double center_in,
                                        int32_t n_out, double pixel_scale_out,
                                        double center_out, int32_t n_transforms,
                                        const dlx_coord_transform* transforms, int32_t method,
                                        int32_t batch, const float* dout, float* din,
                                        double* workspace, void* stream) {
  clear_error();
  DLX_CHECK_ARG(n_in > 0 && n_in <= 46340 && n_out > 0 && n_out <= 46340,
                "n_in and n_out must be in [1, 46340], got %d and %d", n_in, n_out);
  DLX_CHECK_ARG(batch >= 0 && batch <= 65535, "batch must be in [0, 65535], got %d", batch);
  DLX_CHECK_ARG(method == 0 || method == 1, "method must be 0 (nearest) or 1 (linear), got %d",
                method);
  DLX_CHECK_ARG(std::isfinite(pixel_scale_in) && pixel_scale_in > 0 &&
                    std::isfinite(pixel_scale_out) && pixel_scale_out > 0,
                "pixel scales must be > 0");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(dout && din && workspace, "null argument to dlx_interp_transform_vjp");
  TfChain ch{};
  if (int st = build_chain(n_transforms, transforms, ch)) return st;
  return interp_vjp_impl(ch, n_out, pixel_scale_out, center_out, nullptr, nullptr,
                         (long)n_out * n_out, n_in, pixel_scale_in, center_in, method, batch,
                         dout, din, workspace, (hipStream_t)stream);
}

extern "C" int dlx_interp_points_vjp(int32_t n_in, double pixel_scale_in, double center_in,
                                     int64_t npts, const double* qx, const double* qy,
                                     int32_t method, int32_t batch, const float* dout,
                                     float* din, double* workspace, void* stream) {
  clear_error();
  DLX_CHECK_ARG(n_in > 0 && n_in <= 46340, "n_in must be in [1, 46340], got %d", n_in);
  DLX_CHECK_ARG(npts >= 0 && batch >= 0 && batch <= 65535, "bad point count or batch");
  DLX_CHECK_ARG(method == 0 || method == 1, "method must be 0 (nearest) or 1 (linear), got %d",
                method);
  DLX_CHECK_ARG(std::isfinite(pixel_scale_in) && pixel_scale_in > 0, "pixel scale must be > 0");
  if (batch == 0) return DLX_OK;
  DLX_CHECK_ARG(din && workspace && (npts == 0 || (qx && qy && dout)),
                "null argument to dlx_interp_points_vjp");
  return interp_vjp_impl(TfChain{}, 1, 1.0, 0.0, qx, qy, (long)npts, n_in, pixel_scale_in,
                         center_in, method, batch, dout, din, workspace, (hipStream_t)stream);
}
