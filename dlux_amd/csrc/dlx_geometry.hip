// Device builders for the wavelength-independent inputs of the hot path:
// Noll-indexed Zernike bases and static (oversampled, box-averaged) circular
// apertures with a secondary obscuration and spider vanes.  One lane per
// output pixel; FP64 arithmetic (these run once per optical system and feed
// resident device buffers).
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
};

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
  const double rho = hypot(x, y), theta = atan2(y, x);
  const double ap = rho <= 1.0 ? 1.0 : 0.0;
  for (int q = 0; q < nmodes; ++q) {
    double radial = 0.0;
    for (int k = 0; k < z.nterms[q]; ++k) radial += z.c[q][k] * pow(rho, (double)(z.n[q] - 2 * k));
    const int am = z.m[q] < 0 ? -z.m[q] : z.m[q];
    const double trig = z.m[q] >= 0 ? cos(am * theta) : sin(am * theta);
    out[(long)q * npix + p] = (float)(ap * radial * (z.norm[q] * trig));
  }
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
    for (int q = 0; q < nq; ++q) {
      const int j = js[q0 + q];
      DLX_CHECK_ARG(j >= 1, "The Zernike index must be greater than 0.");
      int n, m;
      noll(j, n, m);
      const int nt = (n - m) / 2 + 1;  // k = arange((n - m) // 2 + 1), signed m
      DLX_CHECK_ARG(nt <= kMaxTerms, "Zernike index %d is beyond the supported order", j);
      z.n[q] = n;
      z.m[q] = m;
      z.nterms[q] = nt;
      for (int k = 0; k < nt; ++k) {
        const double sign = (k & 1) ? -1.0 : 1.0;
        z.c[q][k] = sign * fact(std::fabs((double)(n - k))) / fact(k) /
                    fact((double)((n + m) / 2 - k)) / fact((double)((n - m) / 2 - k));
      }
      z.norm[q] = std::sqrt((double)n + 1.0) * (m != 0 ? std::sqrt(2.0) : 1.0);
    }
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
