/*
 * dlux_amd.h -- C ABI of the MI355X (gfx950) diffraction hot path.
 *
 * The reference (dLux 0.16.0) is a pure-Python JAX library with no FFI; each
 * entry point below stands in for the reference routine cited beside it, and
 * is what a jax.ffi handler, a torch binding or a ctypes stub binds.
 *
 * Conventions
 *   - Every entry point returns an int status (DLX_OK == 0).  No exceptions
 *     cross this boundary; dlx_last_error() returns a per-thread message.
 *   - All array pointers are DEVICE pointers owned by the caller.  Complex
 *     arrays are interleaved complex64 (re, im float pairs), row-major, the
 *     layout of jax/torch complex64.  Real arrays are float32 unless noted.
 *   - Per-instance physical parameters are passed as HOST double arrays.
 *   - Work is enqueued on `stream` (a hipStream_t, NULL = default stream) and
 *     is stream-ordered; entry points do not synchronise.
 *   - Entry points that need scratch take (workspace, workspace_bytes); the
 *     matching *_workspace_size() query returns the bytes needed.  The library
 *     never frees caller memory.
 *   - Pupil grids are square N_in x N_in, axis 0 = y, axis 1 = x (the
 *     reference's 'xy' meshgrid order).
 */
#ifndef DLUX_AMD_H
#define DLUX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum dlx_status {
  DLX_OK = 0,
  DLX_ERR_INVALID = 1,     /* bad argument (shape, size, null pointer)        */
  DLX_ERR_HIP = 2,         /* a HIP runtime call failed                       */
  DLX_ERR_WORKSPACE = 3,   /* workspace too small                             */
  DLX_ERR_UNSUPPORTED = 4  /* valid request the library does not implement    */
};

enum dlx_precision {
  DLX_PREC_FP32 = 0,   /* complex64 operands, FP32 MFMA, FP32 accumulate         */
  DLX_PREC_BF16X3 = 1  /* split-BF16 3-term (hi*hi + hi*lo + lo*hi) MFMA, FP32 acc */
};

/* Library version string ("dlux_amd x.y.z gfx950"). */
const char* dlx_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* dlx_last_error(void);

/*
 * MFT geometry shared by a batch of instances (one per wavelength).
 * Mirrors the arguments of dLux's transfer_matrix / calc_nfringes / MFT:
 *   utils/propagation.py:67-77, :130-138, :178-188.
 */
typedef struct dlx_mft_geom {
  int32_t n_in;            /* pupil side length N_in                          */
  int32_t n_out;           /* output side length N_out                        */
  int32_t inverse;         /* 0: exp(-2 pi i ...), 1: exp(+2 pi i ...)         */
  int32_t has_focal;       /* 0: angular (focal_length None), 1: Cartesian     */
  double focal_length;     /* metres, used when has_focal                      */
  double focal_shift;      /* metres, added to focal_length (reference: 0)     */
} dlx_mft_geom;

/*
 * Transfer (DFT) matrices for `batch` instances, written TRANSPOSED relative
 * to the reference: out[b] is N_out x N_in with out[b][l][k] = T_b[k][l],
 *   T_b = exp(-/+ 2 pi i * outer(a, b)),  a_k = (k - (N_in-1)/2 - s_b)/N_in,
 *   b_l = (l - (N_out-1)/2 - s_b) * p_out_b * N_in * dx_b / lambda_b [/(f+df)].
 * The phase argument a_k*b_l is formed in FP64 and reduced to [-1/2, 1/2)
 * turns before an FP32 sincos.
 * Replaces: transfer_matrix, utils/propagation.py:67-127.
 */
int dlx_transfer_matrix(const dlx_mft_geom* geom, int32_t batch,
                        const double* wavelength, const double* pixel_scale_in,
                        const double* pixel_scale_out, const double* shift,
                        void* out, void* stream);

/*
 * Batched two-sided MFT of `batch` complex64 fields (N_in x N_in each) to
 * N_out x N_out: out[b] = c_b * (Y_b^T in[b]) X_b, with X_b = T(shift_x[b]),
 * Y_b = T(shift_y[b]) and c_b = nfringes_b / (N_in N_out).  Shifts are in
 * output pixels (the reference's pixel=True form; the host layer converts
 * pixel=False shifts).  shift_x / shift_y may be NULL (zero).
 * field_scale (host, may be NULL) multiplies out[b] by a real factor
 * (used for sqrt(weight) in polychromatic propagation).
 * Replaces: MFT, utils/propagation.py:178-256 (vmapped over instances as in
 * optical_systems.py:213-216).
 */
size_t dlx_mft_workspace_size(const dlx_mft_geom* geom, int32_t batch,
                              int32_t precision);
int dlx_mft(const dlx_mft_geom* geom, int32_t batch, const void* in, void* out,
            const double* wavelength, const double* pixel_scale_in,
            const double* pixel_scale_out, const double* shift_x,
            const double* shift_y, const double* field_scale, int32_t precision,
            void* workspace, size_t workspace_bytes, void* stream);

/*
 * Pupil field synthesis for a batch of wavelengths, in one pass:
 *   out[b](y,x) = s * amp(y,x) * exp(i*(2 pi/lambda_b * (opd(y,x) + tx*X + ty*Y)
 *                                         + phase(y,x)))
 * with X, Y the pixel-centre coordinates c + (j - (N-1)/2) * dx (metres).
 * amp / opd / phase are float32 N x N and may each be NULL (1 / 0 / 0).
 * If `in` is non-NULL the result multiplies in[b] (complex64, batch x N x N);
 * otherwise it multiplies the reference's initial phasor 1/N^2.
 * Replaces: Wavefront.__init__ phasor (wavefronts.py:111-113), tilt
 * (:370-395), add_opd / add_phase (:332-368), TransmissiveLayer
 * (layers/optical_layers.py:181-186).
 */
int dlx_pupil_field(int32_t n, int32_t batch, const double* wavelength,
                    double pixel_scale, double center, double tilt_x,
                    double tilt_y, const float* amp, const float* opd,
                    const float* phase, double scale, const void* in, void* out,
                    void* stream);

/*
 * OPD = sum_j coeffs[j] * basis[j] (+ base if non-NULL), float32 N x N, for
 * `nsets` coefficient sets at once (coeffs is nsets x nbasis, device f32;
 * out is nsets x npix).  Reads the basis once for all sets.
 * Replaces: eval_basis, utils/math.py:177-196 (BasisLayer.eval_basis,
 * layers/optical_layers.py:304-313).
 */
int dlx_eval_basis(int32_t nbasis, int64_t npix, int32_t nsets,
                   const float* basis, const float* coeffs, const float* base,
                   float* out, void* stream);

/*
 * Sum over pixels of |E|^2 per instance (out: batch floats, device).
 * Replaces: Wavefront.power, wavefronts.py:320-330 (and normalise's
 * reduction, :397-424).
 */
int dlx_power(int64_t npix, int32_t batch, const void* field, float* out,
              void* stream);

/*
 * Intensity: out[b] = |field[b]|^2 (sum_over_batch == 0) or
 * out = sum_b |field[b]|^2 (sum_over_batch == 1), float32.
 * Replaces: Wavefront.psf (wavefronts.py:268-279) and the wavelength sum of
 * OpticalSystem.propagate (optical_systems.py:222-223).
 */
int dlx_intensity(int64_t npix, int32_t batch, const void* field,
                  int32_t sum_over_batch, float* out, void* stream);

/*
 * Fused polychromatic PSF: for each of `nwl` wavelengths,
 *   P_l = amp * exp(i*(2 pi/lambda_l*(opd + tilt) + phase)) * s
 *   E_l = sqrt(w_l) * c_l * Y_l^T P_l X_l
 * and psf = sum_l |E_l|^2 (float32 N_out x N_out).  amp/opd/phase as in
 * dlx_pupil_field (NULL allowed); `scale` is the lambda-independent pupil
 * scale s (initial 1/N^2 times any power normalisation).  In the split-BF16x3
 * form with no phase, the pupil is generated inside the first contraction
 * from amp and opd where its tiles fill the GPU and amp / opd are 16-byte
 * aligned (no pupil array is written); the result is the
 * same bits either way (environment DLUX_AMD_PUPIL_PROLOGUE=0 forces the
 * two-kernel form).
 * Replaces the reference chain OpticalSystem.propagate -> propagate_mono ->
 * layers -> to_focus -> MFT -> psf.sum(0): optical_systems.py:147-223,
 * :391-425, :560-594, :659-680.
 */
size_t dlx_psf_poly_workspace_size(const dlx_mft_geom* geom, int32_t nwl,
                                   int32_t precision);
int dlx_psf_poly(const dlx_mft_geom* geom, int32_t nwl, const double* wavelength,
                 const double* weight, double pixel_scale_in,
                 double pixel_scale_out, double center, double tilt_x,
                 double tilt_y, const float* amp, const float* opd,
                 const float* phase, double scale, float* psf, int32_t precision,
                 void* workspace, size_t workspace_bytes, void* stream);

/*
 * dlx_psf_poly that allocates nothing and takes no workspace: the pupils,
 * transfer matrices and per-wavelength planes live in a fixed 64 MiB device
 * array of the library (one copy per GPU, held until the process ends; its
 * size is dlx_psf_poly_noalloc_scratch_bytes()), the wavelengths run in
 * chunks that fit it, and the transfer matrices are generated every call (the
 * device matrix cache, which allocates, is not used).  Same arguments and
 * result as dlx_psf_poly.  Rejects, with DLX_ERR_INVALID, a geometry whose
 * one-wavelength buffers exceed the array.  Calls that overlap on one GPU
 * (two host threads or two streams) race on the array; callers serialise them.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).
 */
size_t dlx_psf_poly_noalloc_scratch_bytes(void);
int dlx_psf_poly_noalloc(const dlx_mft_geom* geom, int32_t nwl, const double* wavelength,
                         const double* weight, double pixel_scale_in,
                         double pixel_scale_out, double center, double tilt_x,
                         double tilt_y, const float* amp, const float* opd,
                         const float* phase, double scale, float* psf,
                         int32_t precision, void* stream);

/*
 * Vector-Jacobian product of dlx_psf_poly: given dpsf = dL/dpsf (float32
 * N_out x N_out), writes
 *   grad_opd[p]   = sum_l (2 pi/lambda_l) Im(conj(P_l[p]) Pbar_l[p])
 *   grad_phase[p] = sum_l Im(conj(P_l[p]) Pbar_l[p])
 *   grad_amp[p]   = sum_l Re(conj(U_l[p]) Pbar_l[p]),  U_l = P_l / amp (P_l at unit
 *                   amplitude, so defined where amp[p] == 0 too)
 * with Pbar_l = sqrt(w_l) c_l conj(Y_l) (2 dpsf * E_l) X_l^H, the conjugate-
 * transpose contractions of the forward MFT (any output may be NULL, not all).
 * grad_amp is dL/d amp at fixed scale; the cotangent of a power normalisation
 * folded into amp is left to the caller's chain rule.  The per-wavelength
 * fields are recomputed, not stored: the work is four contractions per
 * wavelength (and one more pupil build for grad_amp).  Arguments as dlx_psf_poly.
 * Replaces the reverse-mode derivative JAX builds of the same chain
 * (jax.grad through OpticalSystem.propagate; MFT at utils/propagation.py:178-256,
 * phase/OPD application at wavefronts.py:332-367, |E|^2 at wavefronts.py:268-283).
 */
size_t dlx_psf_poly_vjp_workspace_size(const dlx_mft_geom* geom, int32_t nwl,
                                       int32_t precision);
int dlx_psf_poly_vjp(const dlx_mft_geom* geom, int32_t nwl, const double* wavelength,
                     const double* weight, double pixel_scale_in,
                     double pixel_scale_out, double center, double tilt_x,
                     double tilt_y, const float* amp, const float* opd,
                     const float* phase, double scale, const float* dpsf,
                     float* grad_opd, float* grad_phase, float* grad_amp, int32_t precision,
                     void* workspace, size_t workspace_bytes, void* stream);

/*
 * Transpose of dlx_eval_basis: out[s][j] = sum_p basis[j][p] * grad[s][p]
 * (nsets x nbasis float32), summed in a fixed order (deterministic).
 * Replaces the reverse-mode derivative of eval_basis (utils/math.py:177-196,
 * jnp.tensordot) that jax.grad builds for BasisLayer coefficients
 * (layers/optical_layers.py:304-327).
 */
size_t dlx_project_basis_workspace_size(int32_t nbasis, int64_t npix, int32_t nsets);
int dlx_project_basis(int32_t nbasis, int64_t npix, int32_t nsets, const float* basis,
                      const float* grad, float* out, void* workspace,
                      size_t workspace_bytes, void* stream);

/*
 * One instance of an uneven batched MFT: a complex64 n_in x n_in field
 * (device) propagated at one wavelength with its own pixel scales and output
 * shift (output pixels, per axis), whose intensity times `weight` is added
 * into output image `out_index`.  Geometry flags as in dlx_mft_geom.
 */
typedef struct dlx_mft_instance {
  const void* field;
  int32_t n_in, n_out;
  int32_t out_index;
  int32_t inverse;
  int32_t has_focal;
  int32_t reserved_;
  double focal_length, focal_shift;
  double wavelength, pixel_scale_in, pixel_scale_out;
  double shift_x, shift_y;
  double weight;
} dlx_mft_instance;

/*
 * Uneven batched polychromatic PSFs: for every instance i,
 *   I_i = weight_i * |c_i Y_i^T field_i X_i|^2   (X_i = T(shift_x), Y_i = T(shift_y))
 * and outputs[o] = sum of I_i over the instances with out_index == o
 * (float32 n_out x n_out, all its instances share n_out; an output with no
 * instance is not written).  Shapes may differ between instances; all
 * instances run in grouped persistent launches (split-BF16x3 contractions).
 * Instances that share a field pointer share its conversion.  `instances`
 * and `outputs` are HOST arrays (outputs holds device pointers).  Cannot be
 * captured into a caller's graph (it uploads per-call descriptors); instead,
 * when enabled, the launches themselves run as a HIP graph that the library captures (on a
 * private stream) at the second call with the same workspace, batch shape
 * (sizes, output grouping, shared fields) and matrix state (generated or
 * resident), and launches on `stream` after uploading that call's
 * descriptors (dlx_batched_graphs).
 * Replaces the vmapped source x wavelength propagation of
 * PointSources.model (sources.py:392-411) and OpticalSystem.propagate
 * (optical_systems.py:211-223) with MFT shifts (utils/propagation.py:178-256).
 */
size_t dlx_mft_batched_workspace_size(int32_t n, const dlx_mft_instance* instances,
                                      int32_t n_outputs);
int dlx_mft_batched_psf(int32_t n, const dlx_mft_instance* instances, int32_t n_outputs,
                        float* const* outputs, void* workspace, size_t workspace_bytes,
                        void* stream);

/*
 * Graph replay of dlx_mft_batched_psf's launches: enable = 1 on, 0 off (and
 * every captured graph destroyed after its last launch ran), -1 query.
 * Returns the previous setting.  Default off (environment
 * DLUX_AMD_BATCHED_GRAPHS=1 turns it on): measured on ROCm 7.2, a launch of
 * the captured graph cost more host time than the direct launches it replaces.  At most 64 graphs are kept,
 * least recently used dropped first.  No counterpart in the reference (its
 * launches are XLA's); for serving loops that repeat batch shapes.
 */
int dlx_batched_graphs(int32_t enable);
/* Graphs captured and graph launches since the process started, and graphs held. */
int dlx_batched_graph_stats(uint64_t* captures, uint64_t* launches, int32_t* entries);

/*
 * Noll index j -> radial order n and signed azimuthal order m.
 * Replaces: noll_indices, utils/zernikes.py:99-119.
 */
int dlx_noll_indices(int32_t j, int32_t* n, int32_t* m);

/*
 * Zernike basis on arbitrary coordinates: out[q][p] = Z_{js[q]}(x[p], y[p])
 * on the disk of the given diameter (js: HOST int array; x, y, out: device
 * float32, npix points each; out is nmodes x npix).
 * Replaces: zernike_basis / zernike, utils/zernikes.py:227-315.
 */
int dlx_zernike_basis(int32_t nmodes, const int32_t* js, int64_t npix, const float* x,
                      const float* y, double diameter, float* out, void* stream);

/*
 * Static circular aperture, npix x npix float32: a hard circle of the given
 * diameter, an optional secondary obscuration (secondary_diameter <= 0: none)
 * and n_spiders one-sided vanes of width spider_width at the given angles
 * (degrees, HOST array), evaluated on an oversample x oversample sub-grid of
 * pixel_coords(npix*oversample, array_diameter) and box-averaged.
 * support (may be NULL) receives 1 where any sub-pixel is inside the primary.
 * array_diameter <= 0 means `diameter`.
 * Replaces: circular_aperture (transmission and support), utils/apertures.py:153-242.
 */
int dlx_circular_aperture(int32_t npix, double diameter, int32_t oversample,
                          double secondary_diameter, double array_diameter,
                          double spider_width, int32_t n_spiders,
                          const double* spider_angles_deg, float* transmission,
                          float* support, void* stream);

/*
 * Coordinate transformation of a dynamic aperture (CoordTransform,
 * coordinates.py:230-316): translation, then shear, compression and rotation,
 * each applied when its flag bit is set.  Shear follows
 * utils/coordinates.py:61-78: x += shear_x * x^T and y += shear_y * y^T, where
 * ^T transposes the coordinate plane (pixel (i, j) reads pixel (j, i)).
 * Rotation is rotate_coords (utils/coordinates.py:81-101), in radians.
 */
enum dlx_transform_flags {
  DLX_TF_TRANSLATION = 1,
  DLX_TF_SHEAR = 2,
  DLX_TF_COMPRESSION = 4,
  DLX_TF_ROTATION = 8
};

typedef struct dlx_coord_transform {
  int32_t flags;
  int32_t reserved_;
  double translation[2];
  double shear[2];
  double compression[2];
  double rotation;
} dlx_coord_transform;

/*
 * Soft-edged dynamic aperture (layers/apertures.py:155-641, utils/geometry.py:
 * 229-579): the signed distance to the edge (positive inside; inverted when
 * occulting), clipped to +-clip with clip = pixel_scale * softening / 2, then
 * rescaled by its own minimum and maximum over the grid (soften,
 * utils/geometry.py:49-92; a constant field gives the hard support d > 0).
 *   DLX_AP_CIRCLE       size[0] = radius        (CircularAperture)
 *   DLX_AP_SQUARE       size[0] = width         (SquareAperture)
 *   DLX_AP_RECTANGLE    size = width, height    (RectangularAperture)
 *   DLX_AP_REG_POLYGON  size[0] = rmax, nsides  (RegPolyAperture, 3..64 sides)
 *   DLX_AP_SPIDER       size[0] = arm width, angles_deg[n_angles] (Spider, <= 16 arms)
 * The coordinates are the pixel centres c + (k - (npix-1)/2) * pixel_scale
 * ('xy' order), put through transforms[0 .. n_transforms) in order (an
 * enclosing composite aperture's transformation first).
 */
enum dlx_aperture_kind {
  DLX_AP_CIRCLE = 1,
  DLX_AP_SQUARE = 2,
  DLX_AP_RECTANGLE = 3,
  DLX_AP_REG_POLYGON = 4,
  DLX_AP_SPIDER = 5
};

typedef struct dlx_soft_aperture {
  int32_t kind;
  int32_t occulting;
  int32_t nsides;
  int32_t n_angles;
  int32_t n_transforms;       /* 0..3 */
  int32_t group;              /* see dlx_soft_apertures                         */
  double size[2];
  double softening;           /* pixels, > 0                                    */
  double angles_deg[16];
  dlx_coord_transform transforms[3];
  const double* coords;       /* NULL, or device float64 planes x then y (2 x npix x npix)
                                 of the transformed pixel coordinates, used instead of the
                                 transform chain (a DistortedCoords, coordinates.py:318-357) */
} dlx_soft_aperture;

/*
 * Transmission of n (<= 16) soft apertures on an npix x npix grid:
 *   out = sum over groups g of prod over the apertures of g of T_a
 * (apertures of one group consecutive).  One group is a CompoundAperture
 * (product, layers/apertures.py:1005-1069); one aperture per group is a
 * MultiAperture (sum, :1072-1117); a MultiAperture of CompoundApertures is a
 * group per compound.  (The typedef above describes one aperture.)
 */
size_t dlx_soft_apertures_workspace_size(int32_t npix, int32_t n);
int dlx_soft_apertures(int32_t npix, double pixel_scale, double center, int32_t n,
                      const dlx_soft_aperture* apertures, float* out, void* workspace,
                      size_t workspace_bytes, void* stream);

/*
 * Vector-Jacobian product of dlx_soft_apertures with respect to each
 * aperture's geometry: given dout = dL/dout (npix x npix float32), writes
 * grads[a * 9 + k] = dL/dtheta_k for theta = (size[0], size[1], and the last
 * transform of its chain: translation x, y, shear x, y, compression x, y,
 * rotation); parameters a transform does not have get 0.  soften's extrema
 * contribute through the average over the pixels that attain them (the rule
 * JAX applies to min / max).  Replaces the reverse-mode derivative JAX builds
 * through DynamicAperture.transmission (layers/apertures.py:299-641) and soften
 * (utils/geometry.py:49-92).
 */
size_t dlx_soft_apertures_vjp_workspace_size(int32_t npix, int32_t n);
int dlx_soft_apertures_vjp(int32_t npix, double pixel_scale, double center, int32_t n,
                           const dlx_soft_aperture* apertures, const float* dout, float* grads,
                           void* workspace, size_t workspace_bytes, void* stream);

/*
 * Resampling of `batch` float32 images (n_in x n_in, knots at center_in +
 * (k - (n_in-1)/2) pixel_scale_in) onto an n_out x n_out grid whose pixel
 * coordinates (center_out + (k - (n_out-1)/2) pixel_scale_out) are first sent
 * through the transform chain: method 1 bilinear, 0 the nearest knot; points
 * outside the knots' span take `fill`.  Replaces interpax.interp2d as called by
 * dlu.interp / scale / rotate (utils/interpolation.py:13-108), hence
 * Wavefront.interpolate / scale_to / rotate (wavefronts.py:442-566),
 * PSF.rotate / interpolate (psfs.py:112-157), Interpolate and
 * ApplyInterpolation (layers/optical_layers.py:108-141,
 * layers/detector_layers.py:68-97).  interpax is not available, so this is the
 * textbook bilinear rule.
 */
int dlx_interp_transform(int32_t n_in, double pixel_scale_in, double center_in, int32_t n_out,
                         double pixel_scale_out, double center_out, int32_t n_transforms,
                         const dlx_coord_transform* transforms, int32_t method, float fill,
                         int32_t batch, const float* in, float* out, void* stream);

/*
 * The same sampling at npts explicit points (qx, qy: device float64) of each
 * image: out[b][p].  Replaces interpax.interp2d as called by dlu.interp with
 * arbitrary sample coordinates (utils/interpolation.py:13-50).
 */
int dlx_interp_points(int32_t n_in, double pixel_scale_in, double center_in, int64_t npts,
                      const double* qx, const double* qy, int32_t method, float fill,
                      int32_t batch, const float* in, float* out, void* stream);

/*
 * Adjoints of dlx_interp_transform and dlx_interp_points with respect to the
 * images: din (batch x n_in x n_in float32) = S^T dout, S the sampling
 * matrix (bilinear weights or the nearest knot; points outside the knots'
 * span contribute nothing).  workspace: batch x n_in x n_in doubles (the FP64
 * sums, filled by a kernel; capturable).  Replaces the reverse-mode
 * derivative JAX takes through interpax.interp2d in the resampling helpers
 * (utils/interpolation.py:13-50).
 */
int dlx_interp_transform_vjp(int32_t n_in, double pixel_scale_in, double center_in,
                             int32_t n_out, double pixel_scale_out, double center_out,
                             int32_t n_transforms, const dlx_coord_transform* transforms,
                             int32_t method, int32_t batch, const float* dout, float* din,
                             double* workspace, void* stream);
int dlx_interp_points_vjp(int32_t n_in, double pixel_scale_in, double center_in, int64_t npts,
                          const double* qx, const double* qy, int32_t method, int32_t batch,
                          const float* dout, float* din, double* workspace, void* stream);

/*
 * Zernike basis of an aberrated dynamic aperture (AberratedAperture.calc_basis,
 * layers/apertures.py:735-752; DynamicZernikeBasis, polynomials.py:85-115): the
 * pixel coordinates put through the transforms, divided by `extent`, then
 * Z_j (nsides == 0, utils/zernikes.py:257-294) or the n-sided polike
 * (nsides >= 3, :352-397) on the unit disk.  out: nmodes x npix x npix float32.
 * coords (may be NULL): device float64 planes x then y of the transformed pixel
 * coordinates, used instead of the transforms (a DistortedCoords).
 */
int dlx_aperture_basis(int32_t npix, double pixel_scale, double center, int32_t n_transforms,
                       const dlx_coord_transform* transforms, const double* coords, double extent,
                       int32_t nsides, int32_t nmodes, const int32_t* js, float* out,
                       void* stream);

/*
 * Centred, zero-padded 2-D FFT of `batch` N x N complex64 fields to M x M
 * (M - N even, M a power of two up to 8192):
 *   forward: out = fftshift(fft2(ifftshift(pad_M(in)))) / M
 *   inverse: out = fftshift(ifft2(ifftshift(pad_M(in)))) * M
 * Replaces: FFT, utils/propagation.py:8-64 (Wavefront.propagate_FFT,
 * wavefronts.py:654-727).  Returns DLX_ERR_UNSUPPORTED for other M.
 */
size_t dlx_fft2_workspace_size(int32_t n, int32_t m, int32_t batch);
int dlx_fft2(int32_t n, int32_t m, int32_t batch, int32_t inverse, const void* in,
             void* out, void* workspace, size_t workspace_bytes, void* stream);

/*
 * Angular-spectrum free-space propagation of `batch` N x N complex64 fields,
 * zero-padded to M x M (same M rules as dlx_fft2), over `distance` metres:
 *   out = IFFT{ FFT{pad_M(in)} * H },  H = exp(2 pi i z sqrt(1/l^2 - fx^2 - fy^2))
 * on propagating frequencies and 0 on evanescent ones; fx, fy = (i - M/2)/(M dx).
 * Output keeps the input sampling dx and has M x M pixels (no crop).
 * Replaces: abcdLux.asm.asm_prop as called by ASMPropagator
 * (layers/abcd_propagators.py:444-451).  abcdLux is not available, so this
 * follows the textbook transfer-function method.
 */
size_t dlx_asm_workspace_size(int32_t n, int32_t m, int32_t batch);
int dlx_asm_prop(int32_t n, int32_t m, int32_t batch, const double* wavelength,
                 double pixel_scale, double distance, const void* in, void* out,
                 void* workspace, size_t workspace_bytes, void* stream);

/*
 * Linear canonical transform (Collins integral) of `batch` N_in x N_in complex64
 * fields for the ABCD matrix abcd = {A, B, C, D} (row-major; B != 0), per axis
 *   u_out(x') = sqrt(dx dx' / (i lambda B)) sum_x u_in(x)
 *               exp(i pi / (lambda B) (A x^2 - 2 x x' + D x'^2)),
 * on x = center_in + (k - (N_in-1)/2) pixel_scale_in and
 * x' = center_out + (l - (N_out-1)/2) pixel_scale_out; the 2-D field maps as
 * K u K^T with two complex GEMMs (the MFT's), precision as for dlx_mft.
 * wavelength[b] per field.  Replaces abcdLux.lct.lct_prop as called by
 * MFTPropagator (layers/abcd_propagators.py:245-297); abcdLux is not
 * available, so this follows the textbook Collins integral (the sign, the
 * 1/i phase and the power-preserving sqrt(dx dx') scale are this library's
 * convention).
 */
size_t dlx_lct_workspace_size(int32_t n_in, int32_t n_out, int32_t batch, int32_t precision);
int dlx_lct(int32_t n_in, int32_t n_out, int32_t batch, const double* abcd,
            const double* wavelength, double pixel_scale_in, double center_in,
            double pixel_scale_out, double center_out, const void* in, void* out,
            int32_t precision, void* workspace, size_t workspace_bytes, void* stream);

/*
 * The same transform on the FFT grid: M x M outputs (M - N even, M a power
 * of two up to 8192) at pixel_scale_out[b] = lambda_b |B| / (M pixel_scale_in)
 * (written if non-NULL), centred on center_out; computed as an input chirp,
 * the centred padded FFT of dlx_fft2 (inverse when B < 0) and an output chirp.
 * Replaces abcdLux.lct.lct_prop_fft as called by FFTPropagator
 * (layers/abcd_propagators.py:300-383).
 */
size_t dlx_lct_fft_workspace_size(int32_t n, int32_t m, int32_t batch);
int dlx_lct_fft(int32_t n, int32_t m, int32_t batch, const double* abcd,
                const double* wavelength, double pixel_scale_in, double center_in,
                double center_out, const void* in, void* out, double* pixel_scale_out,
                void* workspace, size_t workspace_bytes, void* stream);

/*
 * Detector chain: the reference's detector layers applied in order to
 * `batch` float32 images (n x n each, stride n*n):
 *   DLX_DET_PIXEL_RESPONSE  img *= map     (map: device f32, size x size, size =
 *                                           the image side at that point)
 *                           ApplyPixelResponse, layers/detector_layers.py:100-131
 *   DLX_DET_CONVOLVE        img = convolve(img, kernel, mode="same"), zero padded,
 *                           kernel device f32 size x size (1 <= size <= 64);
 *                           output index i reads input i + (size-1)/2 - k
 *                           ApplyJitter / PSF.convolve, layers/detector_layers.py:
 *                           134-199, psfs.py:93-110
 *   DLX_DET_CONVOLVE_SEPARABLE  as DLX_DET_CONVOLVE with kernel[ky][kx] =
 *                           data[ky] * data[size + kx] (2*size floats: the column
 *                           then the row factor), computed in two 1-D passes
 *                           (ApplyJitter's Gaussian kernel is such an outer product)
 *   DLX_DET_DOWNSAMPLE      size x size block sums (side must divide by size)
 *                           Downsample / PSF.downsample,
 *                           layers/detector_layers.py:262-293, psfs.py:74-91
 *   DLX_DET_SATURATE        img = min(img, value)
 *                           ApplySaturation, layers/detector_layers.py:202-229
 *   DLX_DET_ADD_CONSTANT    img += value
 *                           AddConstant, layers/detector_layers.py:232-259
 * (LayeredDetector.__call__, detectors.py:103-128.)  Elementwise ops fuse into
 * the load or the store of the nearest convolve / downsample pass, so the
 * chain costs one pass over the image per convolve or downsample.  `value`
 * is used as float32.  dlx_detector_out_side validates the chain and gives
 * the output side; `out` holds batch x out_side^2 floats.
 */
enum dlx_detector_kind {
  DLX_DET_PIXEL_RESPONSE = 1,
  DLX_DET_CONVOLVE = 2,
  DLX_DET_DOWNSAMPLE = 3,
  DLX_DET_SATURATE = 4,
  DLX_DET_ADD_CONSTANT = 5,
  DLX_DET_CONVOLVE_SEPARABLE = 6
};

typedef struct dlx_detector_op {
  int32_t kind;        /* dlx_detector_kind                                   */
  int32_t size;        /* map side, kernel side or downsample factor          */
  double value;        /* threshold or constant                               */
  const float* data;   /* device map or kernel; NULL for the other kinds      */
} dlx_detector_op;

int dlx_detector_out_side(int32_t n, const dlx_detector_op* ops, int32_t nops,
                          int32_t* out_side);
size_t dlx_detector_workspace_size(int32_t n, int32_t batch, const dlx_detector_op* ops,
                                   int32_t nops, int32_t for_vjp);
int dlx_detector_apply(int32_t n, int32_t batch, const float* in, const dlx_detector_op* ops,
                       int32_t nops, float* out, void* workspace, size_t workspace_bytes,
                       void* stream);

/*
 * The polychromatic PSF of dlx_psf_poly (same arguments) seen through the
 * detector chain `ops` (as dlx_detector_apply on the N_out x N_out PSF), with
 * the wavelength sum fused into the chain: the contractions' |E|^2 epilogue
 * writes each wavelength's w_l |E_l|^2 plane, and the chain's first pass
 * (convolve, convolve + downsample, downsample or elementwise) loads the sum
 * over the planes in wavelength order as it reads each pixel, so the PSF is
 * never written or read back.  The result equals dlx_psf_poly followed by
 * dlx_detector_apply.  `out` holds out_side^2 floats (dlx_detector_out_side).
 * Replaces: Telescope.model for a point source (instruments.py:140-171):
 * source.model(optics) -> OpticalSystem.propagate (optical_systems.py:147-223)
 * -> detector.model (detectors.py:103-128, layers/detector_layers.py:100-293).
 */
size_t dlx_psf_poly_detector_workspace_size(const dlx_mft_geom* geom, int32_t nwl,
                                            int32_t precision, const dlx_detector_op* ops,
                                            int32_t nops);
int dlx_psf_poly_detector(const dlx_mft_geom* geom, int32_t nwl, const double* wavelength,
                          const double* weight, double pixel_scale_in,
                          double pixel_scale_out, double center, double tilt_x,
                          double tilt_y, const float* amp, const float* opd,
                          const float* phase, double scale, const dlx_detector_op* ops,
                          int32_t nops, float* out, int32_t precision, void* workspace,
                          size_t workspace_bytes, void* stream);

/*
 * Vector-Jacobian product of dlx_detector_apply with respect to its input
 * images: din = (d out / d in)^T dout.  The chain is re-run forward to keep
 * each pass's input and pre-store values (workspace sized with for_vjp = 1).
 * Adjoints: convolve -> correlation with the same kernel, downsample ->
 * replication, pixel response -> the same map, saturation -> 1 below the
 * threshold, 1/2 at it, 0 above (jnp.minimum's derivative), constant -> 1.
 * Replaces the reverse-mode derivative JAX builds of the detector layers
 * (layers/detector_layers.py:100-293).
 */
int dlx_detector_vjp(int32_t n, int32_t batch, const float* in, const dlx_detector_op* ops,
                     int32_t nops, const float* dout, float* din, void* workspace,
                     size_t workspace_bytes, void* stream);

/*
 * dlx_detector_vjp plus the cotangents of the ops' parameters: grads[k] (a
 * host array of nops device pointers; NULL = not wanted) receives, summed
 * over the batch,
 *   DLX_DET_PIXEL_RESPONSE      dL/dmap     size x size floats
 *   DLX_DET_CONVOLVE            dL/dkernel  size x size floats
 *   DLX_DET_CONVOLVE_SEPARABLE  dL/d(column factor) then dL/d(row factor), 2 size floats
 *   DLX_DET_SATURATE, DLX_DET_ADD_CONSTANT  dL/dvalue, one float (the saturation
 *                               level takes jnp.minimum's 1/2 at ties)
 * A downsample factor has none (a non-NULL entry is an error).
 * Replaces the reverse-mode derivative JAX builds for the detector layers'
 * arrays (ApplyPixelResponse.pixel_response, ApplyJitter's kernel,
 * ApplySaturation.saturation, AddConstant.value;
 * layers/detector_layers.py:100-260, psfs.py:93-110).
 */
size_t dlx_detector_vjp_params_workspace_size(int32_t n, int32_t batch,
                                              const dlx_detector_op* ops, int32_t nops);
int dlx_detector_vjp_params(int32_t n, int32_t batch, const float* in,
                            const dlx_detector_op* ops, int32_t nops, const float* dout,
                            float* din, float* const* grads, void* workspace,
                            size_t workspace_bytes, void* stream);

/*
 * Transfer-matrix cache (no counterpart in the reference: dLux rebuilds the
 * matrices inside every traced call).  dlx_mft, dlx_psf_poly,
 * dlx_psf_poly_vjp and dlx_mft_batched_psf keep the matrices of each
 * (geometry, wavelengths, pixel scales, shifts) batch resident on the device
 * and reuse them; least-recently-used batches are
 * evicted beyond the budget (default 16 GiB, environment DLUX_AMD_CACHE_MB).
 * A budget of 0 disables caching.  The cache is bypassed for streams under
 * graph capture.  dlx_cache_clear synchronises on outstanding uses and frees
 * every entry.  An evicted entry's block is reused without the host waiting:
 * the stream that takes it next waits on the device for the work that last
 * read it.  dlx_cache_bypass(1) makes the calling thread's calls skip the cache
 * (matrices generated into their workspace, nothing allocated, evicted or
 * waited on) until dlx_cache_bypass(0): for request loops whose wavelength
 * grids rarely repeat.
 */
/*
 * Matrix-core rate probe (no counterpart in the reference): the bf16 MFMA
 * throughput (TFLOP/s, into *tflops) this GPU sustains on
 * v_mfma_f32_32x32x16_bf16 from registers, 8 waves per CU, with low-entropy
 * operands (random_operands = 0: the nominal rate) or with full-entropy
 * operands like the split-BF16x3 contractions' (1).  Runs ~40 launches of
 * 1-2 ms on the stream (the first 30 let the clocks settle) and waits for
 * them; workspace >= 512 floats per CU.
 */
int dlx_mfma_rate(int32_t random_operands, double* tflops, void* workspace,
                  size_t workspace_bytes, void* stream);

int dlx_cache_clear(void);
int dlx_cache_set_budget(size_t bytes);
int dlx_cache_bypass(int32_t on);
int dlx_cache_stats(uint64_t* hits, uint64_t* misses, size_t* bytes, int32_t* entries);

/*
 * Diagnostics (no counterpart in the reference): live timing of the MFT
 * contraction launches inside dlx_psf_poly.  While enabled, each stage-1
 * (pupil x transfer matrix) and stage-2 (transfer matrix x intermediate, with
 * the |E|^2 epilogue) launch is bracketed by hipEvents on the caller's stream
 * (at most 8192 launches are recorded).  enable resets the record;
 * dlx_profile_read waits for the recorded events and returns the summed
 * milliseconds and the launch count of one stage (1 or 2).  Not thread-safe.
 */
int dlx_profile_enable(int32_t enable);
int dlx_profile_read(int32_t stage, double* total_ms, int32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* DLUX_AMD_H */
