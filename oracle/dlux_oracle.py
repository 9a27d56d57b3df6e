"""FP64 NumPy oracle for dLux's diffraction hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP path is compared
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it; the product package (``dlux_amd``) never
does, and nothing here is ever the thing that is measured or shipped.

Every function restates, operation for operation, the dLux 0.16.0 routine it
cites (paths relative to the reference's repository root), but computes in
float64/complex128 -- dLux's own code under ``jax_enable_x64``.  dLux itself
cannot be imported here (jax, equinox, zodiax, abcdLux and interpax are absent),
so the restatement is pinned instead by the reference tests' known answers,
closed forms and invariants (``tests/test_oracle.py``).

Parity status per function (all pinned unless marked):
  * nd_coords / pixel_coords / cart2polar / rotate_coords        pinned
  * transfer_matrix / calc_nfringes / mft                        pinned
  * fft / fft_spec / fft_phase_ramp / propagate_fft              pinned
  * pad_to / crop_to / resize / downsample                       pinned
  * noll_indices / zernike_factors / zernike / zernike_basis     pinned
  * circle / spider / combine / circular_aperture                pinned
  * eval_basis, wavefront chain, poly_psf, point_sources_psf     pinned
  * poly_psf_grad (adjoint gradient)                             pinned (FD)
  * gaussian / jitter_kernel / convolve_same / detector           pinned
  * detector_vjp                                                 pinned (FD, adjoint identity)
  * coord_transform / distort_coords / soften / *_distance / soft_aperture / polike  pinned
  * asm_prop (angular spectrum)                                  parity unpinned
      -- its conventions live in abcdLux (absent); checked by textbook
         invariants (z=0 identity, unitarity, z1+z2 composition) and the
         paraxial Gaussian-beam closed form.
  * interp_grid (bilinear / nearest resampling)                 pinned (scipy)
      -- parity with interpax.interp2d (absent) unpinned beyond those rules.
  * abcd_* / compose_abcd / lct / lct_fft (Collins integral)      parity unpinned
      -- against abcdLux (absent); pinned by closed forms: a tilted Gaussian
         beam (the ABCD law), the conjugate plane as the MFT, the FFT form as
         the Collins sum on the FFT grid.
"""

from __future__ import annotations

from math import lgamma, exp

import numpy as np

# ---------------------------------------------------------------------------
# coordinates  (src/dLux/utils/coordinates.py, src/dLux/coordinates.py)
# ---------------------------------------------------------------------------


def nd_coords_1d(n: int, scale: float = 1.0, offset: float = 0.0) -> np.ndarray:
    """One axis of ``nd_coords``: linspace(-(n-1)/2*d - o, (n-1)/2*d - o, n).

    utils/coordinates.py:320-324 (``pixel_fn``).
    """
    start = -(n - 1) / 2 * scale - offset
    end = (n - 1) / 2 * scale - offset
    return np.linspace(start, end, n)


def pixel_coords(npixels: int, diameter: float, fft_style: bool = False) -> np.ndarray:
    """(x, y) meshgrid of pixel centres, 'xy' indexing.  utils/coordinates.py:208-266."""
    d = diameter / npixels
    off = d / 2.0 if (fft_style and npixels % 2 == 0) else 0.0
    xs = nd_coords_1d(npixels, d, off)
    return np.array(np.meshgrid(xs, xs, indexing="xy"))


def cart2polar(coords: np.ndarray) -> np.ndarray:
    """utils/coordinates.py:166-184."""
    x, y = coords
    return np.array([np.hypot(x, y), np.arctan2(y, x)])


def rotate_coords(coords: np.ndarray, rotation: float) -> np.ndarray:
    """utils/coordinates.py:81-101 (rotation in radians)."""
    x, y = coords
    new_x = np.cos(-rotation) * x + np.sin(-rotation) * y
    new_y = -np.sin(-rotation) * x + np.cos(-rotation) * y
    return np.array([new_x, new_y])


def spec_xs(n: int, d: float, c: float = 0.0) -> np.ndarray:
    """``CoordSpec.xs`` = c + (arange(n) - (n-1)/2) * d.  coordinates.py:117-129."""
    return c + (np.arange(n) - (n - 1) / 2) * d


# ---------------------------------------------------------------------------
# MFT  (src/dLux/utils/propagation.py)
# ---------------------------------------------------------------------------


def transfer_matrix(
    wavelength,
    npixels_in,
    pixel_scale_in,
    npixels_out,
    pixel_scale_out,
    shift=0.0,
    focal_length=None,
    focal_shift=0.0,
    inverse=False,
):
    """N_in x N_out DFT matrix exp(-/+ 2 pi i a (x) b).  utils/propagation.py:67-127."""
    fringe_size = wavelength / (pixel_scale_in * npixels_in)
    scale_in = 1.0 / npixels_in
    in_vec = nd_coords_1d(npixels_in, scale_in, shift * scale_in)
    scale_out = pixel_scale_out / fringe_size
    if focal_length is not None:
        scale_out /= focal_length + focal_shift
    out_vec = nd_coords_1d(npixels_out, scale_out, shift * scale_out)
    matrix = -2j * np.pi * np.outer(in_vec, out_vec)
    if inverse:
        matrix *= -1
    return np.exp(matrix)


def calc_nfringes(
    wavelength, npixels_in, pixel_scale_in, npixels_out, pixel_scale_out,
    focal_length=None, focal_shift=0.0,
):
    """utils/propagation.py:130-175."""
    diameter = npixels_in * pixel_scale_in
    fringe_size = wavelength / diameter
    output_size = npixels_out * pixel_scale_out
    if focal_length is not None:
        output_size /= focal_length + focal_shift
    return output_size / fringe_size


def mft(
    phasor,
    wavelength,
    pixel_scale_in,
    npixels_out,
    pixel_scale_out,
    focal_length=None,
    shift=(0.0, 0.0),
    pixel=True,
    inverse=False,
):
    """Two-sided MFT (Y^T E) X times nfringes/(N_in N_out).  utils/propagation.py:178-256."""
    phasor = np.asarray(phasor, np.complex128)
    npixels_in = phasor.shape[-1]
    shift = np.asarray(shift, np.float64).copy()
    if not pixel:
        shift = shift / pixel_scale_out
    x_mat = transfer_matrix(wavelength, npixels_in, pixel_scale_in, npixels_out,
                            pixel_scale_out, shift[0], focal_length, 0.0, inverse)
    y_mat = transfer_matrix(wavelength, npixels_in, pixel_scale_in, npixels_out,
                            pixel_scale_out, shift[1], focal_length, 0.0, inverse)
    out = (y_mat.T @ phasor) @ x_mat
    nfringes = calc_nfringes(wavelength, npixels_in, pixel_scale_in, npixels_out,
                             pixel_scale_out, focal_length)
    return out * np.exp(np.log(nfringes) - (np.log(npixels_in) + np.log(npixels_out)))


# ---------------------------------------------------------------------------
# FFT  (src/dLux/utils/propagation.py:8-64, src/dLux/utils/fourier.py:14-78)
# ---------------------------------------------------------------------------


def fft(phasor, wavelength, pixel_scale, focal_length=None, pad=2, inverse=False):
    """pad -> ifftshift -> fft2 -> fftshift, / N_pad (inverse: ifft2, * N_pad)."""
    phasor = np.asarray(phasor, np.complex128)
    npixels = phasor.shape[-1]
    fringe_size = wavelength / (pixel_scale * npixels)
    new_pixel_scale = fringe_size / pad
    if focal_length is not None:
        new_pixel_scale *= focal_length
    npad = (npixels * (pad - 1)) // 2
    phasor = np.pad(phasor, npad)
    if inverse:
        phasor = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(phasor)))
        phasor = phasor * phasor.shape[-1]
    else:
        phasor = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(phasor)))
        phasor = phasor / phasor.shape[-1]
    return phasor, new_pixel_scale


def fft_spec(npixels_in, pixel_scale_in, wavelength, focal_length=None):
    """utils/fourier.py:14-45."""
    d_out = wavelength / (npixels_in * pixel_scale_in)
    if focal_length is not None:
        d_out = d_out * focal_length
    if npixels_in % 2 != 0:
        return d_out, 0.0
    return d_out, -0.5 * d_out


def fft_phase_ramp(xs, wavelength, shift, focal_length=None, inverse=False):
    """utils/fourier.py:48-78."""
    sign = -1 if inverse else 1
    alpha = wavelength if focal_length is None else wavelength * focal_length
    ramp = np.exp(sign * 2j * np.pi * np.asarray(xs) * shift / alpha)
    return ramp[None, :] * ramp[:, None]


# ---------------------------------------------------------------------------
# array ops  (src/dLux/utils/array_ops.py)
# ---------------------------------------------------------------------------


def pad_to(array, npixels, fill=0.0):
    """utils/array_ops.py:12-50 (parity-preserving centred pad)."""
    n_in = array.shape[-1]
    if n_in % 2 != npixels % 2:
        raise ValueError("parity")
    if npixels < n_in:
        raise ValueError("smaller")
    p = (npixels - n_in) // 2
    return np.pad(array, p, constant_values=fill)


def crop_to(array, npixels):
    """utils/array_ops.py:53-90."""
    n_in = array.shape[-1]
    if n_in % 2 != npixels % 2:
        raise ValueError("parity")
    if npixels > n_in:
        raise ValueError("larger")
    start, stop = (n_in - npixels) // 2, (n_in + npixels) // 2
    return array[start:stop, start:stop]


def resize(array, npixels, fill=0.0):
    """utils/array_ops.py:93-121."""
    n_in = array.shape[-1]
    if npixels == n_in:
        return array
    if npixels < n_in:
        return crop_to(array, npixels)
    return pad_to(array, npixels, fill)


def downsample(array, n, mean=True):
    """utils/array_ops.py:124-162 (two reshape-reduce passes)."""
    method = np.mean if mean else np.sum
    size_in = array.shape[0]
    size_out = size_in // n
    array = method(array.reshape((size_in * size_out, n)), 1)
    array = array.reshape(size_in, size_out).T
    array = method(array.reshape((size_out * size_out, n)), 1)
    return array.reshape(size_out, size_out).T


# ---------------------------------------------------------------------------
# Zernikes  (src/dLux/utils/zernikes.py, src/dLux/utils/math.py)
# ---------------------------------------------------------------------------


def noll_indices(j: int):
    """(n, m) from the Noll index.  utils/zernikes.py:99-119 (integer formula)."""
    n = int(np.ceil(-1 / 2 + np.sqrt(1 + 8 * j) / 2) - 1)
    smallest_j_in_row = n * (n + 1) / 2 + 1
    number_of_shifts = (j - smallest_j_in_row + ~(n & 1) + 2) // 2
    sign_of_shift = -(j & 1) + ~(j & 1) + 2
    base_case = n & 1
    m = int(sign_of_shift * (base_case + number_of_shifts * 2))
    return n, m


def factorial(x: float) -> float:
    """exp(lgamma(x + 1)), 1 at 0; inf at negative integers.  utils/math.py:137-157."""
    if x == 0:
        return 1.0
    if x < 0 and float(x).is_integer():
        return float("inf")
    return exp(lgamma(x + 1.0))


def zernike_factors(j: int):
    """Radial coefficients c_k and powers k, using signed m.  utils/zernikes.py:122-150."""
    if j < 1:
        raise ValueError("The Zernike index must be greater than 0.")
    n, m = noll_indices(j)
    k = np.arange(((n - m) // 2) + 1, dtype=float)
    c = np.array([
        (-1.0) ** kk * factorial(abs(n - kk)) / factorial(kk)
        / factorial(((n + m) // 2) - kk) / factorial(((n - m) // 2) - kk)
        for kk in k
    ])
    return c, k


def zernike(j: int, coordinates, diameter: float = 2.0):
    """Noll Zernike on the disk of given diameter.  utils/zernikes.py:227-257."""
    coords = np.asarray(coordinates, np.float64) / (diameter / 2)
    rho, theta = cart2polar(coords)
    aperture = rho <= 1.0
    n, m = noll_indices(j)
    c, k = zernike_factors(j)
    with np.errstate(divide="ignore", invalid="ignore"):
        rads = rho[:, :, None] ** (abs(n) - 2 * k)[None, None, :]
        radial = (c * rads).sum(axis=2)
    norm = np.sqrt(n + 1)
    if m != 0:
        norm *= np.sqrt(2)
    trig = np.cos if m >= 0 else np.sin
    return aperture * radial * (norm * trig(abs(m) * theta))


def zernike_basis(js, coordinates, diameter: float = 2.0):
    """utils/zernikes.py:297-315."""
    return np.array([zernike(int(j), coordinates, diameter) for j in js])


def eval_basis(basis, coefficients):
    """tensordot over the leading coefficient dims.  utils/math.py:177-196."""
    coefficients = np.asarray(coefficients, np.float64)
    axes = tuple(range(coefficients.ndim))
    return np.tensordot(np.asarray(basis, np.float64), coefficients, axes=(axes, axes))


# ---------------------------------------------------------------------------
# hard masks  (src/dLux/utils/geometry.py, src/dLux/utils/apertures.py)
# ---------------------------------------------------------------------------


def circle(coords, radius, invert=False):
    """Hard circle r < R (or r > R).  utils/geometry.py:95-117, :405-418."""
    dist = cart2polar(coords)[0] - radius
    return (dist > 0).astype(float) if invert else (dist < 0).astype(float)


def spider(coords, width, angles):
    """One-sided vanes: blocked where max(|x'|-w/2, y') < 0.  geometry.py:195-218, :463-485."""
    blocked = np.zeros(coords.shape[1:], bool)
    for angle in np.atleast_1d(angles):
        x, y = rotate_coords(coords, np.deg2rad(angle))
        blocked |= np.maximum(np.abs(x) - width / 2, y) < 0
    return (~blocked).astype(float)


def combine(arrays, oversample=1, use_sum=False):
    """Product (or sum) then downsample.  utils/geometry.py:20-46."""
    arr = np.array(arrays)
    out = arr.sum(0) if use_sum else arr.prod(0)
    return out if oversample == 1 else downsample(out, oversample)


def circular_aperture(
    npixels, diameter, oversample=5, secondary_diameter=None, spider_width=None,
    spider_angles=None, zernike_nolls=None, zernike_oversize=0.01,
):
    """Static circular aperture (+ basis masked by the primary).  utils/apertures.py:153-242."""
    coords = pixel_coords(npixels * oversample, diameter)
    layers = [circle(coords, diameter / 2)]
    if secondary_diameter is not None and secondary_diameter > 0:
        layers.append(circle(coords, secondary_diameter / 2, invert=True))
    if spider_width is not None:
        layers.append(spider(coords, spider_width, spider_angles))
    transmission = combine(layers, oversample)
    if zernike_nolls is None:
        return transmission
    coords = pixel_coords(npixels, diameter)
    basis = zernike_basis(zernike_nolls, coords, diameter * (1.0 + zernike_oversize))
    support = downsample(layers[0], oversample) > 0
    return transmission, basis * support[None]


# ---------------------------------------------------------------------------
# Wavefront chain  (src/dLux/wavefronts.py, src/dLux/optical_systems.py)
# ---------------------------------------------------------------------------


def wavefront_init(npixels: int):
    """ones/N^2 phasor.  wavefronts.py:111-113."""
    return np.ones((npixels, npixels), np.complex128) / npixels ** 2


def tilt_opd(npixels, pixel_scale, angles, center=0.0):
    """theta_x X + theta_y Y on meshgrid(xs, xs).  wavefronts.py:370-395, :584-609."""
    xs = spec_xs(npixels, pixel_scale, center)
    X, Y = np.meshgrid(xs, xs)
    return angles[0] * X + angles[1] * Y


def add_opd(phasor, wavelength, opd):
    """phasor * exp(1j * 2 pi / lambda * opd).  wavefronts.py:332-368."""
    return phasor * np.exp(1j * (2 * np.pi / wavelength) * opd)


def normalise(phasor, mode="power", value=1.0):
    """wavefronts.py:397-424."""
    if mode == "power":
        return phasor * np.sqrt(value / np.sum(np.abs(phasor) ** 2))
    return phasor * np.sqrt(value / np.max(np.abs(phasor) ** 2))


def pupil_field(wavelength, mask, basis, coeffs, diameter, offset=(0.0, 0.0),
                normalise_power=True):
    """Initial wavefront -> tilt -> TransmissiveLayer(mask, normalise) -> BasisLayer(opd).

    optical_systems.py:363-389 + layers/optical_layers.py:181-186, :319-327.
    """
    n = mask.shape[-1]
    d = diameter / n
    E = wavefront_init(n)
    E = add_opd(E, wavelength, tilt_opd(n, d, np.asarray(offset, np.float64)))
    E = E * mask
    if normalise_power:
        E = normalise(E)
    if basis is not None:
        E = add_opd(E, wavelength, eval_basis(basis, coeffs))
    return E


def poly_psf(mask, basis, coeffs, wavelengths, weights, diameter, npixels_out,
             pixel_scale_out, offset=(0.0, 0.0), normalise_power=True,
             return_fields=False):
    """Sum_l w_l |MFT(P_l)|^2 -- OpticalSystem.propagate.  optical_systems.py:147-223, :659-680."""
    wavelengths = np.atleast_1d(np.asarray(wavelengths, np.float64))
    weights = (np.ones_like(wavelengths) / len(wavelengths) if weights is None
               else np.atleast_1d(np.asarray(weights, np.float64)))
    n = mask.shape[-1]
    d = diameter / n
    psf = np.zeros((npixels_out, npixels_out))
    fields = []
    for wl, w in zip(wavelengths, weights):
        P = pupil_field(wl, mask, basis, coeffs, diameter, offset, normalise_power)
        E = mft(P, wl, d, npixels_out, pixel_scale_out) * w ** 0.5
        fields.append(E)
        psf += np.abs(E) ** 2
    return (psf, fields) if return_fields else psf


def point_sources_psf(mask, basis, coeffs, wavelengths, spectrum_weights, positions,
                      fluxes, diameter, npixels_out, pixel_scale_out):
    """PointSources.model: weights normalised, x flux, sum over sources.  sources.py:392-411."""
    w = np.asarray(spectrum_weights, np.float64)
    w = w / w.sum()
    out = 0.0
    for pos, flux in zip(positions, fluxes):
        out = out + poly_psf(mask, basis, coeffs, wavelengths, w * flux, diameter,
                             npixels_out, pixel_scale_out, offset=pos)
    return out


# ---------------------------------------------------------------------------
# Gradient of the phase-retrieval loss (JAX AD of the chain above, written out)
# ---------------------------------------------------------------------------


def gaussian_loss(psf, data, flux):
    """0.5 * sum((flux * psf - data)^2 / max(data, 1))."""
    var = np.maximum(data, 1.0)
    return 0.5 * np.sum((flux * psf - data) ** 2 / var)


def poly_psf_vjp(mask, basis, coeffs, wavelengths, weights, diameter, npixels_out,
                 pixel_scale_out, G, offset=(0.0, 0.0), normalise_power=True):
    """(dL/dOPD, dL/dphase) of poly_psf for the cotangent G = dL/dPSF, in FP64.

    With P_l = A exp(i k_l OPD) and E_l = sqrt(w_l) s Y^T P_l X:
      Ebar_l = 2 G E_l
      Pbar_l = sqrt(w_l) s conj(Y) Ebar_l X^H
      dL/dOPD = sum_l k_l Im(conj(P_l) Pbar_l)   (real-cotangent convention)
      dL/dphase = sum_l Im(conj(P_l) Pbar_l)
    The reference differentiates the same chain by JAX reverse mode
    (optical_systems.py:147-223, utils/propagation.py:178-256,
    wavefronts.py:332-367); this is its closed form.
    """
    wavelengths = np.atleast_1d(np.asarray(wavelengths, np.float64))
    weights = np.atleast_1d(np.asarray(weights, np.float64))
    n = mask.shape[-1]
    d = diameter / n
    g_opd = np.zeros((n, n))
    g_phase = np.zeros((n, n))
    for wl, w in zip(wavelengths, weights):
        P = pupil_field(wl, mask, basis, coeffs, diameter, offset, normalise_power)
        Y = transfer_matrix(wl, n, d, npixels_out, pixel_scale_out)
        s = np.exp(np.log(calc_nfringes(wl, n, d, npixels_out, pixel_scale_out))
                   - (np.log(n) + np.log(npixels_out)))
        E = (w ** 0.5) * s * (Y.T @ P @ Y)
        Ebar = 2.0 * G * E
        Pbar = (w ** 0.5) * s * (np.conj(Y) @ Ebar @ np.conj(Y).T)
        h = np.imag(np.conj(P) * Pbar)
        g_opd += (2 * np.pi / wl) * h
        g_phase += h
    return g_opd, g_phase


def poly_psf_grad(mask, basis, coeffs, wavelengths, weights, diameter, npixels_out,
                  pixel_scale_out, data, flux):
    """(loss, dloss/dcoeffs) of the Gaussian loss by the adjoint chain, in FP64.

    G = dL/dPSF = flux (flux PSF - data) / max(data, 1); then poly_psf_vjp and
    dL/dc_j = sum_px Z_j dL/dOPD.
    """
    psf = poly_psf(mask, basis, coeffs, wavelengths, weights, diameter, npixels_out,
                   pixel_scale_out)
    var = np.maximum(data, 1.0)
    loss = 0.5 * np.sum((flux * psf - data) ** 2 / var)
    G = flux * (flux * psf - data) / var
    g_opd, _ = poly_psf_vjp(mask, basis, coeffs, wavelengths, weights, diameter, npixels_out,
                            pixel_scale_out, G)
    grad = np.tensordot(basis, g_opd, axes=([1, 2], [0, 1]))
    return loss, grad


# ---------------------------------------------------------------------------
# Angular spectrum (textbook; abcdLux.asm.asm_prop is absent -> parity unpinned)
# ---------------------------------------------------------------------------


def asm_prop(u_in, dx, wavelength, z, npad):
    """U_z = IFFT{ FFT{pad(U)} * exp(i 2 pi z sqrt(1/l^2 - fx^2 - fy^2)) }, evanescent -> 0.

    Replaces layers/abcd_propagators.py:444-451 (abcdLux.asm.asm_prop, crop=False).
    Centred pad (utils/array_ops.py:12-50 semantics), centred FFT (ifftshift before,
    fftshift after), output keeps the input sampling dx and has npad pixels.
    """
    u = pad_to(np.asarray(u_in, np.complex128), npad)
    U = np.fft.fft2(np.fft.ifftshift(u))
    f = np.fft.fftfreq(npad, d=dx)
    FX, FY = np.meshgrid(f, f)
    arg = 1.0 / wavelength ** 2 - FX ** 2 - FY ** 2
    H = np.where(arg > 0, np.exp(2j * np.pi * z * np.sqrt(np.maximum(arg, 0.0))), 0.0)
    return np.fft.fftshift(np.fft.ifft2(U * H))


# ---------------------------------------------------------------------------
# Linear canonical transforms  (layers/abcd_propagators.py:23-383; the
# transforms themselves live in abcdLux, absent: restated from the Collins
# integral -- parity unpinned against abcdLux, pinned by closed forms)
# ---------------------------------------------------------------------------


def abcd_free_space(z):
    return np.array([[1.0, float(z)], [0.0, 1.0]])


def abcd_lens(f):
    return np.array([[1.0, 0.0], [-1.0 / float(f), 1.0]])


def abcd_mirror(radius):
    return np.array([[1.0, 0.0], [-2.0 / float(radius), 1.0]])


def abcd_conjugate_plane(f):
    return np.array([[0.0, float(f)], [-1.0 / float(f), 0.0]])


def compose_abcd(mats):
    """Elements in propagation order: M = M_n ... M_2 M_1."""
    out = np.eye(2)
    for m in mats:
        out = np.asarray(m, float) @ out
    return out


def lct_kernel(wavelength, xs_in, xs_out, abcd):
    """1-D Collins kernel K[l, k] = sqrt(dx dx' / (i lambda B))
    exp(i pi / (lambda B) (A x_k^2 - 2 x_k x'_l + D x'_l^2)) (B != 0)."""
    A, B, D = abcd[0][0], abcd[0][1], abcd[1][1]
    x = np.asarray(xs_in, np.float64)[None, :]
    xp = np.asarray(xs_out, np.float64)[:, None]
    dx = xs_in[1] - xs_in[0] if len(xs_in) > 1 else 1.0
    dxp = xs_out[1] - xs_out[0] if len(xs_out) > 1 else 1.0
    pref = np.sqrt(abs(dx * dxp) / (wavelength * abs(B))) * np.exp(-1j * np.pi / 4 * np.sign(B))
    return pref * np.exp(1j * np.pi / (wavelength * B) * (A * x ** 2 - 2 * x * xp + D * xp ** 2))


def lct(u_in, wavelength, xs_in, xs_out, abcd):
    """The 2-D separable LCT of a field sampled on xs_in (both axes) to xs_out:
    u_out = K u_in K^T (the Collins integral as a Riemann sum, prefactor
    dx dx' / (i lambda B))."""
    K = lct_kernel(wavelength, xs_in, xs_out, abcd)
    return K @ np.asarray(u_in, np.complex128) @ K.T


def lct_fft(u_in, wavelength, dx, c_in, c_out, npad, abcd):
    """The same transform sampled on the FFT grid of npad points,
    d_out = lambda |B| / (npad dx), centred on c_out, computed as chirp ->
    centred padded FFT -> chirp.  Returns (field, d_out)."""
    A, B, D = abcd[0][0], abcd[0][1], abcd[1][1]
    n = u_in.shape[-1]
    sg = np.sign(B)
    lb = wavelength * B
    d_out = wavelength * abs(B) / (npad * dx)
    xs = spec_xs(n, dx, c_in)
    xo = spec_xs(npad, d_out, c_out)
    # the centred FFT pairs offsets from index npad // 2; the grids are
    # symmetric about (npad - 1) / 2: delta apart
    de = npad // 2 - (npad - 1) / 2
    t_in = (A / (2 * lb) * xs ** 2 + (-c_out / lb - sg * de / (npad * dx)) * xs
            + c_out * c_in / lb + sg * de * c_in / (npad * dx) + sg * de ** 2 / npad)
    t_out = (D / (2 * lb) * xo ** 2 + (-c_in / lb - sg * de / (npad * d_out)) * xo
             + c_in * c_out / lb + sg * de * c_out / (npad * d_out) + sg * de ** 2 / npad
             - c_in * c_out / lb - sg * de ** 2 / npad - sg / 8)
    q_in = np.exp(2j * np.pi * t_in)
    q_out = np.exp(2j * np.pi * t_out)
    u = np.asarray(u_in, np.complex128) * q_in[:, None] * q_in[None, :]
    u = pad_to(u, npad)
    if sg > 0:
        F = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(u))) / npad
    else:
        F = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(u))) * npad
    return F * q_out[:, None] * q_out[None, :], d_out


def propagate_fft(phasor, wavelength, pixel_scale, center=0.0, pad=2, focal_length=None,
                  spec_c=None, inverse=False):
    """Wavefront.propagate_FFT with an optional output centre.  wavefronts.py:654-727.

    Returns (phasor, pixel_scale, center).
    """
    n_in = phasor.shape[-1]
    n_out = n_in * pad
    d_fft, c_fft = fft_spec(n_out, pixel_scale, wavelength, focal_length)
    if spec_c is not None:
        in_ramp = fft_phase_ramp(spec_xs(n_in, pixel_scale, center), wavelength,
                                 c_fft - spec_c, focal_length, inverse)
        shift = fft_spec(n_out, d_fft, wavelength, focal_length)[1]
        out_ramp = fft_phase_ramp(spec_xs(n_out, d_fft, spec_c), wavelength, shift,
                                  focal_length, inverse)
        c_out = spec_c
    else:
        in_ramp, out_ramp, c_out = 1.0, 1.0, c_fft
    out, ps = fft(np.asarray(phasor) * in_ramp, wavelength, pixel_scale, focal_length, pad,
                  inverse)
    return out * out_ramp, ps, c_out


# ---------------------------------------------------------------------------
# Detector layers  (src/dLux/layers/detector_layers.py, src/dLux/psfs.py,
# src/dLux/utils/math.py, src/dLux/utils/array_ops.py)
# ---------------------------------------------------------------------------


def gaussian(mean=0.0, std=1.0, npixels=64, extent=5.0):
    """Separable normalised Gaussian on linspace(-extent, extent, n) per axis,
    pdf with loc=mean, scale=std, outer product in 'xy' order, divided by its
    sum.  utils/math.py:20-67 (the grid spans +-extent in the pdf's own units)."""
    mean = np.broadcast_to(np.asarray(mean, np.float64), (2,))
    std = np.broadcast_to(np.asarray(std, np.float64), (2,))
    npix = (npixels, npixels) if np.ndim(npixels) == 0 else tuple(npixels)
    axes = [np.linspace(-extent, extent, n) for n in npix]
    g1 = [np.exp(-0.5 * ((a - m) / s) ** 2) / (s * np.sqrt(2 * np.pi))
          for a, m, s in zip(axes, mean, std)]
    kernel = np.array(np.meshgrid(*g1, indexing="xy")).prod(0)
    return kernel / np.sum(kernel)


def jitter_kernel(sigma, kernel_size=9, oversample=3):
    """ApplyJitter.kernel: gaussian on kernel_size*oversample pixels, then a
    sum-downsample by oversample.  layers/detector_layers.py:180-196."""
    k = gaussian(0.0, [sigma, sigma], kernel_size * oversample)
    return downsample(k, oversample, mean=False)


def convolve_same(data, kernel):
    """Direct 2-D convolution, output the size of `data`, zero padded:
    out[i, j] = sum_{a,b} data[i + (K0-1)//2 - a, j + (K1-1)//2 - b] kernel[a, b]
    (scipy / jax.scipy.signal.convolve mode="same", method="direct", the call
    of PSF.convolve, psfs.py:93-110)."""
    data = np.asarray(data, np.float64)
    kernel = np.asarray(kernel, np.float64)
    n0, n1 = data.shape
    k0, k1 = kernel.shape
    o0, o1 = (k0 - 1) // 2, (k1 - 1) // 2
    padded = np.zeros((n0 + k0 - 1, n1 + k1 - 1))
    # padded[u, v] = data[u - (k0 - 1 - o0), v - (k1 - 1 - o1)]
    padded[k0 - 1 - o0:k0 - 1 - o0 + n0, k1 - 1 - o1:k1 - 1 - o1 + n1] = data
    out = np.zeros((n0, n1))
    for a in range(k0):
        for b in range(k1):
            # data index i + o0 - a  ->  padded index i + (k0 - 1) - a
            out += kernel[a, b] * padded[k0 - 1 - a:k0 - 1 - a + n0, k1 - 1 - b:k1 - 1 - b + n1]
    return out


def correlate_same_adjoint(g, kernel):
    """Adjoint of convolve_same in its first argument:
    din[q] = sum_{a,b} g[q0 - (K0-1)//2 + a, q1 - (K1-1)//2 + b] kernel[a, b]."""
    g = np.asarray(g, np.float64)
    kernel = np.asarray(kernel, np.float64)
    n0, n1 = g.shape
    k0, k1 = kernel.shape
    o0, o1 = (k0 - 1) // 2, (k1 - 1) // 2
    padded = np.zeros((n0 + k0 - 1, n1 + k1 - 1))
    padded[o0:o0 + n0, o1:o1 + n1] = g
    out = np.zeros((n0, n1))
    for a in range(k0):
        for b in range(k1):
            out += kernel[a, b] * padded[a:a + n0, b:b + n1]
    return out


def detector(psf, layers):
    """LayeredDetector.__call__ (detectors.py:103-128) over a list of layers:
    ("pixel_response", map)  psf * map                  detector_layers.py:130-131
    ("convolve", kernel)     convolve_same(psf, kernel) detector_layers.py:198-199
    ("convolve_separable", (ky, kx))  the same with kernel = outer(ky, kx)
    ("downsample", k)        block sums                 detector_layers.py:292-293
    ("saturate", t)          minimum(psf, t)            detector_layers.py:228-229
    ("add", c)               psf + c                    detector_layers.py:258-259"""
    x = np.asarray(psf, np.float64)
    for layer in layers:
        kind = layer[0]
        if kind == "pixel_response":
            x = x * np.asarray(layer[1], np.float64)
        elif kind == "convolve":
            x = convolve_same(x, layer[1])
        elif kind == "convolve_separable":
            x = convolve_same(x, np.outer(layer[1][0], layer[1][1]))
        elif kind == "downsample":
            x = downsample(x, int(layer[1]), mean=False)
        elif kind == "saturate":
            x = np.minimum(x, float(layer[1]))
        elif kind == "add":
            x = x + float(layer[1])
        else:
            raise ValueError(f"unknown detector layer {kind!r}")
    return x


def detector_vjp(psf, layers, g):
    """(d detector / d psf)^T g, the reverse pass JAX builds: adjoint convolution
    (correlation), block replication, map multiply, minimum's derivative
    (1 below, 1/2 at, 0 above the threshold), identity for the constant."""
    xs = [np.asarray(psf, np.float64)]
    for layer in layers:
        xs.append(detector(xs[-1], [layer]))
    g = np.asarray(g, np.float64)
    for layer, x in zip(reversed(layers), reversed(xs[:-1])):
        kind = layer[0]
        if kind == "pixel_response":
            g = g * np.asarray(layer[1], np.float64)
        elif kind == "convolve":
            g = correlate_same_adjoint(g, layer[1])
        elif kind == "convolve_separable":
            g = correlate_same_adjoint(g, np.outer(layer[1][0], layer[1][1]))
        elif kind == "downsample":
            k = int(layer[1])
            g = np.repeat(np.repeat(g, k, axis=0), k, axis=1)
        elif kind == "saturate":
            t = float(layer[1])
            g = g * np.where(x < t, 1.0, np.where(x == t, 0.5, 0.0))
    return g


# ---------------------------------------------------------------------------
# Dynamic apertures  (src/dLux/layers/apertures.py, src/dLux/utils/geometry.py,
# src/dLux/coordinates.py, src/dLux/utils/coordinates.py, utils/zernikes.py)
# ---------------------------------------------------------------------------


def gen_powers(degree):
    """(x powers, y powers) of the 2-D monomials (utils/coordinates.py:104-133)."""
    n = degree * (degree + 1) // 2
    vals = np.arange(n)
    tri = lambda k: k * (k + 1) // 2
    ypows = vals - np.repeat(tri(np.arange(degree)), np.arange(1, degree + 1))
    xpows = np.flip(vals - np.repeat(n - np.flip(tri(np.arange(1, degree + 1))),
                                     np.arange(degree, 0, -1)))
    return np.array([xpows, ypows])


def distort_coords(coords, coeffs, pows):
    """coords + sum_m coeffs[:, m] x^px_m y^py_m (utils/coordinates.py:136-163)."""
    coords = np.asarray(coords, np.float64)
    pows = np.asarray(pows)
    basis = np.prod(coords[:, None] ** pows.reshape(pows.shape + (1,) * (coords.ndim - 1)), 0)
    return coords + np.tensordot(np.asarray(coeffs, np.float64), basis, axes=(-1, 0))


def coord_transform(coords, translation=None, rotation=None, compression=None, shear=None,
                    distortion=None, order=None):
    """A CoordTransform, or with `distortion` (and `order`) a DistortedCoords
    (coordinates.py:230-357)."""
    if distortion is not None:
        return distort_coords(coords, distortion, gen_powers(order + 1)[:, 1:])
    return _coord_transform(coords, translation, rotation, compression, shear)


def _coord_transform(coords, translation=None, rotation=None, compression=None, shear=None):
    """CoordTransform.__call__: translation, shear, compression, rotation, in
    that order (coordinates.py:306-316).  Shear adds each coordinate plane's
    transpose times the factor (utils/coordinates.py:61-78)."""
    coords = np.asarray(coords, np.float64)
    if translation is not None:
        coords = coords - np.asarray(translation, float)[:, None, None]
    if shear is not None:
        coords = coords + np.transpose(coords, (0, 2, 1)) * np.asarray(shear, float)[:, None, None]
    if compression is not None:
        coords = coords * np.asarray(compression, float)[:, None, None]
    if rotation is not None:
        coords = rotate_coords(coords, float(rotation))
    return coords


def soften(distances, clip_dist, invert=False):
    """Clip to +-clip_dist and rescale by the extrema; a constant field gives
    the support d > 0 (utils/geometry.py:49-92)."""
    d = np.asarray(distances, np.float64)
    if invert:
        d = -d
    d = np.clip(d, -clip_dist, clip_dist)
    lo, hi = d.min(), d.max()
    if hi == lo:
        return (d > 0).astype(float)
    return (d - lo) / (hi - lo)


def circ_distance(coords, radius):
    """utils/geometry.py:402-418."""
    return cart2polar(coords)[0] - radius


def square_distance(coords, width):
    """utils/geometry.py:421-437."""
    return np.max(np.abs(coords), axis=0) - width / 2


def rectangle_distance(coords, width, height):
    """utils/geometry.py:440-460."""
    return np.maximum(np.abs(coords[0]) - width / 2, np.abs(coords[1]) - height / 2)


def spider_distance(coords, width, angle):
    """One arm: max(|x'| - w/2, y') after rotating by the arm angle
    (utils/geometry.py:463-484)."""
    x, y = rotate_coords(coords, np.deg2rad(angle))
    return np.maximum(np.abs(x) - width / 2, y)


def reg_polygon_distance(coords, nsides, radius):
    """Max over edges of the signed line distance; edge slopes by nandiv
    (+inf where the edge is vertical: its default fill), the vertical rule for
    an infinite slope and the horizontal rule for slope 0
    (utils/geometry.py:487-579, utils/math.py:228-250)."""
    t = np.linspace(-np.pi, np.pi, nsides, endpoint=False)
    x, y = radius * np.cos(t), radius * np.sin(t)
    j = np.roll(np.arange(nsides), -1)
    out = None
    for i in range(nsides):
        den = x[j[i]] - x[i]
        m = np.inf if den == 0 else (y[j[i]] - y[i]) / den
        if m == np.inf:
            val = -np.sign(x[i]) * (x[i] - coords[0])
        elif m == 0:
            val = -np.sign(y[i]) * (y[i] - coords[1])
        else:
            dist = (m * coords[0] - coords[1] + y[i] - m * x[i]) / np.sqrt(m ** 2 + 1)
            val = -np.sign(y[i] - m * x[i]) * dist
        out = val if out is None else np.maximum(out, val)
    return out


def soft_aperture(kind, coords, pixel_scale, softening=1.0, occulting=False, transforms=(),
                  **shape):
    """DynamicAperture.transmission for kind in circle / square / rectangle /
    reg_polygon / spider (layers/apertures.py:299-641): transforms applied in
    order, clip = pixel_scale * softening / 2, distances negated (positive
    inside) then softened, inverted when occulting."""
    for t in transforms:
        coords = coord_transform(coords, **t)
    clip = pixel_scale * softening / 2
    if kind == "circle":
        d = -circ_distance(coords, shape["radius"])
    elif kind == "square":
        d = -square_distance(coords, shape["width"])
    elif kind == "rectangle":
        d = -rectangle_distance(coords, shape["width"], shape["height"])
    elif kind == "reg_polygon":
        d = -reg_polygon_distance(coords, shape["nsides"], shape["rmax"])
    elif kind == "spider":
        d = -np.min([spider_distance(coords, shape["width"], a)
                     for a in np.atleast_1d(shape["angles"])], axis=0)
    else:
        raise ValueError(kind)
    return soften(d, clip, occulting)


def polike(nsides, j, coordinates, diameter=2.0):
    """Zernike on an n-sided aperture: (1/r_a) Z_j(coords / r_a) with
    r_a = cos(pi/n) / cos(u), u the angle within the wedge (utils/zernikes.py:318-349)."""
    coords = np.asarray(coordinates, np.float64) / (diameter / 2)
    theta = cart2polar(coords)[1]
    alpha = np.pi / nsides
    phi = theta + alpha
    wedge = np.floor((phi + alpha) / (2.0 * alpha))
    u = phi - wedge * (2 * alpha)
    r_alpha = np.cos(alpha) / np.cos(u)
    return 1 / r_alpha * zernike(j, coords / r_alpha)


# ---------------------------------------------------------------------------
# Resampling  (utils/interpolation.py:13-108; interpax.interp2d is absent:
# restated as the textbook bilinear / nearest rules, pinned against scipy)
# ---------------------------------------------------------------------------


def interp_grid(image, n, d, c, xs, ys, method="linear", fill=0.0):
    """`image` (n x n, knots at c + (k - (n-1)/2) d on both axes) sampled at
    (xs, ys); fill outside the knots' span."""
    image = np.asarray(image, np.float64)
    h = (n - 1) / 2.0
    u = (np.asarray(xs, np.float64) - c) / d + h
    v = (np.asarray(ys, np.float64) - c) / d + h
    inside = (u >= 0) & (u <= n - 1) & (v >= 0) & (v <= n - 1)
    uc, vc = np.clip(u, 0, n - 1), np.clip(v, 0, n - 1)
    if method == "nearest":
        # the nearer knot; at an exact half-pixel tie the upper one (interpax's
        # nearest rule keeps the right neighbour unless the left is strictly nearer)
        rn = np.minimum(np.floor(vc + 0.5).astype(int), n - 1)
        cn = np.minimum(np.floor(uc + 0.5).astype(int), n - 1)
        out = image[rn, cn]
    else:
        c0 = np.minimum(np.floor(uc).astype(int), max(n - 2, 0))
        r0 = np.minimum(np.floor(vc).astype(int), max(n - 2, 0))
        c1, r1 = np.minimum(c0 + 1, n - 1), np.minimum(r0 + 1, n - 1)
        tu, tv = uc - c0, vc - r0
        top = image[r0, c0] + (image[r0, c1] - image[r0, c0]) * tu
        bot = image[r1, c0] + (image[r1, c1] - image[r1, c0]) * tu
        out = top + (bot - top) * tv
    return np.where(inside, out, fill)
