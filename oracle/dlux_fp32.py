"""FP32 NumPy restatement of dLux's polychromatic PSF path: the CPU baseline.

TEST INFRASTRUCTURE ONLY.  This module is timed by ``bench.py``'s
``cpu_baseline`` leg (as "CPU restatement of dLux", not dLux itself, which
needs JAX) and checked by ``tests/``; the product package never imports it.
The FP64 oracle (``dlux_oracle``) stays the error reference.

It follows dLux's default (x64 off) arithmetic and operation order:

  * ``Wavefront.__init__``: phasor = ones / N^2, complex64 (wavefronts.py:111-113)
  * tilt by the source offset: opd = theta . (X, Y), phasor *= exp(i k opd)
    (wavefronts.py:370-395, optical_systems.py:363-389)
  * ``TransmissiveLayer``: phasor *= mask, then power normalisation
    (layers/optical_layers.py:181-186, wavefronts.py:397-424)
  * ``BasisLayer``: opd = tensordot(basis, coeffs) once per PSF (it does not
    depend on the wavelength), phasor *= exp(i k opd)
    (layers/optical_layers.py:319-327, utils/math.py:177-196)
  * ``MFT``: FP32 transfer matrices, (Y^T E) X, times exp(log nf - log N_in -
    log N_out) (utils/propagation.py:67-127, 229-254)
  * ``propagate``: E * sqrt(w), |E|^2 summed over wavelengths
    (optical_systems.py:211-223)

with the wavelength loop batched the way ``filter_vmap`` batches it: one
stacked (n_lambda, N, N) complex64 array per step, matmuls broadcast over it.
The gradient (phase retrieval) is the same chain's reverse mode written out
in FP32 (forward, then Ebar = 2 G E, Pbar = conj(Y) Ebar X^H, dL/dOPD =
sum_l k_l Im(conj(P_l) Pbar_l), dL/dc = Z^T dL/dOPD), which is what JAX's
reverse mode computes for this chain.
"""

from __future__ import annotations

import numpy as np

F32 = np.float32
C64 = np.complex64


def _cis(arg):
    """exp(i arg) for a float32 array, as complex64 (vectorised cos and sin)."""
    out = np.empty(arg.shape, C64)
    np.cos(arg, out=out.real)
    np.sin(arg, out=out.imag)
    return out


def _coords_1d(n, scale, offset=0.0):
    """nd_coords' pixel_fn in float32 (utils/coordinates.py:320-324)."""
    start = F32(-(n - 1) / 2 * scale - offset)
    end = F32((n - 1) / 2 * scale - offset)
    return np.linspace(start, end, n, dtype=F32)


def transfer_matrices(wavelengths, n_in, pixel_scale_in, n_out, pixel_scale_out, shift=0.0):
    """(n_lambda, N_in, N_out) complex64 exp(-2 pi i a (x) b), FP32 phase arguments
    (utils/propagation.py:67-127)."""
    wl = np.asarray(wavelengths, F32)
    fringe = wl / F32(pixel_scale_in * n_in)
    in_vec = _coords_1d(n_in, 1.0 / n_in, shift / n_in)
    scale_out = F32(pixel_scale_out) / fringe
    base = (np.arange(n_out, dtype=F32) - F32((n_out - 1) / 2))
    out_vec = (base[None, :] - F32(shift)) * scale_out[:, None]
    arg = F32(-2 * np.pi) * in_vec[None, :, None] * out_vec[:, None, :]
    return _cis(arg)


def mft_norm(wavelengths, n_in, pixel_scale_in, n_out, pixel_scale_out):
    """exp(log nfringes - (log N_in + log N_out)) in float32 (utils/propagation.py:130-175, :254)."""
    wl = np.asarray(wavelengths, F32)
    nf = F32(n_out * pixel_scale_out) / (wl / F32(n_in * pixel_scale_in))
    return np.exp(np.log(nf) - (np.log(F32(n_in)) + np.log(F32(n_out)))).astype(F32)


def pupil_fields(wavelengths, mask, opd, diameter, offset=(0.0, 0.0), normalise=True):
    """(n_lambda, N, N) complex64 pupil fields: init, tilt, mask, normalise, OPD."""
    n = mask.shape[-1]
    d = diameter / n
    wl = np.asarray(wavelengths, F32)
    k = (F32(2 * np.pi) / wl)[:, None, None]
    E = np.full((len(wl), n, n), 1.0 / n ** 2, dtype=C64)
    if offset[0] or offset[1]:
        xs = _coords_1d(n, d)
        X, Y = np.meshgrid(xs, xs)
        tilt = F32(offset[0]) * X + F32(offset[1]) * Y
        E = E * _cis(k * tilt[None])
    E = E * mask.astype(F32)[None]
    if normalise:
        power = np.sum(np.abs(E) ** 2, axis=(-2, -1), keepdims=True)
        E = E * np.sqrt(F32(1.0) / power).astype(F32)
    if opd is not None:
        E = E * _cis(k * opd[None])
    return E


def eval_basis(basis, coeffs):
    """tensordot(basis, coeffs) in float32 (utils/math.py:177-196)."""
    return np.tensordot(np.asarray(coeffs, F32), basis, axes=(0, 0)).astype(F32)


def poly_psf(mask, basis, coeffs, wavelengths, weights, diameter, n_out, pixel_scale_out,
             offset=(0.0, 0.0)):
    """sum_l w_l |MFT(P_l)|^2 in FP32 (optical_systems.py:147-223)."""
    wl = np.asarray(wavelengths, F32)
    w = np.asarray(weights, F32)
    n = mask.shape[-1]
    opd = eval_basis(basis, coeffs) if basis is not None else None
    P = pupil_fields(wl, mask, opd, diameter, offset)
    T = transfer_matrices(wl, n, diameter / n, n_out, pixel_scale_out)
    E = np.matmul(np.matmul(T.transpose(0, 2, 1), P), T)
    E = E * (mft_norm(wl, n, diameter / n, n_out, pixel_scale_out) * np.sqrt(w))[:, None, None]
    return np.sum(np.abs(E) ** 2, axis=0)


def poly_psf_grad(mask, basis, coeffs, wavelengths, weights, diameter, n_out, pixel_scale_out,
                  data, flux):
    """(loss, dloss/dcoeffs) of 0.5 sum (flux psf - d)^2 / max(d, 1) in FP32:
    the forward chain, then its reverse mode (the adjoint contractions)."""
    wl = np.asarray(wavelengths, F32)
    w = np.asarray(weights, F32)
    n = mask.shape[-1]
    opd = eval_basis(basis, coeffs)
    P = pupil_fields(wl, mask, opd, diameter)
    T = transfer_matrices(wl, n, diameter / n, n_out, pixel_scale_out)
    a = (mft_norm(wl, n, diameter / n, n_out, pixel_scale_out) * np.sqrt(w))[:, None, None]
    E = a * np.matmul(np.matmul(T.transpose(0, 2, 1), P), T)
    psf = np.sum(np.abs(E) ** 2, axis=0)
    d = np.asarray(data, F32)
    var = np.maximum(d, F32(1.0))
    resid = F32(flux) * psf - d
    loss = F32(0.5) * np.sum(resid ** 2 / var)
    G = F32(flux) * resid / var
    Eb = F32(2.0) * G[None] * E
    Tc = np.conj(T)
    Pb = a * np.matmul(np.matmul(Tc, Eb), Tc.transpose(0, 2, 1))
    k = (F32(2 * np.pi) / wl)[:, None, None]
    g_opd = np.sum(k * np.imag(np.conj(P) * Pb), axis=0)
    return loss, np.tensordot(basis, g_opd, axes=([1, 2], [0, 1])).astype(F32)


def mft_intensity(field, wavelength, pixel_scale_in, n_out, pixel_scale_out, shift=(0.0, 0.0)):
    """w-less |MFT(field)|^2 of one (source, wavelength) instance with an output
    shift in pixels (the uneven batch's unit of work)."""
    n = field.shape[-1]
    X = transfer_matrices([wavelength], n, pixel_scale_in, n_out, pixel_scale_out, shift[0])[0]
    Y = transfer_matrices([wavelength], n, pixel_scale_in, n_out, pixel_scale_out, shift[1])[0]
    E = (Y.T @ field.astype(C64)) @ X
    E = E * mft_norm([wavelength], n, pixel_scale_in, n_out, pixel_scale_out)[0]
    return np.abs(E) ** 2
