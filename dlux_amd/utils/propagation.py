"""MFT / FFT propagation utilities (utils/propagation.py:8-256) on the HIP path.

Inputs may be single fields (N, N) or a leading batch (B, N, N) with per-
instance wavelengths / pixel scales -- the batched form is what dLux obtains by
vmapping these functions over wavelength (optical_systems.py:213-216).
"""

from __future__ import annotations

import numpy as np
import torch

from .. import runtime

__all__ = ["FFT", "MFT", "transfer_matrix", "calc_nfringes"]


def FFT(phasor, wavelength, pixel_scale, focal_length=None, pad: int = 2,
        inverse: bool = False):
    """Centred, unitary-normalised FFT with zero padding.  utils/propagation.py:8-64.

    Returns (phasor, new_pixel_scale); the output has npixels * pad pixels.
    """
    f = runtime.as_complex(phasor)
    npixels = f.shape[-1]
    wl = np.asarray(wavelength, np.float64)
    fringe_size = wl / (np.asarray(pixel_scale, np.float64) * npixels)
    new_pixel_scale = fringe_size / pad
    if focal_length is not None:
        new_pixel_scale = new_pixel_scale * focal_length
    npad = npixels + 2 * ((npixels * (pad - 1)) // 2)
    out = runtime.fft2(f, npad, inverse=inverse)
    return out, new_pixel_scale


def transfer_matrix(wavelength, npixels_in, pixel_scale_in, npixels_out, pixel_scale_out,
                    shift=0.0, focal_length=None, focal_shift=0.0, inverse=False):
    """(N_in, N_out) complex64 DFT matrix exp(-/+2 pi i a (x) b).  utils/propagation.py:67-127."""
    W = runtime.transfer_matrix_T(np.atleast_1d(wavelength), int(npixels_in), pixel_scale_in,
                                  int(npixels_out), pixel_scale_out, shifts=shift,
                                  focal_length=focal_length, focal_shift=focal_shift,
                                  inverse=inverse)
    W = W.transpose(-1, -2)
    return W[0] if np.ndim(wavelength) == 0 else W


def calc_nfringes(wavelength, npixels_in, pixel_scale_in, npixels_out, pixel_scale_out,
                  focal_length=None, focal_shift=0.0):
    """Number of fringes across the output plane.  utils/propagation.py:130-175."""
    diameter = npixels_in * np.asarray(pixel_scale_in, np.float64)
    fringe_size = np.asarray(wavelength, np.float64) / diameter
    output_size = npixels_out * np.asarray(pixel_scale_out, np.float64)
    if focal_length is not None:
        output_size = output_size / (focal_length + focal_shift)
    return output_size / fringe_size


def MFT(phasor, wavelength, pixel_scale_in, npixels_out, pixel_scale_out, focal_length=None,
        shift=(0.0, 0.0), pixel: bool = True, inverse: bool = False, field_scale=None,
        precision: str | None = None) -> torch.Tensor:
    """Two-sided matrix Fourier transform.  utils/propagation.py:178-256.

    ``shift`` is (2,) or (B, 2) in output pixels (``pixel=True``) or in output
    pixel-scale units.  ``field_scale`` (scalar or (B,)) multiplies the result.
    """
    shift = np.asarray(shift, np.float64)
    pout = np.asarray(pixel_scale_out, np.float64)
    if not pixel:
        shift = shift / (pout[..., None] if pout.ndim else pout)
    sx = shift[..., 0]
    sy = shift[..., 1]
    return runtime.mft(phasor, wavelength, pixel_scale_in, int(npixels_out), pout,
                       focal_length=focal_length, shift_x=sx, shift_y=sy, inverse=inverse,
                       field_scale=field_scale, precision=precision)
