"""FFT output sampling and centring ramps (utils/fourier.py:14-78)."""

from __future__ import annotations

import numpy as np
import torch

from .. import runtime

__all__ = ["fft_spec", "fft_phase_ramp"]


def fft_spec(npixels_in, pixel_scale_in, wavelength, focal_length=None):
    """Native FFT output pixel scale and centre offset.  utils/fourier.py:14-45."""
    d_out = wavelength / (npixels_in * pixel_scale_in)
    if focal_length is not None:
        d_out = d_out * focal_length
    if npixels_in % 2 != 0:
        return d_out, 0.0
    return d_out, -0.5 * d_out


def fft_phase_ramp_1d(xs, wavelength, shift, focal_length=None, inverse=False) -> torch.Tensor:
    """The 1-D factor exp(+/-2 pi i xs shift / alpha) of fft_phase_ramp (complex64, on the
    device); the phase is formed in FP64 on the host."""
    sign = -1 if inverse else 1
    alpha = wavelength if focal_length is None else wavelength * focal_length
    xs = xs.detach().cpu().numpy() if isinstance(xs, torch.Tensor) else np.asarray(xs)
    ramp = np.exp(sign * 2j * np.pi * xs.astype(np.float64) * shift / alpha)
    return torch.as_tensor(ramp.astype(np.complex64), device=runtime.device())


def fft_phase_ramp(xs, wavelength, shift, focal_length=None, inverse=False) -> torch.Tensor:
    """Separable ramp exp(+/-2 pi i xs shift / alpha) as an (n, n) outer product.

    utils/fourier.py:48-78.  The phase is formed in FP64 on the host (n values)
    and the outer product on the device.
    """
    r = fft_phase_ramp_1d(xs, wavelength, shift, focal_length, inverse)
    return r[None, :] * r[:, None]
