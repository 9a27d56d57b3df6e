"""Detector layers (layers/detector_layers.py:23-293): PSF -> PSF transforms.

Each layer is one op of the HIP detector chain (dlx_detector_apply): applied
alone it is one launch; a ``LayeredDetector`` hands its whole layer list to
the chain, which fuses the elementwise layers into the convolution and
downsampling passes.  Gradients flow to the PSF through dlx_detector_vjp.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import runtime
from ..base import Base
from ..psfs import PSF
from ..utils.math import gaussian

__all__ = ["DetectorLayer", "ApplyInterpolation", "ApplyPixelResponse", "ApplyJitter",
           "ApplySaturation", "AddConstant", "Downsample"]


class DetectorLayer(Base):
    """Base detector layer (detector_layers.py:23-65)."""

    def __call__(self, psf: PSF) -> PSF:  # pragma: no cover - abstract
        raise NotImplementedError

    def apply(self, psf: PSF) -> PSF:
        """Backwards-compatible alias of __call__ (detector_layers.py:51-65)."""
        return self(psf)

    # (op name, argument) of the HIP chain, or None if the layer has no chain form
    def _det_op(self):
        return None

    # factor applied to the PSF pixel scale
    def _scale_factor(self) -> int:
        return 1

    def _run(self, psf: PSF) -> PSF:
        data = runtime.detector(psf.data, [self._det_op()])
        return PSF(data, psf.pixel_scale * self._scale_factor())

    def _device(self, key, make):
        """A device tensor built by make(), kept while `key` stays the same
        (host arrays are uploaded once, not on every call).  `key` holds the
        objects or values the tensor was made from; a layer updated through
        set() gets new ones and so a new tensor."""
        cached = self.__dict__.get("_dev")
        if cached is not None and len(cached[0]) == len(key) and all(
                a is b or (np.isscalar(a) and a == b) for a, b in zip(cached[0], key)):
            return cached[1]
        t = make()
        self.__dict__["_dev"] = (key, t)
        return t


def _scalar(v):
    """A scalar layer parameter: a torch tensor is kept (gradients can reach
    it), anything else becomes a float64 numpy scalar."""
    return v if isinstance(v, torch.Tensor) else np.asarray(v, float)


def _op_scalar(v):
    """The op argument of a scalar parameter: the tensor itself when it
    requires grad, else a Python float."""
    if isinstance(v, torch.Tensor):
        return v if v.requires_grad else float(v.detach().cpu())
    return float(v)


class ApplyInterpolation(DetectorLayer):
    """Resample the PSF through a coordinate transformation
    (detector_layers.py:68-97); not fused into the detector chain."""

    def __init__(self, transformation, method: str = "linear", fill: float = 0.0):
        from ..coordinates import BaseCoordTransform
        if not isinstance(transformation, BaseCoordTransform):
            raise TypeError("transformation must be a BaseCoordTransform.")
        self.transformation = transformation
        self.method = str(method)
        self.fill = np.asarray(fill, dtype=float)

    def __call__(self, psf: PSF) -> PSF:
        return psf.interpolate(self.transformation, method=self.method, fill=float(self.fill))


class ApplyPixelResponse(DetectorLayer):
    """psf * pixel_response (detector_layers.py:100-131)."""

    def __init__(self, pixel_response):
        if isinstance(pixel_response, torch.Tensor):
            self.pixel_response = pixel_response.to(torch.float32)
        else:
            self.pixel_response = np.asarray(pixel_response, dtype=np.float32)
        if self.pixel_response.ndim != 2:
            raise ValueError("pixel_response must be a 2d array.")

    def _det_op(self):
        resp = self.pixel_response
        if isinstance(resp, torch.Tensor) and (resp.is_cuda or resp.requires_grad):
            # a tracked host tensor is uploaded on every call (its graph is per call)
            return ("pixel_response", resp if resp.is_cuda else runtime.as_float(resp))
        return ("pixel_response", self._device((resp,), lambda: runtime.as_float(resp)))

    def __call__(self, psf: PSF) -> PSF:
        return self._run(psf)


class ApplyJitter(DetectorLayer):
    """Convolution with a normalised Gaussian of standard deviation `sigma`
    pixels: gaussian on kernel_size * oversample points, sum-downsampled by
    oversample (detector_layers.py:134-199)."""

    def __init__(self, sigma, kernel_size: int = 9, oversample: int = 3):
        self.kernel_size = int(kernel_size)
        self.sigma = _scalar(sigma)
        self.oversample = int(oversample)
        if self.kernel_size <= 0:
            raise ValueError("kernel_size must be greater than 0.")

    @property
    def kernel(self) -> np.ndarray:
        """detector_layers.py:180-196 (a float32 host array of side kernel_size)."""
        s = float(self.sigma)
        k = gaussian(mean=np.zeros(2), std=np.array([s, s]),
                     npixels=self.kernel_size * self.oversample).astype(np.float64)
        n, f = self.kernel_size, self.oversample
        return k.reshape(n, f, n, f).sum((1, 3)).astype(np.float32)

    def _factor(self) -> np.ndarray:
        """The 1-D factor of `kernel`: the Gaussian and its normalisation are
        separable, and so is the block sum, so kernel = outer(f, f) with f the
        normalised block sums of the 1-D pdf on linspace(-5, 5, n * oversample)."""
        s = float(self.sigma)
        n, f = self.kernel_size, self.oversample
        g = np.exp(-0.5 * (np.linspace(-5.0, 5.0, n * f) / s) ** 2)
        g = g.reshape(n, f).sum(1)
        return (g / g.sum()).astype(np.float32)

    def _factor_tracked(self) -> torch.Tensor:
        """_factor as a torch expression of a sigma that requires grad (the
        gradient reaches sigma through dlx_detector_vjp_params' kernel
        cotangent)."""
        s = self.sigma.to(runtime.device(), torch.float64)
        n, f = self.kernel_size, self.oversample
        xs = torch.linspace(-5.0, 5.0, n * f, dtype=torch.float64, device=s.device)
        g = torch.exp(-0.5 * (xs / s) ** 2).reshape(n, f).sum(1)
        return (g / g.sum()).float()

    def _det_op(self):
        # two 1-D passes (2K multiply-adds per pixel) instead of K^2
        if isinstance(self.sigma, torch.Tensor) and self.sigma.requires_grad \
                and torch.is_grad_enabled():
            f = self._factor_tracked()
            return ("convolve_separable", (f, f))
        key = (float(self.sigma), self.kernel_size, self.oversample)
        # both factors in one device buffer, so the call passes it as is
        ff = self._device(key, lambda: runtime.as_float(np.tile(self._factor(), 2)))
        return ("convolve_separable", (ff[:self.kernel_size], ff[self.kernel_size:]))

    def __call__(self, psf: PSF) -> PSF:
        return self._run(psf)


class ApplySaturation(DetectorLayer):
    """minimum(psf, threshold) (detector_layers.py:202-229)."""

    def __init__(self, threshold):
        self.threshold = _scalar(threshold)

    def _det_op(self):
        return ("saturate", _op_scalar(self.threshold))

    def __call__(self, psf: PSF) -> PSF:
        return self._run(psf)


class AddConstant(DetectorLayer):
    """psf + value (detector_layers.py:232-259)."""

    def __init__(self, value):
        self.value = _scalar(value)

    def _det_op(self):
        return ("add", _op_scalar(self.value))

    def __call__(self, psf: PSF) -> PSF:
        return self._run(psf)


class Downsample(DetectorLayer):
    """Sum over kernel_size x kernel_size blocks; the pixel scale grows by
    kernel_size (detector_layers.py:262-293, psfs.py:74-91)."""

    def __init__(self, kernel_size: int):
        self.kernel_size = int(kernel_size)
        if self.kernel_size <= 0:
            raise ValueError("kernel_size must be greater than 0.")

    def _det_op(self):
        return ("downsample", self.kernel_size)

    def _scale_factor(self) -> int:
        return self.kernel_size

    def __call__(self, psf: PSF) -> PSF:
        return self._run(psf)

