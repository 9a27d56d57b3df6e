"""PSF container (psfs.py:14-274): a float32 device image plus its pixel scale.

The detector-space operations (downsample, convolve, min and the arithmetic
operators) run on the device; downsample and convolve are passes of the HIP
detector chain (dlx_detector_apply).
"""

from __future__ import annotations

import numpy as np
import torch

from .base import Base

__all__ = ["PSF"]


class PSF(Base):
    """Intensity image plus its pixel scale (psfs.py:14-72)."""

    def __init__(self, data, pixel_scale):
        self.data = data if isinstance(data, torch.Tensor) else torch.as_tensor(
            np.asarray(data, np.float32))
        self.pixel_scale = np.asarray(pixel_scale, float)

    @property
    def npixels(self) -> int:
        return self.data.shape[-1]

    @property
    def ndim(self) -> int:
        return self.pixel_scale.ndim

    def downsample(self, n: int) -> "PSF":
        """Sum over n x n blocks; pixel scale times n (psfs.py:74-91)."""
        from . import runtime
        return self.set(data=runtime.detector(self.data, [("downsample", int(n))]),
                        pixel_scale=self.pixel_scale * n)

    def convolve(self, other, method: str = "auto") -> "PSF":
        """convolve(data, other, mode="same"), direct summation (psfs.py:93-110).
        `method` is accepted for signature parity; "fft" is not implemented."""
        if method not in ("auto", "direct"):
            raise NotImplementedError(f"convolution method {method!r} is not implemented")
        from . import runtime
        return self.set(data=runtime.detector(self.data, [("convolve", other)]))

    def min(self, path: str, value) -> "PSF":
        """minimum(leaf, value), as zodiax's Base.min (ApplySaturation's call)."""
        if path != "data":
            return self.set(path, np.minimum(self.get(path), value))
        from . import runtime
        return self.set(data=runtime.detector(self.data, [("saturate", float(value))]))

    def rotate(self, angle: float, method: str = "linear") -> "PSF":
        """The PSF rotated by `angle` radians about its centre (psfs.py:112-128)."""
        from .utils.interpolation import rotate
        return self.set(data=rotate(self.data, angle, method))

    def interpolate(self, transformation, method: str = "linear", fill: float = 0.0) -> "PSF":
        """The PSF sampled at transformation(pixel_coords(n, n * pixel_scale))
        (psfs.py:130-157)."""
        from . import runtime
        from .coordinates import BaseCoordTransform
        if not isinstance(transformation, BaseCoordTransform):
            raise TypeError("transformation must be a BaseCoordTransform.")
        from .coordinates import CoordTransform, _chain_coords, _untracked
        _untracked(transformation, "PSF.interpolate")
        d = float(np.mean(self.pixel_scale))
        if isinstance(transformation, CoordTransform):
            return self.set(data=runtime.interp_transform(
                self.data, d, 0.0, self.npixels, d, 0.0, [transformation._struct()], method,
                fill))
        q = _chain_coords((transformation,), self.npixels, d, 0.0, runtime.device())
        return self.set(data=runtime.interp_points(self.data, d, 0.0, q[0], q[1], method, fill))

    def resize(self, npixels: int) -> "PSF":
        """psfs.py:159-173."""
        from .utils.array_ops import resize
        return self.set(data=resize(self.data, npixels))

    def flip(self, axis) -> "PSF":
        """psfs.py:175-190 (axis 0 = y, axis 1 = x)."""
        axis = (axis,) if np.ndim(axis) == 0 else tuple(axis)
        return self.set(data=torch.flip(self.data, [a % self.data.ndim for a in axis]))

    def _unified_op(self, other, op: str) -> "PSF":
        """psfs.py:192-234: None leaves the PSF unchanged; a PSF contributes its data."""
        if other is None:
            return self
        if isinstance(other, PSF):
            other = other.data
        if not isinstance(other, (torch.Tensor, np.ndarray, float, int, complex, np.generic)):
            raise TypeError(
                f"Unsupported type for {op}: {type(other)}. Must be an array, "
                "PSF, or None."
            )
        if isinstance(other, np.ndarray):
            other = torch.as_tensor(other.astype(np.float32), device=self.data.device)
        if op == "add":
            return self.add("data", other)
        if op == "subtract":
            return self.add("data", -other)
        if op == "multiply":
            return self.multiply("data", other)
        return self.divide("data", other)

    def __add__(self, other):
        return self._unified_op(other, "add")

    def __sub__(self, other):
        return self._unified_op(other, "subtract")

    def __mul__(self, other):
        return self._unified_op(other, "multiply")

    def __truediv__(self, other):
        return self._unified_op(other, "divide")

    __iadd__ = __add__
    __isub__ = __sub__
    __imul__ = __mul__
    __itruediv__ = __truediv__
