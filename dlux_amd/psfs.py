"""PSF container (psfs.py:14-72)."""

from __future__ import annotations

import numpy as np
import torch

from .base import Base

__all__ = ["PSF"]


class PSF(Base):
    """Intensity image plus its pixel scale."""

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

    def __mul__(self, other):
        return self.multiply("data", other)

    def __add__(self, other):
        return self.add("data", other)
