"""Coordinate / sampling specifications (coordinates.py:23-157)."""

from __future__ import annotations

import numpy as np

from .base import Base

__all__ = ["Spec", "PadSpec", "CoordSpec"]


class Spec(Base):
    """Abstract sampling specification."""


class PadSpec(Spec):
    """Integer pad / crop factors relative to the input grid (coordinates.py:34-71)."""

    def __init__(self, pad=1, crop=1, c=0.0):
        self.pad = int(pad)
        self.crop = int(crop)
        self.c = np.asarray(c, float)


class CoordSpec(Spec):
    """Explicit (n, d, c) sampling (coordinates.py:74-157)."""

    def __init__(self, n=None, d=None, c=0.0):
        self.n = n
        self.d = None if d is None else np.asarray(d, float)
        self.c = None if c is None else np.asarray(c, float)

    @property
    def xs(self) -> np.ndarray:
        """c + (arange(n) - (n - 1)/2) * d  (host FP64)."""
        if self.d is None:
            raise ValueError("d must be specified to calculate coordinates.")
        return self.c + (np.arange(self.n) - (self.n - 1) / 2) * self.d

    @property
    def fov(self):
        if self.d is None:
            raise ValueError("d must be specified to calculate FOV.")
        return self.n * self.d

    @property
    def extent(self):
        if self.d is None:
            raise ValueError("d must be specified to calculate extent.")
        return self.c - (self.n / 2) * self.d, self.c + (self.n / 2) * self.d
