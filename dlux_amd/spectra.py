"""Spectra (spectra.py:20-117): wavelengths plus weights normalised to sum 1."""

from __future__ import annotations

import numpy as np

from .base import Base

__all__ = ["Spectrum"]


class Spectrum(Base):
    """Wavelengths (m) and relative weights, normalised to sum 1 (spectra.py:56-117)."""

    def __init__(self, wavelengths, weights=None):
        self.wavelengths = np.asarray(wavelengths, float)
        if weights is None:
            weights = np.ones(self.wavelengths.shape) / self.wavelengths.shape[-1]
        weights = np.asarray(weights, float)
        if weights.ndim == 2:
            self.weights = weights / weights.sum(-1)[:, None]
        else:
            self.weights = weights / weights.sum()
        if self.weights.ndim == 1:
            if self.wavelengths.shape != self.weights.shape:
                raise ValueError("wavelengths and weights must have the same shape.")
        elif self.wavelengths.shape != self.weights.shape[-1:]:
            raise ValueError("wavelengths and weights must have the same trailing shape.")

    def normalise(self) -> "Spectrum":
        s = self.weights.sum(-1)[:, None] if self.weights.ndim == 2 else self.weights.sum()
        return self.divide("weights", s)
