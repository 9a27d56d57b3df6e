"""Spectra (spectra.py:12-178): wavelengths plus weights normalised to sum 1.

Spectra are host data (a few dozen wavelengths and weights); the optical
systems hand them to the device as per-wavelength kernel arguments.
"""

from __future__ import annotations

import numpy as np

from .base import Base

__all__ = ["BaseSpectrum", "SimpleSpectrum", "Spectrum", "PolySpectrum"]


class BaseSpectrum(Base):
    """Abstract spectrum (spectra.py:12-25)."""

    def normalise(self):  # pragma: no cover - abstract
        raise NotImplementedError


class SimpleSpectrum(BaseSpectrum):
    """A spectrum defined on a 1-D array of wavelengths (spectra.py:28-53)."""

    def __init__(self, wavelengths):
        self.wavelengths = np.asarray(wavelengths, float)
        if self.wavelengths.ndim != 1:
            raise ValueError("wavelengths must be a 1d array.")


class Spectrum(SimpleSpectrum):
    """Wavelengths (m) and relative weights, normalised to sum 1 (spectra.py:56-117)."""

    def __init__(self, wavelengths, weights=None):
        super().__init__(wavelengths)
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


class PolySpectrum(SimpleSpectrum):
    """Weights from the polynomial c0 + c1 x + ... + cn x^n at each
    wavelength, normalised to sum 1 (spectra.py:120-178)."""

    def __init__(self, wavelengths, coefficients):
        super().__init__(wavelengths)
        self.coefficients = np.asarray(coefficients, float)
        if self.coefficients.ndim != 1:
            raise ValueError("Coefficients must be a 1d array.")

    def _eval_weight(self, wavelength):
        return sum(c * wavelength ** i for i, c in enumerate(self.coefficients))

    @property
    def weights(self) -> np.ndarray:
        w = np.array([self._eval_weight(x) for x in self.wavelengths], float)
        return w / w.sum()

    def normalise(self) -> "PolySpectrum":
        return self
