"""Point sources (sources.py:269-411)."""

from __future__ import annotations

import numpy as np

from .base import Base
from .psfs import PSF
from .spectra import Spectrum

__all__ = ["PointSource", "PointSources"]


def _validate_return_mode(return_wf, return_psf):
    if return_wf and return_psf:
        raise ValueError(
            "Cannot return both Wavefront and PSF objects. Choose one: "
            "return_wf=True for Wavefront, or return_psf=True for PSF."
        )


class Source(Base):
    """A spectrum-carrying source (sources.py:27-268)."""

    def __init__(self, wavelengths=None, weights=None, spectrum=None):
        if spectrum is not None:
            self.spectrum = spectrum
        else:
            if wavelengths is None:
                raise ValueError("Either wavelengths or spectrum must be specified.")
            self.spectrum = Spectrum(wavelengths, weights)

    @property
    def wavelengths(self):
        return self.spectrum.wavelengths

    @property
    def weights(self):
        return self.spectrum.weights

    def normalise(self):
        return self.set("spectrum", self.spectrum.normalise())


class PointSource(Source):
    """One point source: position (rad), flux, spectrum (sources.py:269-327)."""

    def __init__(self, wavelengths=None, position=np.zeros(2), flux=1.0, weights=None,
                 spectrum=None):
        self.position = np.asarray(position, float)
        if self.position.shape != (2,):
            raise ValueError("position must be a 1d array of shape (2,).")
        self.flux = np.asarray(flux, float)
        super().__init__(wavelengths=wavelengths, weights=weights, spectrum=spectrum)

    def model(self, optics, return_wf: bool = False, return_psf: bool = False):
        _validate_return_mode(return_wf, return_psf)
        src = self.normalise()
        weights = src.weights * src.flux
        return optics.propagate(src.wavelengths, src.position, weights, return_wf, return_psf)


class PointSources(Source):
    """Sources sharing one spectrum, with per-source positions and fluxes (sources.py:330-411)."""

    def __init__(self, wavelengths=None, position=np.zeros((1, 2)), flux=None, weights=None,
                 spectrum=None):
        super().__init__(wavelengths=wavelengths, weights=weights, spectrum=spectrum)
        self.position = np.asarray(position, float)
        if self.position.ndim != 2:
            raise ValueError("position must be a 2d array.")
        if flux is None:
            self.flux = np.ones(len(self.position))
        else:
            self.flux = np.asarray(flux, float)
            if self.flux.ndim != 1:
                raise ValueError("flux must be a 1d array.")
            if len(self.flux) != len(self.position):
                raise ValueError("Length of flux must be equal to length of positions.")

    def model(self, optics, return_wf: bool = False, return_psf: bool = False):
        _validate_return_mode(return_wf, return_psf)
        src = self.normalise()
        weights = src.weights[None, :] * src.flux[:, None]
        if return_wf:
            return [optics.propagate(src.wavelengths, pos, w, return_wf=True)
                    for pos, w in zip(src.position, weights)]
        # every (source, wavelength) pair in one batched call where the system allows it
        batched = getattr(optics, "_sources_psf", None)
        out = batched(src.wavelengths, src.position, weights) if batched else None
        if out is not None:
            psf, p_out = out
            return PSF(psf, p_out) if return_psf else psf
        psf = None
        for pos, w in zip(src.position, weights):
            p = optics.propagate(src.wavelengths, pos, w)
            psf = p if psf is None else psf + p
        if return_psf:
            return PSF(psf, getattr(optics, "psf_pixel_scale", 1.0))
        return psf
