"""Instruments (instruments.py:19-256): optics + source(s) + optional detector.

``Telescope.model`` is the reference's full forward model: the sources
through the optics as point-source images (one uneven batched MFT launch for
a whole scene where the optics lower), then the detector chain.  ``Dither``
models every dither position in the same single batched launch -- one output
image per dither -- and runs the detector chain over the stack of images.
"""

from __future__ import annotations

import numpy as np
import torch

from .base import Base
from .detectors import BaseDetector
from .optical_systems import OpticalSystem
from .psfs import PSF
from .sources import BaseSource, Scene

__all__ = ["Instrument", "Telescope", "Dither"]


class Instrument(Base):
    """Abstract instrument (instruments.py:19-34)."""

    def model(self, return_psf: bool = False):  # pragma: no cover - abstract
        raise NotImplementedError


class Telescope(Instrument):
    """instruments.py:37-171."""

    def __init__(self, optics, source, detector=None):
        if not isinstance(optics, OpticalSystem):
            raise TypeError(
                f"optics must be an OpticalSystem instance, "
                f"got {type(optics).__name__}. "
                "Use a dLux optics type from dLux.optical_systems."
            )
        self.optics = optics
        if isinstance(source, BaseSource):
            self.source = source
        elif isinstance(source, tuple):
            if len(source) != 2 or not isinstance(source[1], BaseSource):
                raise TypeError(
                    f"source tuple must be of the form (key: str, Source: Source), "
                    f"got {type(source).__name__} with length {len(source)}."
                )
            self.source = source[1]
        else:
            self.source = Scene(source)
        if not isinstance(detector, BaseDetector) and detector is not None:
            raise TypeError(
                f"detector must be a Detector instance, got {type(detector).__name__}. "
                "Use a dLux detector type from dLux.detectors."
            )
        self.detector = detector

    def __getattr__(self, key):
        """Attributes of the optics, source and detector, in that order (instruments.py:116-138)."""
        d = self.__dict__
        for name in ("optics", "source", "detector"):
            obj = d.get(name)
            if obj is None:
                continue
            try:
                return getattr(obj, key)
            except AttributeError:
                continue
        raise AttributeError(f"{type(self).__name__} has no attribute {key!r}")

    def _detect(self, psf: PSF, return_psf: bool):
        if self.detector is not None:
            return self.detector.model(psf, return_psf=return_psf)
        return psf if return_psf else psf.data

    def _fused_detector_model(self):
        """(image, pixel scale) of a point source through optics that lower to the
        fused PSF and a LayeredDetector whose layers all have a chain form, from
        one dlx_psf_poly_detector call (the wavelength sum fused into the
        detector's first pass); None otherwise."""
        from .detectors import LayeredDetector
        from .sources import PointSource
        if type(self.source) is not PointSource or not isinstance(self.detector,
                                                                   LayeredDetector):
            return None
        ops, scale = self.detector._chain()
        fused = getattr(self.optics, "_fused_psf", None)
        if ops is None or fused is None:
            return None
        src = self.source.normalise()
        weights = np.asarray(src.weights * src.flux, float)
        wavelengths = np.atleast_1d(np.asarray(src.wavelengths, float))
        out = fused(wavelengths, np.atleast_1d(weights), np.asarray(src.position, float),
                    detector_ops=ops)
        if out is None:
            return None
        img, p_out = out
        return img, p_out * scale

    def model(self, return_psf: bool = False):
        """instruments.py:140-171."""
        fused = self._fused_detector_model()
        if fused is not None:
            img, ps = fused
            return PSF(img, ps) if return_psf else img
        psfs = self.optics.model(self.source, return_psf=True)
        data = psfs.data
        if psfs.ndim > 0:
            data = data.sum(tuple(range(psfs.ndim)))
        psf = PSF(data, np.mean(psfs.pixel_scale))
        return self._detect(psf, return_psf)


class Dither(Telescope):
    """A Telescope modelled at several source offsets (instruments.py:174-256)."""

    def __init__(self, dithers, optics, source, detector=None):
        self.dithers = np.asarray(dithers, float)
        if self.dithers.ndim != 2 or self.dithers.shape[1] != 2:
            raise ValueError("dithers must be an array of shape (ndithers, 2)")
        super().__init__(optics=optics, source=source, detector=detector)

    def _batched_images(self):
        """(ndithers, n, n) images and p_out from one batched launch, or None."""
        batched = getattr(self.optics, "_sources_psf", None)
        if batched is None or not hasattr(self.source, "position"):
            return None
        try:
            terms = self.source._point_terms()
        except NotImplementedError:
            return None
        groups = [(wl, pos + d, w, k) for k, d in enumerate(self.dithers)
                  for wl, pos, w in terms]
        out = batched(groups, n_outputs=len(self.dithers))
        if out is None:
            return None
        images, p_out = out
        return torch.stack(images), p_out

    def model(self, return_psf: bool = False):
        """instruments.py:231-256: every dither moves the source position."""
        if not hasattr(self.source, "position"):
            raise AttributeError(
                f"{type(self.source).__name__} has no attribute 'position' to dither")
        got = self._batched_images()
        if got is None:
            outs = [Telescope.model(self.add("source.position", d), return_psf=True)
                    for d in self.dithers]
            psf = PSF(torch.stack([o.data for o in outs]),
                      np.stack([np.asarray(o.pixel_scale) for o in outs]))
            return psf if return_psf else psf.data
        images, p_out = got
        psf = PSF(images, p_out)
        if self.detector is not None:
            # one chain over the stack: the dithers are the batch of the launch
            psf = self.detector.model(psf, return_psf=True)
        psf = PSF(psf.data, np.full(len(self.dithers), float(psf.pixel_scale)))
        return psf if return_psf else psf.data
