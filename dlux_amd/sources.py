"""Sources (sources.py:27-850): point sources, binaries, resolved sources and scenes.

Resolved sources are a fused polychromatic PSF followed by the convolve pass
of the HIP detector kernels.  Every other source reduces to point terms: for each star, its on-sky
position and its spectral weights already multiplied by its flux.  A
``Scene`` concatenates the terms of its members, so a whole scene -- every
(star, wavelength) instance, whatever wavelength grid each member uses --
becomes one uneven batched MFT launch (``dlx_mft_batched_psf``) when the
optical system can lower it; otherwise each star goes through
``optics.propagate``, which is the fused polychromatic path of a single star.
"""

from __future__ import annotations

import numpy as np
import torch

from .base import Base
from .psfs import PSF
from .spectra import BaseSpectrum, Spectrum
from .utils.source import fluxes_from_contrast, positions_from_sep

__all__ = ["BaseSource", "Source", "PointSource", "PointSources", "BinarySource",
           "ResolvedSource", "PointResolvedSource", "Scene"]


def _validate_return_mode(return_wf, return_psf):
    """sources.py:29-46."""
    if return_wf and return_psf:
        raise ValueError(
            "Cannot return both Wavefront and PSF objects. Choose one: "
            "set return_wf=True for Wavefront, or return_psf=True for PSF object."
        )


def _as_wavelengths_1d(wavelengths):
    """sources.py:49-68."""
    if wavelengths is None:
        return None
    wavelengths = np.asarray(wavelengths, dtype=float)
    if wavelengths.ndim != 1:
        raise ValueError("wavelengths must be a 1d array.")
    return wavelengths


def _infer_n_wavelengths(wavelengths, spectrum) -> int:
    """sources.py:71-103."""
    wavelengths = _as_wavelengths_1d(wavelengths)
    if wavelengths is not None:
        return len(wavelengths)
    if spectrum is None:
        raise ValueError(
            "wavelengths must be provided when weights is None and no spectrum "
            "is supplied."
        )
    if not isinstance(spectrum, BaseSpectrum):
        raise TypeError("spectrum must be a dLux BaseSpectrum instance from dLux.spectra.")
    return len(np.asarray(spectrum.wavelengths, dtype=float))


def _as_position_2d(position) -> np.ndarray:
    """sources.py:106-129."""
    position = np.asarray(position, dtype=float)
    if position.shape != (2,):
        raise ValueError(
            f"position must be a 1d array of shape (2,), got shape {position.shape}. "
            "Pass position=[x, y] where x, y are on-sky angles in radians."
        )
    return position


def _stack_wavefronts(wfs):
    """One Wavefront whose leaves carry a leading star axis (the layout of the
    reference's vmapped propagate, sources.py:401-405)."""
    out = wfs[0].set(phasor=torch.stack([w.phasor for w in wfs]))
    return out.set(wavelength=np.stack([np.asarray(w.wavelength) for w in wfs]),
                   pixel_scale=np.stack([np.asarray(w.pixel_scale) for w in wfs]),
                   center=np.stack([np.asarray(w.center) for w in wfs]))


def _terms_psf(optics, terms):
    """Sum over point terms of sum_l w[s, l] |E_{s,l}|^2, plus the output pixel
    scale.  One batched launch where the optics allow, else a loop over stars."""
    batched = getattr(optics, "_sources_psf", None)
    if batched is not None:
        out = batched([t + (0,) for t in terms], n_outputs=1)
        if out is not None:
            psfs, p_out = out
            return psfs[0], p_out
    psf, scales = None, []
    for wavelengths, positions, weights in terms:
        for pos, w in zip(positions, weights):
            p = optics.propagate(wavelengths, pos, w, return_psf=True)
            psf = p.data if psf is None else psf + p.data
            scales.append(np.mean(p.pixel_scale))
    return psf, float(np.mean(scales))


def _model_terms(optics, terms, return_wf, return_psf):
    if return_wf:
        wfs = [optics.propagate(wl, pos, w, return_wf=True)
               for wl, positions, weights in terms for pos, w in zip(positions, weights)]
        return _stack_wavefronts(wfs)
    psf, p_out = _terms_psf(optics, terms)
    return PSF(psf, p_out) if return_psf else psf


class BaseSource(Base):
    """Abstract source (sources.py:132-178)."""

    def normalise(self):  # pragma: no cover - abstract
        raise NotImplementedError

    def model(self, optics, return_wf: bool = False, return_psf: bool = False):
        raise NotImplementedError

    def _point_terms(self):
        """[(wavelengths (L,), positions (S, 2), weights (S, L))] of the
        normalised source, fluxes folded into the weights."""
        raise NotImplementedError(f"{type(self).__name__} has no point-source form")


class Source(BaseSource):
    """A spectrum-carrying source (sources.py:181-266)."""

    def __init__(self, wavelengths=None, weights=None, spectrum=None):
        if spectrum is not None:
            if not isinstance(spectrum, BaseSpectrum):
                raise TypeError(
                    f"spectrum must be a dLux BaseSpectrum instance from "
                    f"dLux.spectra, "
                    f"got {type(spectrum).__name__}. "
                    "Use a dLux spectrum type from dLux.spectra."
                )
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

    def _point_terms(self):
        src = self.normalise()
        return [(src.wavelengths, src.position[None], (src.weights * src.flux)[None])]

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

    def _point_terms(self):
        src = self.normalise()
        return [(src.wavelengths, src.position, src.weights[None, :] * src.flux[:, None])]

    def model(self, optics, return_wf: bool = False, return_psf: bool = False):
        _validate_return_mode(return_wf, return_psf)
        return _model_terms(optics, self._point_terms(), return_wf, return_psf)


class BinarySource(Source):
    """Two stars from a mean position, mean flux, separation, position angle and
    contrast; each star has its own spectral weights (sources.py:524-635)."""

    def __init__(self, wavelengths=None, position=np.zeros(2), mean_flux=1.0,
                 separation=0.0, position_angle=np.pi / 2, contrast=1.0, spectrum=None,
                 weights=None):
        wavelengths = _as_wavelengths_1d(wavelengths)
        if weights is None:
            weights = np.ones((2, _infer_n_wavelengths(wavelengths, spectrum)))
        self.position = _as_position_2d(position)
        self.mean_flux = np.asarray(mean_flux, float)
        self.separation = np.asarray(separation, float)
        self.position_angle = np.asarray(position_angle, float)
        self.contrast = np.asarray(contrast, float)
        super().__init__(wavelengths=wavelengths, spectrum=spectrum, weights=weights)

    def _point_terms(self):
        src = self.normalise()
        positions = positions_from_sep(src.position, src.separation, src.position_angle)
        flux = fluxes_from_contrast(src.mean_flux, src.contrast)
        weights = np.broadcast_to(src.weights, (2, len(src.wavelengths))) * flux[:, None]
        return [(src.wavelengths, positions, weights)]

    def model(self, optics, return_wf: bool = False, return_psf: bool = False):
        _validate_return_mode(return_wf, return_psf)
        return _model_terms(optics, self._point_terms(), return_wf, return_psf)


_NO_WF_MESSAGE = (
    "Wavefront information cannot be preserved through convolution. "
    "Convolution can only operate on PSFs (incoherent light). "
    "Please use return_wf=False to get the PSF array or return_psf=True "
    "to get a PSF object."
)


def _convolve_distribution(psf, distribution):
    """convolve(psf, distribution, mode="same") on the device (the convolve
    pass of the detector kernels; square distributions up to 64 px)."""
    from . import runtime
    d = np.asarray(distribution, np.float32)
    if d.ndim != 2 or d.shape[0] != d.shape[1]:
        raise NotImplementedError(
            f"only square distributions are supported, got shape {d.shape}")
    return runtime.detector(psf, [("convolve", d)])


class ResolvedSource(PointSource):
    """A point source's PSF convolved with a normalised intensity distribution
    (sources.py:414-521)."""

    def __init__(self, wavelengths=None, position=np.zeros(2), flux=1.0,
                 distribution=np.ones((3, 3)), weights=None, spectrum=None):
        distribution = np.asarray(distribution, dtype=float)
        self.distribution = distribution / distribution.sum()
        if self.distribution.ndim != 2:
            raise ValueError("distribution must be a 2d array.")
        super().__init__(position=position, flux=flux, spectrum=spectrum,
                         wavelengths=wavelengths, weights=weights)

    def normalise(self):
        """Spectrum normalised; distribution floored at 0 and normalised (sources.py:475-488)."""
        floor = np.maximum(self.distribution, 0.0)
        return self.set(["spectrum", "distribution"],
                        [self.spectrum.normalise(), floor / floor.sum()])

    def _point_terms(self):
        raise NotImplementedError("a resolved source has no point-source form")

    def model(self, optics, return_wf: bool = False, return_psf: bool = False):
        _validate_return_mode(return_wf, return_psf)
        if return_wf:
            raise NotImplementedError(_NO_WF_MESSAGE)
        src = self.normalise()
        psf = optics.propagate(src.wavelengths, src.position, src.weights * src.flux,
                               return_psf=True)
        conv = _convolve_distribution(psf.data, src.distribution)
        return PSF(conv, float(np.mean(psf.pixel_scale))) if return_psf else conv


class PointResolvedSource(ResolvedSource):
    """A point source plus a resolved component about it sharing its spectrum:
    flux split by contrast, each with its own spectral weights (sources.py:638-748)."""

    def __init__(self, wavelengths=None, position=np.zeros(2), flux=1.0,
                 distribution=np.ones((3, 3)), contrast=1.0, weights=None, spectrum=None):
        wavelengths = _as_wavelengths_1d(wavelengths)
        if weights is None:
            weights = np.ones((2, _infer_n_wavelengths(wavelengths, spectrum)))
        self.contrast = np.asarray(contrast, float)
        super().__init__(wavelengths=wavelengths, position=position, flux=flux,
                         distribution=distribution, spectrum=spectrum, weights=weights)

    def model(self, optics, return_wf: bool = False, return_psf: bool = False):
        _validate_return_mode(return_wf, return_psf)
        if return_wf:
            raise NotImplementedError(_NO_WF_MESSAGE)
        src = self.normalise()
        flux = fluxes_from_contrast(src.flux, src.contrast)
        weights = np.broadcast_to(src.weights, (2, len(src.wavelengths))) * flux[:, None]
        # the reference propagates with the default weights 1/L (sources.py:728) and then
        # weights each wavelength's PSF, so every term carries that extra 1/L
        weights = weights / len(src.wavelengths)
        point = optics.propagate(src.wavelengths, src.position, weights[0], return_psf=True)
        resolved = optics.propagate(src.wavelengths, src.position, weights[1])
        psf = point.data + _convolve_distribution(resolved, src.distribution)
        return PSF(psf, float(np.mean(point.pixel_scale))) if return_psf else psf


class Scene(BaseSource):
    """Several sources modelled together (sources.py:751-850).  Members are
    keyed by class name (de-duplicated with _0, _1, ...) or by the key of a
    (key, source) tuple; their point terms go into one batched launch."""

    def __init__(self, sources):
        from .optical_systems import list2dictionary
        if isinstance(sources, BaseSource):
            sources = [sources]
        elif isinstance(sources, tuple):
            sources = list(sources)
        self.sources = list2dictionary(sources, False, BaseSource)

    def __getattr__(self, key):
        sources = self.__dict__.get("sources")
        if sources is not None and key in sources:
            return sources[key]
        raise AttributeError(
            f"{type(self).__name__} has no attribute {key!r} "
            f"(sources: {list(sources) if sources is not None else []})")

    def normalise(self):
        return self.set("sources", {k: s.normalise() for k, s in self.sources.items()})

    def _point_terms(self):
        return [t for s in self.sources.values() for t in s._point_terms()]

    def model(self, optics, return_wf: bool = False, return_psf: bool = False):
        _validate_return_mode(return_wf, return_psf)
        if return_wf:
            return {k: s.model(optics, return_wf=True) for k, s in self.sources.items()}
        try:
            terms = self._point_terms()
        except NotImplementedError:
            terms = None
        if terms is not None:
            return _model_terms(optics, terms, False, return_psf)
        outs = [s.model(optics, return_psf=True) for s in self.sources.values()]
        psf = outs[0].data
        for o in outs[1:]:
            psf = psf + o.data
        if return_psf:
            return PSF(psf, float(np.mean([np.mean(o.pixel_scale) for o in outs])))
        return psf
