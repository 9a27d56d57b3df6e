"""FFT / MFT propagator layers (layers/propagators.py:14-217), the ABCD
elements and linear-canonical-transform propagators
(layers/abcd_propagators.py:23-383) and the angular-spectrum propagator
(layers/abcd_propagators.py:386-459)."""

from __future__ import annotations

import numpy as np
import torch

from .. import runtime
from ..base import Base
from ..coordinates import CoordSpec, PadSpec
from ..utils.array_ops import crop_to
from ..wavefronts import Wavefront
from .optical_layers import OpticalLayer

__all__ = ["Propagator", "FFT", "MFT", "ASMPropagator", "ABCDElement", "ABCDFreeSpace",
           "ABCDLens", "ABCDMirror", "ABCDConjugatePlane", "ABCDPropagator", "MFTPropagator",
           "FFTPropagator", "Fraunhofer", "Fresnel"]


class Propagator(OpticalLayer):
    """focal_length (None = angular) and direction (propagators.py:14-54)."""

    def __init__(self, focal_length=None, inverse: bool = False):
        if focal_length is not None:
            focal_length = float(np.asarray(focal_length, float))
        self.focal_length = focal_length
        self.inverse = bool(inverse)


class FFT(Propagator):
    """Padded FFT propagation, then crop (propagators.py:57-142)."""

    def __init__(self, focal_length=None, inverse: bool = False, pad: int = 1, crop: int = 1,
                 center: bool = True):
        super().__init__(focal_length, inverse)
        self.pad = int(pad)
        self.crop = int(crop)
        self.center = bool(center)

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        spec = CoordSpec(c=0.0) if self.center else None
        size_out = wavefront.npixels * self.pad // self.crop
        return wavefront.propagate_FFT(pad=self.pad, focal_length=self.focal_length,
                                       inverse=self.inverse, spec_out=spec).resize(size_out)


class MFT(Propagator):
    """MFT propagation to (npixels, pixel_scale) (propagators.py:145-217)."""

    def __init__(self, npixels: int, pixel_scale, focal_length=None, inverse: bool = False):
        super().__init__(focal_length, inverse)
        self.pixel_scale = np.asarray(pixel_scale, float)
        self.npixels = int(npixels)

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        return wavefront.propagate(npixels=self.npixels, pixel_scale=self.pixel_scale,
                                   focal_length=self.focal_length, inverse=self.inverse)


class ASMPropagator(OpticalLayer):
    """Angular-spectrum free-space propagation (abcd_propagators.py:386-459).

    The propagation itself (abcdLux.asm.asm_prop in the reference) is the HIP
    transfer-function method dlx_asm_prop: pad to n_padded, FFT, multiply by
    exp(2 pi i z sqrt(1/l^2 - f^2)) (0 for evanescent waves), inverse FFT.
    """

    def __init__(self, distance, spec):
        self.distance = float(np.asarray(distance, float))
        if isinstance(spec, CoordSpec):
            if spec.d is not None or spec.c is not None:
                raise ValueError("ASMPropagator CoordSpec can not specify d or c.")
        self.spec = spec

    def __getattr__(self, key):
        spec = self.__dict__.get("spec")
        if spec is not None and hasattr(spec, key):
            return getattr(spec, key)
        raise AttributeError(f"{type(self).__name__} has no attribute {key!r}")

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        if isinstance(self.spec, CoordSpec):
            n_padded = self.spec.n
        else:
            n_padded = wavefront.npixels * self.spec.pad
        wf = wavefront.propagate_ASM(self.distance, n_padded)
        if isinstance(self.spec, PadSpec) and self.spec.crop > 1:
            n_out = wf.npixels // self.spec.crop
            wf = wf.set(phasor=crop_to(wf.phasor, n_out))
        return wf


# ---------------------------------------------------------------------------
# ABCD elements and LCT propagators (layers/abcd_propagators.py:23-383).  The
# matrices are the textbook ray-transfer matrices; the transforms are the
# Collins integral, run on the device by dlx_lct (two complex GEMMs) and
# dlx_lct_fft (chirp, centred FFT, chirp).
# ---------------------------------------------------------------------------


class ABCDElement(Base):
    """An optical element with a 2 x 2 ray-transfer matrix (abcd_propagators.py:23-33)."""

    @property
    def abcd(self) -> np.ndarray:  # pragma: no cover - abstract
        raise NotImplementedError


class ABCDFreeSpace(ABCDElement):
    """[[1, z], [0, 1]] (abcd_propagators.py:36-66)."""

    def __init__(self, distance):
        self.distance = np.asarray(distance, float)

    @property
    def abcd(self) -> np.ndarray:
        return np.array([[1.0, float(self.distance)], [0.0, 1.0]])


class ABCDLens(ABCDElement):
    """Thin lens [[1, 0], [-1/f, 1]] (abcd_propagators.py:69-111)."""

    def __init__(self, focal_length):
        self.focal_length = np.asarray(focal_length, float)

    @property
    def abcd(self) -> np.ndarray:
        return np.array([[1.0, 0.0], [-1.0 / float(self.focal_length), 1.0]])


class ABCDMirror(ABCDElement):
    """Curved mirror of radius R, [[1, 0], [-2/R, 1]] (abcd_propagators.py:114-144)."""

    def __init__(self, radius):
        self.radius = np.asarray(radius, float)

    @property
    def abcd(self) -> np.ndarray:
        return np.array([[1.0, 0.0], [-2.0 / float(self.radius), 1.0]])


class ABCDConjugatePlane(ABCDElement):
    """Front to back focal plane of a lens, [[0, f], [-1/f, 0]]: the
    pupil-focal Fourier relation (abcd_propagators.py:147-186)."""

    def __init__(self, focal_length):
        self.focal_length = np.asarray(focal_length, float)

    @property
    def abcd(self) -> np.ndarray:
        f = float(self.focal_length)
        return np.array([[0.0, f], [-1.0 / f, 0.0]])


def compose_abcd(mats) -> np.ndarray:
    """Elements in propagation order: M = M_n ... M_2 M_1."""
    out = np.eye(2)
    for m in mats:
        out = np.asarray(m, float) @ out
    return out


class ABCDPropagator(OpticalLayer):
    """A propagator through a sequence of ABCD elements (abcd_propagators.py:192-242)."""

    def __init__(self, ABCDs, spec):
        from ..optical_systems import list2dictionary
        self.ABCDs = list2dictionary(ABCDs, True, allowed_types=(ABCDElement,))
        self.spec = spec

    def __getattr__(self, key):
        d = self.__dict__
        spec, abcds = d.get("spec"), d.get("ABCDs")
        if spec is not None and hasattr(spec, key):
            return getattr(spec, key)
        if abcds is not None:
            if key in abcds:
                return abcds[key]
            for el in abcds.values():
                if hasattr(el, key):
                    return getattr(el, key)
        raise AttributeError(f"{type(self).__name__} has no attribute {key!r} "
                             f"(elements: {list(abcds) if abcds is not None else []})")

    @property
    def abcd(self) -> np.ndarray:
        return compose_abcd([m.abcd for m in self.ABCDs.values()])


def _wf_grid(wavefront):
    return float(np.ravel(wavefront.pixel_scale)[0]), float(np.ravel(wavefront.center)[0])


class MFTPropagator(ABCDPropagator):
    """LCT to the grid of a CoordSpec (n, d, c) by the Collins integral as two
    complex GEMMs (dlx_lct).  Replaces abcdLux.lct.lct_prop
    (abcd_propagators.py:245-297)."""

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        spec = self.spec
        d_in, c_in = _wf_grid(wavefront)
        c_out = 0.0 if spec.c is None else float(np.ravel(spec.c)[0])
        field = runtime.lct(wavefront.phasor, wavefront._wl_list(), self.abcd, d_in, c_in,
                            int(spec.n), float(np.ravel(spec.d)[0]), c_out)
        return wavefront.set(phasor=field, pixel_scale=np.asarray(spec.d, float),
                             center=np.atleast_1d(np.asarray(c_out, float)))


class FFTPropagator(ABCDPropagator):
    """LCT on the FFT grid of n_padded points, d_out = lambda |B| / (n_padded
    d_in), centred on the spec's c, then the PadSpec crop
    (abcd_propagators.py:300-383).  dlx_lct_fft (chirp, centred FFT, chirp)
    for power-of-two grids; the same transform by dlx_lct's GEMMs otherwise."""

    def __init__(self, ABCDs, spec):
        if isinstance(spec, CoordSpec) and spec.d is not None:
            raise ValueError("FFTPropagator CoordSpec can not specify d.")
        super().__init__(ABCDs=ABCDs, spec=spec)

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        spec = self.spec
        n_pad = int(spec.n) if isinstance(spec, CoordSpec) else wavefront.npixels * spec.pad
        d_in, c_in = _wf_grid(wavefront)
        c_out = 0.0 if spec.c is None else float(np.ravel(spec.c)[0])
        abcd = self.abcd
        wls = wavefront._wl_list()
        if n_pad & (n_pad - 1) == 0 and n_pad <= 8192 and (n_pad - wavefront.npixels) % 2 == 0:
            field, ps = runtime.lct_fft(wavefront.phasor, wls, abcd, d_in, c_in, n_pad, c_out)
        else:
            ph = wavefront.phasor if wavefront.phasor.ndim == 3 else wavefront.phasor[None]
            ps = np.array([wl * abs(abcd[0, 1]) / (n_pad * d_in) for wl in wls])
            field = torch.stack([runtime.lct(ph[i], wl, abcd, d_in, c_in, n_pad, ps[i], c_out)
                                 for i, wl in enumerate(wls)])
            if wavefront.phasor.ndim == 2:
                field = field[0]
        if isinstance(spec, PadSpec) and spec.crop > 1:
            field = crop_to(field, field.shape[-1] // spec.crop)
        ps = np.asarray(ps, float)
        if wavefront.phasor.ndim == 2:
            ps = np.asarray(np.ravel(ps)[0], float)
        return wavefront.set(phasor=field, pixel_scale=ps,
                             center=np.atleast_1d(np.asarray(c_out, float)))


class Fraunhofer(ABCDPropagator):
    """Placeholder, as in the reference (abcd_propagators.py:462-475)."""

    def __init__(self):
        raise NotImplementedError("Fraunhofer propagator is not yet implemented.")


class Fresnel(ABCDPropagator):
    """Placeholder, as in the reference (abcd_propagators.py:478-490)."""

    def __init__(self):
        raise NotImplementedError("Fresnel propagator is not yet implemented.")
