"""FFT / MFT propagator layers (layers/propagators.py:14-217) and the
angular-spectrum propagator (layers/abcd_propagators.py:386-459)."""

from __future__ import annotations

import numpy as np

from ..coordinates import CoordSpec, PadSpec
from ..utils.array_ops import crop_to
from ..wavefronts import Wavefront
from .optical_layers import OpticalLayer

__all__ = ["Propagator", "FFT", "MFT", "ASMPropagator"]


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
