"""Optical layers (layers/optical_layers.py:80-371): the drop-in surface.

Each layer is ``__call__(wavefront) -> wavefront``.  Layers whose action is a
lambda-independent amplitude / OPD / phase / normalisation also expose
``_lower(plan)``, which folds them into a fused-propagation plan (see
``optical_systems``): the optical system then runs the whole chain as one
dlx_psf_poly call instead of layer by layer.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import runtime
from ..base import Base
from ..wavefronts import Wavefront
from ..utils.math import eval_basis as _eval_basis

__all__ = ["OpticalLayer", "Interpolate", "TransmissiveLayer", "AberratedLayer", "BasisLayer",
           "Tilt", "Normalise"]


def _as_float(value):
    """A float32 device tensor (layer data lives on the GPU, like dLux's arrays)."""
    if value is None:
        return None
    if isinstance(value, torch.Tensor) and value.is_complex():
        raise TypeError("expected a real array")
    return runtime.as_float(value)


def _shape(value):
    return tuple(value.shape) if hasattr(value, "shape") else np.shape(value)


class OpticalLayer(Base):
    """Base optical layer: ``__call__(wavefront) -> wavefront``."""

    def __call__(self, wavefront: Wavefront) -> Wavefront:  # pragma: no cover
        raise NotImplementedError

    # Layers that cannot be folded into a fused plan return False.
    def _lower(self, plan) -> bool:
        return False


class Interpolate(OpticalLayer):
    """Resample the wavefront through a coordinate transformation
    (optical_layers.py:108-141)."""

    def __init__(self, transformation, method: str = "linear", complex: bool = True,
                 fill: float = 0.0):
        from ..coordinates import BaseCoordTransform
        if not isinstance(transformation, BaseCoordTransform):
            raise TypeError("transformation must be a BaseCoordTransform.")
        self.transformation = transformation
        self.method = str(method)
        self.complex = bool(complex)
        self.fill = np.asarray(fill, dtype=float)

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        return wavefront.interpolate(self.transformation, method=self.method,
                                     complex=self.complex, fill=float(self.fill))


class TransmissiveLayer(OpticalLayer):
    """E *= transmission, optional power normalisation (optical_layers.py:144-186)."""

    def __init__(self, transmission=None, normalise: bool = False):
        self.transmission = _as_float(transmission)
        self.normalise = bool(normalise)

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        wavefront *= self.transmission
        if self.normalise:
            wavefront = wavefront.normalise()
        return wavefront

    def _lower(self, plan) -> bool:
        plan.mul_amp(self.transmission)
        if self.normalise:
            plan.normalise()
        return True


class AberratedLayer(OpticalLayer):
    """Static OPD and/or phase (optical_layers.py:189-236)."""

    def __init__(self, opd=None, phase=None):
        if opd is not None and phase is not None and _shape(opd) != _shape(phase):
            raise ValueError(
                "opd and phase must have the same shape. Got "
                f"shapes {_shape(opd)} and {_shape(phase)}."
            )
        self.opd = _as_float(opd)
        self.phase = _as_float(phase)

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        wavefront = wavefront.add_opd(self.opd)
        return wavefront.add_phase(self.phase)

    def _lower(self, plan) -> bool:
        plan.add_opd(self.opd)
        plan.add_phase(self.phase)
        return True


class BasisLayer(OpticalLayer):
    """Basis . coefficients applied as OPD, phase or amplitude (optical_layers.py:239-327)."""

    def __init__(self, basis=None, coefficients=None, effect: str = "opd",
                 coefficient_shape=None):
        if effect not in ("opd", "phase", "amplitude"):
            raise ValueError("effect must be 'opd', 'phase', or 'amplitude'.")
        if basis is not None:
            bshape = _shape(basis)
            if coefficients is None:
                if coefficient_shape is None:
                    coefficient_shape = tuple(bshape[:-2])
                coefficients = np.zeros(coefficient_shape, np.float32)
            else:
                coefficient_shape = _shape(coefficients)
            if tuple(bshape[: len(coefficient_shape)]) != tuple(coefficient_shape):
                raise ValueError("The coefficient shape must match the leading basis dimensions.")
            basis = _as_float(basis)
            coefficients = _as_float(coefficients)
        self.basis = basis
        self.coefficients = coefficients
        self.effect = effect

    def eval_basis(self) -> torch.Tensor:
        """tensordot(basis, coefficients) on the device (dlx_eval_basis)."""
        return _eval_basis(self.basis, self.coefficients)

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        out = self.eval_basis()
        if self.effect == "phase":
            return wavefront.add_phase(out)
        if self.effect == "opd":
            return wavefront.add_opd(out)
        return wavefront * (1 + out)

    def _lower(self, plan) -> bool:
        out = self.eval_basis()
        if self.effect == "phase":
            plan.add_phase(out)
        elif self.effect == "opd":
            plan.add_opd(out)
        else:
            plan.mul_amp(1 + out)
        return True


class Tilt(OpticalLayer):
    """Tilt by (x, y) angles in radians (optical_layers.py:330-359)."""

    def __init__(self, angles):
        self.angles = np.asarray(angles, dtype=float)
        if self.angles.shape != (2,):
            raise ValueError("angles must be a 1d array of shape (2,).")

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        return wavefront.tilt(self.angles)

    def _lower(self, plan) -> bool:
        plan.add_tilt(self.angles)
        return True


class Normalise(OpticalLayer):
    """Normalise to unit power (optical_layers.py:362-371)."""

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        return wavefront.normalise()

    def _lower(self, plan) -> bool:
        plan.normalise()
        return True
