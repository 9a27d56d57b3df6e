"""Optic and BasisOptic (layers/optics.py:17-175)."""

from __future__ import annotations

from ..wavefronts import Wavefront
from .optical_layers import TransmissiveLayer, AberratedLayer, BasisLayer, _shape

__all__ = ["Optic", "BasisOptic"]


class Optic(TransmissiveLayer, AberratedLayer):
    """transmission * E, then OPD, then phase, optional normalise (optics.py:17-96)."""

    def __init__(self, transmission=None, opd=None, phase=None, normalise: bool = False):
        if transmission is not None:
            ts = _shape(transmission)
            if opd is not None and ts != _shape(opd):
                raise ValueError(
                    "transmission and opd must have the same shape. Got "
                    f"shapes {ts} and {_shape(opd)}."
                )
            if phase is not None and ts != _shape(phase):
                raise ValueError(
                    "transmission and phase must have the same shape. Got "
                    f"shapes {ts} and {_shape(phase)}."
                )
        AberratedLayer.__init__(self, opd, phase)
        TransmissiveLayer.__init__(self, transmission, normalise)

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        wavefront *= self.transmission
        wavefront = wavefront.add_opd(self.opd)
        wavefront = wavefront.add_phase(self.phase)
        if self.normalise:
            wavefront = wavefront.normalise()
        return wavefront

    def _lower(self, plan) -> bool:
        plan.mul_amp(self.transmission)
        plan.add_opd(self.opd)
        plan.add_phase(self.phase)
        if self.normalise:
            plan.normalise()
        return True


class BasisOptic(TransmissiveLayer, BasisLayer):
    """transmission, then the basis effect, optional normalise (optics.py:99-175)."""

    def __init__(self, basis, transmission=None, coefficients=None, normalise: bool = False,
                 effect: str = "opd", coefficient_shape=None):
        TransmissiveLayer.__init__(self, transmission, normalise)
        BasisLayer.__init__(self, basis, coefficients, effect, coefficient_shape)

    def __call__(self, wavefront: Wavefront) -> Wavefront:
        wavefront *= self.transmission
        wavefront = BasisLayer.__call__(self, wavefront)
        if self.normalise:
            wavefront = wavefront.normalise()
        return wavefront

    def _lower(self, plan) -> bool:
        plan.mul_amp(self.transmission)
        BasisLayer._lower(self, plan)
        if self.normalise:
            plan.normalise()
        return True
