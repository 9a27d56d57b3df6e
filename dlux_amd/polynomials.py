"""Zernike basis parameterisations (polynomials.py:20-117).

Every basis is evaluated on the device: on arbitrary coordinates with
dlx_zernike_basis (and, for the n-sided polike, the wedge radius computed on
the device first), and on a wavefront's own pixel grid with
dlx_aperture_basis, which generates the coordinates in the kernel.
"""

from __future__ import annotations

import numpy as np

from . import runtime
from .base import Base
from .parametric import CoordBasis, ExplicitBasis
from .utils import zernikes as zk

__all__ = ["DynamicZernike", "ZernikeBasis", "DynamicZernikeBasis"]


class DynamicZernike(Base):
    """One Noll Zernike, evaluated on demand (polynomials.py:20-52)."""

    def __init__(self, j: int):
        self.j = int(j)
        if self.j < 1:
            raise ValueError("The Zernike index must be greater than 0.")
        self.name = zk.zernike_name(self.j)
        self.n, self.m = zk.noll_indices(self.j)
        self._c, self._k = zk.zernike_factors(self.j)

    def calculate(self, coordinates, nsides: int = 0, diameter: float = 2.0):
        if nsides == 0:
            return zk.zernike(self.j, coordinates, diameter)
        return zk.polike(nsides, self.j, coordinates, diameter)


class _ZernikeBasis:
    @staticmethod
    def get_indices(js=None, radial_orders=None) -> list:
        """polynomials.py:55-69."""
        if (js is None) == (radial_orders is None):
            raise ValueError("Provide exactly one of js or radial_orders.")
        if js is not None:
            indices = [int(j) for j in np.atleast_1d(np.asarray(js))]
        else:
            indices = zk.radial_orders_to_indices(radial_orders)
        if not indices:
            raise ValueError("At least one Zernike mode must be selected.")
        if any(j < 1 for j in indices):
            raise ValueError("Zernike indices must be greater than zero.")
        return indices


class ZernikeBasis(_ZernikeBasis, ExplicitBasis):
    """A Zernike basis sampled once on given coordinates (polynomials.py:72-82)."""

    def __init__(self, coordinates, js=None, radial_orders=None, coefficients=None,
                 diameter=2.0):
        js = self.get_indices(js, radial_orders)
        basis = zk.zernike_basis(js, coordinates, diameter)
        super().__init__(basis, coefficients, (len(js),))


class DynamicZernikeBasis(_ZernikeBasis, CoordBasis):
    """A Zernike (or n-sided polike) basis evaluated from the coordinate
    context at every call (polynomials.py:85-115).  From a wavefront the
    diameter defaults to the wavefront's; from explicit coordinates, to 2."""

    def __init__(self, js=None, radial_orders=None, coefficients=None, nsides=0):
        js = self.get_indices(js, radial_orders)
        self.zernikes = [DynamicZernike(j) for j in js]
        coefficients = np.zeros(len(js)) if coefficients is None else coefficients
        self._set_coefficients(coefficients, (len(js),))
        self.nsides = int(nsides)
        if self.nsides != 0 and self.nsides < 3:
            raise ValueError("nsides must be zero or greater than two.")

    @property
    def js(self) -> list:
        return [z.j for z in self.zernikes]

    def calculate_basis(self, *, wavefront=None, coordinates=None, diameter=None, **kwargs):
        from_wavefront = coordinates is None and wavefront is not None
        if diameter is None:
            diameter = wavefront.diameter if from_wavefront else 2.0
        diameter = float(np.asarray(diameter).reshape(-1)[0])
        if from_wavefront:
            # the wavefront's own pixel grid: coordinates generated in the kernel
            dx = float(np.mean(wavefront.pixel_scale))
            centre = float(np.asarray(wavefront.center).reshape(-1)[0])
            return runtime.aperture_basis(wavefront.npixels, dx, self.js, (), diameter / 2.0,
                                          self.nsides, centre)
        coordinates = self.get_coordinates(wavefront=wavefront, coordinates=coordinates)
        if self.nsides == 0:
            return zk.zernike_basis(self.js, coordinates, diameter)
        return zk.polike_basis(self.nsides, self.js, coordinates, diameter)
