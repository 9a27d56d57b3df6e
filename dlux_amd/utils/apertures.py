"""Static aperture builders (utils/apertures.py:153-242) on the device."""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib, runtime
from .coordinates import pixel_coords
from .zernikes import zernike_basis

__all__ = ["circular_aperture"]


def circular_aperture(npixels: int, diameter: float, oversample: int = 5,
                      secondary_diameter: float | None = None,
                      array_diameter: float | None = None, spider_width: float | None = None,
                      spider_angles=None, zernike_nolls=None, zernike_oversize: float = 0.01,
                      return_support: bool = False):
    """Oversampled circular aperture (+ Zernike basis masked by the primary support).

    Kernels: dlx_circular_aperture (transmission, support) and dlx_zernike_basis.
    """
    if (spider_width is None) != (spider_angles is None):
        raise ValueError(
            "`spider_width` and `spider_angles` must both be provided or both be None."
        )
    dev = runtime.device()
    trans = torch.empty((npixels, npixels), dtype=torch.float32, device=dev)
    support = torch.empty((npixels, npixels), dtype=torch.float32, device=dev)
    angles = np.atleast_1d(np.asarray(spider_angles if spider_angles is not None else [],
                                      np.float64))
    st = _lib.load().dlx_circular_aperture(
        int(npixels), float(diameter), int(oversample),
        float(secondary_diameter) if secondary_diameter else 0.0,
        float(array_diameter) if array_diameter else 0.0,
        float(spider_width) if spider_width is not None else 0.0, int(angles.size),
        _lib.dptr(angles) if angles.size else None, trans.data_ptr(), support.data_ptr(),
        runtime.stream_handle())
    _lib.check(st, "dlx_circular_aperture")
    if zernike_nolls is None:
        return trans
    coord_diam = diameter if array_diameter is None else array_diameter
    coords = pixel_coords(npixels, diameter=coord_diam)
    basis = zernike_basis(zernike_nolls, coords, diameter * (1.0 + zernike_oversize))
    basis = basis * support[None]
    if return_support:
        return trans, basis, support > 0
    return trans, basis
