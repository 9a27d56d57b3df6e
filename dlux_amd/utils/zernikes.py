"""Zernike polynomials (utils/zernikes.py:99-315); bases are evaluated on the device."""

from __future__ import annotations

import ctypes
from math import lgamma, exp

import numpy as np
import torch

from .. import _lib, runtime

__all__ = ["noll_indices", "zernike_factors", "zernike", "zernike_basis", "zernike_name"]

_NAMES = {
    1: "Piston", 2: "Tilt X", 3: "Tilt Y", 4: "Defocus", 5: "Astig 45", 6: "Astig 0",
    7: "Coma Y", 8: "Coma X", 9: "Trefoil Y", 10: "Trefoil X", 11: "Spherical",
}


def zernike_name(j: int) -> str:
    return _NAMES.get(int(j), f"Zernike {int(j)}")


def noll_indices(j: int) -> tuple[int, int]:
    """(n, m) of Noll index j (the C ABI's dlx_noll_indices).  utils/zernikes.py:99-119."""
    n, m = ctypes.c_int32(), ctypes.c_int32()
    _lib.check(_lib.load().dlx_noll_indices(int(j), ctypes.byref(n), ctypes.byref(m)),
               "dlx_noll_indices")
    return int(n.value), int(m.value)


def _factorial(x: float) -> float:
    if x == 0:
        return 1.0
    if x < 0 and float(x).is_integer():
        return float("inf")
    return exp(lgamma(x + 1.0))


def zernike_factors(j: int):
    """Radial coefficients c and powers k (signed m).  utils/zernikes.py:122-150."""
    if j < 1:
        raise ValueError("The Zernike index must be greater than 0.")
    n, m = noll_indices(j)
    k = np.arange(((n - m) // 2) + 1, dtype=float)
    c = np.array([(-1.0) ** kk * _factorial(abs(n - kk)) / _factorial(kk)
                  / _factorial(((n + m) // 2) - kk) / _factorial(((n - m) // 2) - kk)
                  for kk in k])
    return c, k


def zernike_basis(js, coordinates, diameter: float = 2.0) -> torch.Tensor:
    """(len(js), *grid) float32 device basis.  utils/zernikes.py:297-315."""
    js = [int(j) for j in np.atleast_1d(np.asarray(js))]
    for j in js:
        if j < 1:
            raise ValueError("The Zernike index must be greater than 0.")
    coords = runtime.as_float(coordinates)
    if coords.shape[0] != 2:
        raise ValueError("coordinates must have a leading (x, y) axis of length 2")
    grid = tuple(coords.shape[1:])
    npix = int(np.prod(grid))
    x = coords[0].contiguous()
    y = coords[1].contiguous()
    out = torch.empty((len(js),) + grid, dtype=torch.float32, device=coords.device)
    arr = (ctypes.c_int32 * len(js))(*js)
    st = _lib.load().dlx_zernike_basis(len(js), arr, npix, x.data_ptr(), y.data_ptr(),
                                       float(diameter), out.data_ptr(),
                                       runtime.stream_handle())
    _lib.check(st, "dlx_zernike_basis")
    return out


def zernike(j: int, coordinates, diameter: float = 2.0) -> torch.Tensor:
    """One Noll Zernike on the disk of the given diameter.  utils/zernikes.py:227-257."""
    return zernike_basis([j], coordinates, diameter)[0]
