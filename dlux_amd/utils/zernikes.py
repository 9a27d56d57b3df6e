"""Zernike polynomials (utils/zernikes.py:99-315); bases are evaluated on the device."""

from __future__ import annotations

import ctypes
from math import lgamma, exp

import numpy as np
import torch

from .. import _lib, runtime

__all__ = ["noll_indices", "zernike_factors", "zernike", "zernike_basis", "zernike_name",
           "radial_orders_to_indices", "polike", "polike_basis"]

# Names of the first 36 Noll modes (the table of utils/zernikes.py:22-67; a
# naming convention, so the same strings).
_NAMES = {
    1: "Piston", 2: "Tilt X", 3: "Tilt Y", 4: "Defocus", 5: "Astig X", 6: "Astig Y",
    7: "Coma X", 8: "Coma Y", 9: "Trefoil X", 10: "Trefoil Y", 11: "Spherical",
    12: "2nd Astig X", 13: "2nd Astig Y", 14: "Quadrafoil X", 15: "Quadrafoil Y",
    16: "2nd Coma X", 17: "2nd Coma Y", 18: "2nd Trefoil X", 19: "2nd Trefoil Y",
    20: "Pentafoil X", 21: "Pentafoil Y", 22: "2nd Spherical", 23: "3rd Coma X",
    24: "3rd Coma Y", 25: "3rd Astig X", 26: "3rd Astig Y", 27: "Hexafoil X",
    28: "Hexafoil Y", 29: "4th Coma X", 30: "4th Coma Y", 31: "4th Astig X",
    32: "4th Astig Y", 33: "3rd Trefoil X", 34: "3rd Trefoil Y", 35: "Heptafoil X",
    36: "Heptafoil Y",
}


def zernike_name(j: int) -> str:
    """utils/zernikes.py:81-96."""
    j = int(j)
    return _NAMES[j] if 1 <= j <= 36 else f"Zernike {j}"


def radial_orders_to_indices(radial_orders) -> list:
    """Every Noll index of the given radial orders, in order
    (utils/zernikes.py:70-78)."""
    indices = []
    for order in radial_orders:
        order = int(order)
        if order < 0:
            raise ValueError("radial_orders must contain non-negative integers.")
        first = order * (order + 1) // 2 + 1
        indices.extend(range(first, first + order + 1))
    return indices


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


def _wedge_radius(nsides: int, coords: torch.Tensor) -> torch.Tensor:
    """r_alpha = cos(alpha) / cos(u) of the n-gon wedge each point lies in
    (utils/zernikes.py:318-349), on the device in float64."""
    alpha = np.pi / nsides
    phi = torch.atan2(coords[1], coords[0]) + alpha
    wedge = torch.floor((phi + alpha) / (2.0 * alpha))
    u = phi - wedge * (2 * alpha)
    return np.cos(alpha) / torch.cos(u)


def polike_basis(nsides: int, js, coordinates, diameter: float = 2.0) -> torch.Tensor:
    """Zernikes on the regular n-gon: (1/r_alpha) Z_j(coords / r_alpha) on
    coordinates scaled by diameter/2 (utils/zernikes.py:318-349, :398-420)."""
    if nsides < 3:
        raise ValueError(f"nsides must be >= 3, not {nsides}.")
    c = runtime.as_float(coordinates).double() / (float(diameter) / 2.0)
    r_a = _wedge_radius(int(nsides), c)
    basis = zernike_basis(js, c / r_a, 2.0)
    return (basis.double() / r_a).float()


def polike(nsides: int, j: int, coordinates, diameter: float = 2.0) -> torch.Tensor:
    """One Noll Zernike on the regular n-gon (utils/zernikes.py:318-349)."""
    return polike_basis(nsides, [j], coordinates, diameter)[0]
