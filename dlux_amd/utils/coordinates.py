"""Pixel-centre coordinate grids on the device (utils/coordinates.py:81-340).

These are wavelength-independent set-up arrays; they are produced as float32
device tensors in the reference's 'xy' meshgrid order (coords[0] = x varies
along axis 1).
"""

from __future__ import annotations

import numpy as np
import torch

from .. import runtime


def _cast_tuple(v, name):
    if isinstance(v, (int, np.integer)):
        return (int(v),)
    if isinstance(v, (tuple, list)):
        return tuple(int(x) for x in v)
    raise TypeError(f"{name} must be an int or tuple of ints.")


def nd_coords(npixels, pixel_scales=1.0, offsets=0.0, indexing: str = "xy") -> torch.Tensor:
    """Pixel-centre coordinates linspace(-(n-1)/2 d - o, (n-1)/2 d - o, n) per axis.

    utils/coordinates.py:269-340.
    """
    if indexing not in ["xy", "ij"]:
        raise ValueError("indexing must be either 'xy' or 'ij'.")
    npix = _cast_tuple(npixels, "npixels")
    scales = np.atleast_1d(np.asarray(pixel_scales, np.float64))
    offs = np.atleast_1d(np.asarray(offsets, np.float64))
    ndim = max(len(npix), scales.size, offs.size)
    if len(npix) != ndim:
        npix = npix * ndim
    scales = np.broadcast_to(scales, (ndim,))
    offs = np.broadcast_to(offs, (ndim,))
    lins = []
    for n, d, o in zip(npix, scales, offs):
        lins.append(np.linspace(-(n - 1) / 2 * d - o, (n - 1) / 2 * d - o, n))
    pos = np.array(np.meshgrid(*lins, indexing=indexing))
    return torch.as_tensor(np.squeeze(pos).astype(np.float32), device=runtime.device())


def pixel_coords(npixels: int, diameter: float = None, radius: float = None,
                 pixel_scale: float = None, polar: bool = False,
                 fft_style: bool = False) -> torch.Tensor:
    """(x, y) coordinates of each pixel centre.  utils/coordinates.py:208-266."""
    supplied = sum(v is not None for v in (diameter, radius, pixel_scale))
    if supplied != 1:
        raise ValueError("Exactly one of diameter, radius, or pixel_scale must be provided.")
    if diameter is not None:
        d = diameter / npixels
    elif radius is not None:
        d = 2 * radius / npixels
    else:
        d = pixel_scale
    off = (d / 2.0, d / 2.0) if (fft_style and npixels % 2 == 0) else (0.0, 0.0)
    coords = nd_coords((npixels,) * 2, (d,) * 2, offsets=off, indexing="xy")
    return cart2polar(coords) if polar else coords


def cart2polar(coordinates: torch.Tensor) -> torch.Tensor:
    """utils/coordinates.py:166-184."""
    x, y = coordinates
    return torch.stack([torch.hypot(x, y), torch.atan2(y, x)])


def polar2cart(coordinates: torch.Tensor) -> torch.Tensor:
    """utils/coordinates.py:187-205."""
    r, phi = coordinates
    return torch.stack([r * torch.cos(phi), r * torch.sin(phi)])


def translate_coords(coords, translation):
    """coords - translation per axis (utils/coordinates.py:23-39)."""
    t = torch.as_tensor(np.asarray(translation, np.float32), device=coords.device)
    return coords - t[:, None, None]


def compress_coords(coords, compress):
    """coords * compress per axis (utils/coordinates.py:42-58)."""
    c = torch.as_tensor(np.asarray(compress, np.float32), device=coords.device)
    return coords * c[:, None, None]


def shear_coords(coords, shear):
    """coords + transpose(coords) * shear per axis, with each coordinate plane
    transposed (utils/coordinates.py:61-78)."""
    s = torch.as_tensor(np.asarray(shear, np.float32), device=coords.device)
    return coords + coords.transpose(1, 2) * s[:, None, None]


def rotate_coords(coords: torch.Tensor, rotation: float) -> torch.Tensor:
    """utils/coordinates.py:81-101."""
    x, y = coords
    c, s = float(np.cos(-rotation)), float(np.sin(-rotation))
    return torch.stack([c * x + s * y, -s * x + c * y])


def gen_powers(degree: int):
    """(x powers, y powers) of the 2-D monomials up to total degree - 1, in the
    reference's order (utils/coordinates.py:104-133)."""
    n = int(degree) * (int(degree) + 1) // 2
    vals = np.arange(n)
    tri = lambda k: k * (k + 1) // 2
    ydiffs = np.repeat(tri(np.arange(degree)), np.arange(1, degree + 1))
    ypows = vals - ydiffs
    xdiffs = np.repeat(n - np.flip(tri(np.arange(1, degree + 1))), np.arange(degree, 0, -1))
    xpows = np.flip(vals - xdiffs)
    return np.array([xpows, ypows])


def polynomial_basis(coords, pows):
    """x^px y^py for each column of pows, over coords (2, ...)
    (utils/coordinates.py:159-163)."""
    pows = np.asarray(pows)
    c = coords if isinstance(coords, torch.Tensor) else torch.as_tensor(np.asarray(coords))
    p = torch.as_tensor(pows, device=c.device).reshape(pows.shape + (1,) * (c.ndim - 1))
    return torch.prod(c[:, None] ** p, dim=0)


def distort_coords(coords, coeffs, pows):
    """coords + tensordot(coeffs, polynomial_basis(coords, pows))
    (utils/coordinates.py:136-156)."""
    c = coords if isinstance(coords, torch.Tensor) else torch.as_tensor(np.asarray(coords))
    k = coeffs if isinstance(coeffs, torch.Tensor) else torch.as_tensor(np.asarray(coeffs))
    k = k.to(device=c.device, dtype=c.dtype)
    return c + torch.tensordot(k, polynomial_basis(c, pows).to(c.dtype), dims=([1], [0]))
