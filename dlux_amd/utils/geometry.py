"""Hard-edged mask primitives on device coordinate grids (utils/geometry.py:20-485)."""

from __future__ import annotations

import numpy as np
import torch

from .coordinates import cart2polar, rotate_coords
from .array_ops import downsample

__all__ = ["combine", "circle", "spider", "circ_distance", "spider_distance"]


def combine(arrays, oversample: int = 1, use_sum: bool = False) -> torch.Tensor:
    """Product (or sum) of masks, then block-mean by `oversample`.  geometry.py:20-46."""
    arr = torch.stack([torch.as_tensor(a) for a in arrays])
    out = arr.sum(0) if use_sum else arr.prod(0)
    return out if oversample == 1 else downsample(out, oversample)


def circ_distance(coords, radius):
    """geometry.py:405-418."""
    return cart2polar(coords)[0] - radius


def circle(coords, radius: float, invert: bool = False) -> torch.Tensor:
    """Hard circle r < R (r > R if invert).  geometry.py:95-117."""
    d = circ_distance(coords, radius)
    return (d > 0).float() if invert else (d < 0).float()


def spider_distance(coords, width: float, angle: float):
    """max(|x'| - w/2, y') after rotation by `angle` degrees.  geometry.py:463-485."""
    x, y = rotate_coords(coords, float(np.deg2rad(angle)))
    return torch.maximum(torch.abs(x) - width / 2, y)


def spider(coords, width: float, angles) -> torch.Tensor:
    """1 outside every one-sided vane.  geometry.py:195-218."""
    blocked = torch.zeros(coords.shape[1:], dtype=torch.bool, device=coords.device)
    for a in np.atleast_1d(np.asarray(angles, np.float64)):
        blocked |= spider_distance(coords, width, a) < 0
    return (~blocked).float()
