"""Image resampling (utils/interpolation.py:13-108) on the device.

interpax is not available: "linear" is the textbook bilinear rule on the knot
grid and "nearest" takes the nearest knot; points outside the knots' span
take the fill value (interp2d's extrap = fill).  Knot grids are regular
(pixel_coords / nd_coords), so a transformation of them runs inside the
kernel (dlx_interp_transform); explicit sample points use dlx_interp_points.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import runtime

__all__ = ["interp", "scale", "rotate"]


def _grid(knot_coords):
    """(n, spacing, centre) of a regular (2, n, n) knot grid."""
    c = knot_coords if isinstance(knot_coords, torch.Tensor) else \
        torch.as_tensor(np.asarray(knot_coords, np.float64))
    if c.ndim != 3 or c.shape[0] != 2 or c.shape[1] != c.shape[2]:
        raise ValueError(f"knot coordinates must be a (2, n, n) grid, got {tuple(c.shape)}")
    x = c[0, 0].double().cpu().numpy()
    n = x.size
    d = float(x[1] - x[0]) if n > 1 else 1.0
    return n, d, float(0.5 * (x[0] + x[-1]))


def interp(image, knot_coords, sample_coords, method: str = "linear", fill: float = 0.0):
    """`image` (n x n, knots on the regular grid knot_coords) sampled at
    sample_coords (2, ...) (utils/interpolation.py:13-50)."""
    n, d, c = _grid(knot_coords)
    sc = sample_coords if isinstance(sample_coords, torch.Tensor) else \
        torch.as_tensor(np.asarray(sample_coords, np.float64))
    return runtime.interp_points(image, d, c, sc[0], sc[1], method, fill)


def scale(array, npixels: int, ratio: float, method: str = "linear"):
    """Paraxial resampling of a square array to npixels with sampling ratio
    `ratio` (utils/interpolation.py:53-79): the output grid pixel_coords(npixels,
    1), compressed by ratio * npixels / n_in, sampled from pixel_coords(n_in, 1)."""
    from ..coordinates import CoordTransform
    n_in = int(array.shape[-1])
    r = float(ratio) * int(npixels) / n_in
    tf = CoordTransform(compression=[r, r])._struct()
    return runtime.interp_transform(array, 1.0 / n_in, 0.0, int(npixels), 1.0 / int(npixels), 0.0,
                                    [tf], method)


def rotate(array, angle, method: str = "linear"):
    """A square array rotated by `angle` (radians) about its centre
    (utils/interpolation.py:82-108): nd_coords (unit pixels) through
    rotate_coords."""
    from ..coordinates import CoordTransform
    n = int(array.shape[-1])
    tf = CoordTransform(rotation=float(np.asarray(angle)))._struct()
    return runtime.interp_transform(array, 1.0, 0.0, n, 1.0, 0.0, [tf], method)
