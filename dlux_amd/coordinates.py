"""Coordinate / sampling specifications and transformations (coordinates.py:23-316)."""

from __future__ import annotations

import numpy as np

from .base import Base

__all__ = ["Spec", "PadSpec", "CoordSpec", "BaseCoordTransform", "CoordTransform",
           "DistortedCoords"]


class Spec(Base):
    """Abstract sampling specification."""


class PadSpec(Spec):
    """Integer pad / crop factors relative to the input grid (coordinates.py:34-71)."""

    def __init__(self, pad=1, crop=1, c=0.0):
        self.pad = int(pad)
        self.crop = int(crop)
        self.c = np.asarray(c, float)


class CoordSpec(Spec):
    """Explicit (n, d, c) sampling (coordinates.py:74-157)."""

    def __init__(self, n=None, d=None, c=0.0):
        self.n = n
        self.d = None if d is None else np.asarray(d, float)
        self.c = None if c is None else np.asarray(c, float)

    @property
    def xs(self) -> np.ndarray:
        """c + (arange(n) - (n - 1)/2) * d  (host FP64)."""
        if self.d is None:
            raise ValueError("d must be specified to calculate coordinates.")
        return self.c + (np.arange(self.n) - (self.n - 1) / 2) * self.d

    @property
    def fov(self):
        if self.d is None:
            raise ValueError("d must be specified to calculate FOV.")
        return self.n * self.d

    @property
    def extent(self):
        if self.d is None:
            raise ValueError("d must be specified to calculate extent.")
        return self.c - (self.n / 2) * self.d, self.c + (self.n / 2) * self.d


class BaseCoordTransform(Base):
    """Abstract coordinate transformation (coordinates.py:160-227)."""

    def calculate(self, npix: int, diameter: float):
        from .utils.coordinates import pixel_coords
        return self(pixel_coords(npix, diameter))

    def __call__(self, coords):  # pragma: no cover - abstract
        raise NotImplementedError

    def apply(self, coords):
        return self(coords)


class CoordTransform(BaseCoordTransform):
    """Translation, then shear, compression and rotation (coordinates.py:230-316).

    The dynamic apertures do not call ``__call__``: they hand the parameters
    to the HIP kernels (``_struct``), which apply the same steps per pixel;
    ``__call__`` transforms a (2, n, n) device coordinate grid for direct use.
    """

    def __init__(self, translation=None, rotation=None, compression=None, shear=None):
        def check(v, shape, name):
            if v is None:
                return None
            if _is_tensor(v):   # kept as given, so gradients can reach it
                if tuple(v.shape) != shape:
                    raise ValueError(f"{name} must have shape {shape}.")
                return v
            v = np.asarray(v, dtype=float)
            if v.shape != shape:
                raise ValueError(f"{name} must have shape {shape}.")
            return v
        self.translation = check(translation, (2,), "translation")
        self.rotation = check(rotation, (), "rotation")
        self.compression = check(compression, (2,), "compression")
        self.shear = check(shear, (2,), "shear")

    def __call__(self, coords):
        from .utils import coordinates as uc
        if self.translation is not None:
            coords = uc.translate_coords(coords, self.translation)
        if self.shear is not None:
            coords = uc.shear_coords(coords, self.shear)
        if self.compression is not None:
            coords = uc.compress_coords(coords, self.compression)
        if self.rotation is not None:
            coords = uc.rotate_coords(coords, float(self.rotation))
        return coords

    def _struct(self):
        """The ``dlx_coord_transform`` of this transformation."""
        from . import _lib
        t = _lib.CoordTransformC()
        if self.translation is not None:
            t.flags |= _lib.DLX_TF_TRANSLATION
            t.translation[:] = list(_floats(self.translation))
        if self.shear is not None:
            t.flags |= _lib.DLX_TF_SHEAR
            t.shear[:] = list(_floats(self.shear))
        if self.compression is not None:
            t.flags |= _lib.DLX_TF_COMPRESSION
            t.compression[:] = list(_floats(self.compression))
        if self.rotation is not None:
            t.flags |= _lib.DLX_TF_ROTATION
            t.rotation = _floats(self.rotation)[0]
        return t

    def _key(self):
        return tuple(None if v is None else _floats(v)
                     for v in (self.translation, self.shear, self.compression, self.rotation))

    # (parameter slot, attribute, element) of the transform's gradient slots
    # in the dynamic-aperture geometry VJP
    _SLOTS = ((2, "translation", 0), (3, "translation", 1), (4, "shear", 0), (5, "shear", 1),
              (6, "compression", 0), (7, "compression", 1), (8, "rotation", None))


def _is_tensor(v) -> bool:
    return type(v).__module__.startswith("torch") and hasattr(v, "requires_grad")


def _floats(v) -> tuple:
    """The values of a parameter (numpy array, number or torch tensor) as a
    tuple of Python floats."""
    if _is_tensor(v):
        return tuple(v.detach().double().cpu().reshape(-1).tolist())
    return tuple(float(x) for x in np.atleast_1d(np.asarray(v, float)).reshape(-1))


class DistortedCoords(BaseCoordTransform):
    """A 2-D polynomial distortion coords + sum_m c[:, m] x^px_m y^py_m over the
    monomials of gen_powers(order + 1) without the constant
    (coordinates.py:318-357).  On the device its coordinates are evaluated as
    FP64 planes that the aperture and resampling kernels read in place of a
    transform chain."""

    def __init__(self, order: int = 1, distortion=None):
        from .utils.coordinates import gen_powers
        self.powers = np.array(gen_powers(int(order) + 1))[:, 1:]
        if distortion is None:
            distortion = np.zeros(self.powers.shape)
        if not _is_tensor(distortion):
            distortion = np.asarray(distortion, dtype=float)
        if tuple(distortion.shape) != self.powers.shape:
            raise ValueError("distortion shape must match powers shape.")
        self.distortion = distortion

    def __call__(self, coords):
        from .utils.coordinates import distort_coords
        return distort_coords(coords, self.distortion, self.powers)

    def _key(self):
        return ("distorted", tuple(self.powers.ravel().tolist()), _floats(self.distortion))


def _untracked(tf, what: str) -> None:
    """Raise when a parameter of `tf` is a tensor that requires grad while
    gradients are recorded: the resampling kernels take the transformation
    as constants."""
    import torch
    if not torch.is_grad_enabled():
        return
    for f in ("translation", "rotation", "compression", "shear", "distortion"):
        v = getattr(tf, f, None)
        if _is_tensor(v) and v.requires_grad:
            raise NotImplementedError(f"gradients with respect to the transformation of {what} "
                                      "are not implemented (the images' gradient is)")


def _chain_coords(tfs, npix: int, pixel_scale: float, centre: float, dev):
    """The (2, npix, npix) float64 device planes of the pixel grid
    centre + (k - (npix-1)/2) pixel_scale put through the transforms in order
    (CoordTransform: translation, shear of the transposed planes, compression,
    rotation; DistortedCoords: the polynomial)."""
    import torch
    xs = centre + (torch.arange(npix, dtype=torch.float64, device=dev) - (npix - 1) / 2.0) \
        * pixel_scale
    gy, gx = torch.meshgrid(xs, xs, indexing="ij")
    c = torch.stack([gx, gy])
    for t in tfs:
        if isinstance(t, CoordTransform):
            if t.translation is not None:
                c = c - torch.tensor(_floats(t.translation), dtype=torch.float64,
                                     device=dev)[:, None, None]
            if t.shear is not None:
                sh = torch.tensor(_floats(t.shear), dtype=torch.float64, device=dev)
                c = c + c.transpose(1, 2) * sh[:, None, None]
            if t.compression is not None:
                c = c * torch.tensor(_floats(t.compression), dtype=torch.float64,
                                     device=dev)[:, None, None]
            if t.rotation is not None:
                r = _floats(t.rotation)[0]
                co, si = np.cos(-r), np.sin(-r)
                c = torch.stack([co * c[0] + si * c[1], -si * c[0] + co * c[1]])
        elif isinstance(t, DistortedCoords):
            d = torch.tensor(np.asarray(_floats(t.distortion)).reshape(t.powers.shape),
                             dtype=torch.float64, device=dev)
            px = torch.as_tensor(t.powers[0], device=dev)
            py = torch.as_tensor(t.powers[1], device=dev)
            basis = c[0][None] ** px[:, None, None] * c[1][None] ** py[:, None, None]
            c = c + torch.tensordot(d, basis, dims=([1], [0]))
        else:
            raise TypeError(f"unsupported coordinate transformation {type(t).__name__}")
    return c.contiguous()
