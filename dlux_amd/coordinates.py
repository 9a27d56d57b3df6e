"""Coordinate / sampling specifications and transformations (coordinates.py:23-316)."""

from __future__ import annotations

import numpy as np

from .base import Base

__all__ = ["Spec", "PadSpec", "CoordSpec", "BaseCoordTransform", "CoordTransform"]


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
