"""Dynamic apertures (layers/apertures.py:26-1117) generated on the device.

Soft-edged transmissions come from dlx_soft_apertures: per aperture a signed
edge distance on the (transformed) pixel grid, clipped and rescaled by its own
extrema (the reference's ``soften``), combined as a product (CompoundAperture)
or a sum (MultiAperture).  Aberrated apertures evaluate their Zernike (or
n-sided polike) basis on the aperture's own transformed, extent-normalised
coordinates with dlx_aperture_basis.  Transmissions and bases are cached per
(grid, parameters), so a system whose aperture does not change re-uses them
and one whose aperture moves regenerates them on the device, never on the host.

All of them lower into the fused pupil plan (amplitude, normalisation and OPD
terms), so an optical system built from them keeps the single dlx_psf_poly
path.  Gradients flow to AberratedAperture coefficients and, through
dlx_soft_apertures_vjp, to the geometry of the soft apertures: a size
(radius, width, height, rmax) or a parameter of the innermost transformation
that applies to an aperture, given as a torch tensor that requires grad.
Gradients with respect to softening, spider angles, an outer transformation
under an aperture's own, or the geometry of an aberrated aperture's basis
raise NotImplementedError.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib, runtime
from ..polynomials import DynamicZernikeBasis
from ..coordinates import (BaseCoordTransform, CoordTransform, DistortedCoords, _chain_coords,
                           _floats, _is_tensor)
from .optical_layers import OpticalLayer

__all__ = ["ApertureLayer", "BaseDynamicAperture", "DynamicAperture", "CircularAperture",
           "SquareAperture", "RectangularAperture", "RegPolyAperture", "Spider",
           "AberratedAperture", "CompositeAperture", "CompoundAperture", "MultiAperture"]


def _grid_of(coords):
    """(npix, centre) of a (2, n, n) pixel-centre grid (the coordinates the
    reference's apertures are evaluated on: wavefront.coordinates())."""
    c = coords if isinstance(coords, torch.Tensor) else torch.as_tensor(np.asarray(coords))
    if c.ndim != 3 or c.shape[0] != 2 or c.shape[1] != c.shape[2]:
        raise ValueError(f"coordinates must be a (2, n, n) pixel grid, got {tuple(c.shape)}")
    x = c[0, 0]
    centre = float(0.5 * (x[0] + x[-1]))
    return int(c.shape[-1]), centre


def _param(v):
    """A geometric parameter: a torch tensor is kept (gradients can reach it),
    anything else becomes a float64 numpy array."""
    return v if _is_tensor(v) else np.asarray(v, float)


def _tracked(v) -> bool:
    return _is_tensor(v) and v.requires_grad


def _grad_sources(items):
    """(tensors, map) for the geometry VJP: the distinct tracked tensors and,
    per (aperture, slot) that feeds one, (aperture, slot, tensor index,
    element or None)."""
    params, smap = [], []
    for a, (_, _, src) in enumerate(items):
        for slot, t, elem in src:
            for i, q in enumerate(params):
                if q is t:
                    break
            else:
                params.append(t)
                i = len(params) - 1
            smap.append((a, slot, i, elem))
    return params, smap


class ApertureLayer(OpticalLayer):
    """Base aperture: the ``normalise`` flag (layers/apertures.py:26-71)."""

    def __init__(self, normalise: bool = False):
        self.normalise = bool(normalise)

    def transmission(self, coords, pixel_scale):  # pragma: no cover - abstract
        raise NotImplementedError

    # --- device generation -------------------------------------------------
    def _descs(self, transforms):
        """dlx_soft_aperture descriptors of this aperture under the enclosing
        transforms, as [(group key, descriptor)]."""
        raise NotImplementedError

    def _cache(self, key, make):
        cached = self.__dict__.get("_gen")
        if cached is not None and cached[0] == key:
            return cached[1]
        value = make()
        self.__dict__["_gen"] = (key, value)
        return value

    def _grid_transmission(self, npix: int, pixel_scale: float, centre: float = 0.0):
        items = self._descs(())
        params, smap = _grad_sources(items)
        keep = _explicit_coords(items, npix, pixel_scale, centre)
        if params and torch.is_grad_enabled():
            out = runtime.soft_apertures_tracked(npix, pixel_scale, _grouped(items), centre,
                                                 params, smap, keep)
        else:
            key = ("t", npix, float(pixel_scale), float(centre), self._param_key())
            out = self._cache(key, lambda: runtime.soft_apertures(
                npix, pixel_scale, _grouped(items), centre))
        del keep
        return out

    def _param_key(self):
        raise NotImplementedError


def _explicit_coords(items, npix, pixel_scale, centre):
    """Descriptors whose chain holds a DistortedCoords get their coordinates as
    FP64 device planes (returned, to keep them alive for the call)."""
    keep = []
    for _, d, _ in items:
        chain = getattr(d, "_chain", None)
        if chain is not None:
            c = _chain_coords(chain, npix, float(pixel_scale), float(centre), runtime.device())
            keep.append(c)
            d.coords = c.data_ptr()
    return keep


def _grouped(descs):
    """Descriptors with consecutive integer group ids (same key -> same group)."""
    out, ids = [], {}
    for key, d, _ in descs:
        d.group = ids.setdefault(key, len(ids))
        out.append(d)
    return out


class BaseDynamicAperture(ApertureLayer):
    """An aperture with an optional coordinate transformation (layers/apertures.py:74-152)."""

    def __init__(self, transformation=None, normalise: bool = False):
        super().__init__(normalise=normalise)
        if transformation is not None and not isinstance(transformation, BaseCoordTransform):
            raise TypeError(
                f"transformation must be a BaseCoordTransform instance, "
                f"got {type(transformation).__name__}. Import with: "
                "from dLux.coordinates import BaseCoordTransform. "
                "Use a BaseCoordTransform-compatible type from this module."
            )
        self.transformation = transformation

    def __getattr__(self, key):
        tf = self.__dict__.get("transformation")
        if tf is not None and hasattr(tf, key):
            return getattr(tf, key)
        raise AttributeError(f"{type(self).__name__} has no attribute {key!r}")

    def _transforms(self, outer):
        return tuple(outer) + ((self.transformation,) if self.transformation is not None
                               else ())

    def _tkey(self):
        return None if self.transformation is None else self.transformation._key()

    def transmission(self, coords, pixel_scale):
        npix, centre = _grid_of(coords)
        return self._grid_transmission(npix, float(pixel_scale), centre)

    def __call__(self, wavefront):
        wavefront *= self._grid_transmission(wavefront.npixels, float(np.mean(
            wavefront.pixel_scale)), float(np.asarray(wavefront.center).reshape(-1)[0]))
        if self.normalise:
            return wavefront.normalise()
        return wavefront

    def _lower(self, plan) -> bool:
        plan.mul_amp(self._grid_transmission(plan.n, plan.dx))
        if self.normalise:
            plan.normalise()
        return True


class DynamicAperture(BaseDynamicAperture):
    """Soft edges and occultation (layers/apertures.py:155-237)."""

    _kind = 0

    def __init__(self, transformation=None, occulting: bool = False, softening=1.0,
                 normalise: bool = False):
        super().__init__(transformation=transformation, normalise=normalise)
        self.occulting = bool(occulting)
        self.softness = _param(softening)
        if _floats(self.softness)[0] <= 0:
            raise ValueError("softening must be greater than 0.")

    def _fill(self, d):  # pragma: no cover - per shape
        raise NotImplementedError

    def _descs(self, transforms, group=None):
        d = _lib.SoftAperture()
        d.kind = self._kind
        d.occulting = int(self.occulting)
        d.softening = _floats(self.softness)[0]
        tfs = self._transforms(transforms)
        if all(isinstance(t, CoordTransform) for t in tfs):
            if len(tfs) > 3:
                raise NotImplementedError("at most three nested coordinate transformations")
            d.n_transforms = len(tfs)
            for k, t in enumerate(tfs):
                d.transforms[k] = t._struct()
        else:   # the kernels read the coordinates as planes (see _explicit_coords)
            d.n_transforms = 0
            d._chain = tfs
        self._fill(d)
        return [(id(self) if group is None else group, d, self._sources(tfs))]

    def _sizes(self):
        """(slot, parameter) of the sizes the geometry VJP differentiates."""
        return ()

    def _sources(self, tfs):
        """(slot, tensor, element) of the tracked parameters of this aperture."""
        if torch.is_grad_enabled():
            if _tracked(self.softness):
                raise NotImplementedError("gradients with respect to softening are not "
                                          "implemented")
            for t in tfs[:-1]:
                if any(_tracked(getattr(t, f, None)) for _, f, _ in CoordTransform._SLOTS):
                    raise NotImplementedError(
                        "gradients with respect to a transformation are implemented for the "
                        "innermost transformation that applies to an aperture only")
        src = [(slot, v, None) for slot, v in self._sizes() if _tracked(v)]
        if any(isinstance(t, DistortedCoords) for t in tfs):
            if torch.is_grad_enabled() and any(
                    _tracked(getattr(t, f, None)) for t in tfs
                    for f in ("distortion", "translation", "shear", "compression", "rotation")):
                raise NotImplementedError("gradients with respect to a transformation chain "
                                          "holding a DistortedCoords are not implemented")
            return src
        if tfs and isinstance(tfs[-1], CoordTransform):
            for slot, f, elem in CoordTransform._SLOTS:
                v = getattr(tfs[-1], f)
                if _tracked(v):
                    src.append((slot, v, elem))
        return src

    def _param_key(self):
        return (type(self).__name__, self.occulting, _floats(self.softness), self._tkey(),
                self._shape_key())


class CircularAperture(DynamicAperture):
    """layers/apertures.py:240-311."""

    _kind = _lib.DLX_AP_CIRCLE

    def __init__(self, radius, transformation=None, occulting: bool = False, softening=1.0,
                 normalise: bool = False):
        super().__init__(transformation, occulting, softening, normalise)
        self.radius = _param(radius)

    def _fill(self, d):
        d.size[0] = _floats(self.radius)[0]

    def _sizes(self):
        return ((0, self.radius),)

    def _shape_key(self):
        return _floats(self.radius)

    @property
    def extent(self):
        return self.radius

    @property
    def nsides(self) -> int:
        return 0


class SquareAperture(DynamicAperture):
    """layers/apertures.py:314-387."""

    _kind = _lib.DLX_AP_SQUARE

    def __init__(self, width, transformation=None, occulting: bool = False, softening=1.0,
                 normalise: bool = False):
        super().__init__(transformation, occulting, softening, normalise)
        self.width = _param(width)

    def _fill(self, d):
        d.size[0] = _floats(self.width)[0]

    def _sizes(self):
        return ((0, self.width),)

    def _shape_key(self):
        return _floats(self.width)

    @property
    def extent(self):
        return np.sqrt(2) * self.width

    @property
    def nsides(self) -> int:
        return 4


class RectangularAperture(DynamicAperture):
    """layers/apertures.py:390-472 (height first, as the reference)."""

    _kind = _lib.DLX_AP_RECTANGLE

    def __init__(self, height, width, transformation=None, occulting: bool = False,
                 softening=1.0, normalise: bool = False):
        super().__init__(transformation, occulting, softening, normalise)
        self.height = _param(height)
        self.width = _param(width)

    def _fill(self, d):
        d.size[0] = _floats(self.width)[0]
        d.size[1] = _floats(self.height)[0]

    def _sizes(self):
        return ((0, self.width), (1, self.height))

    def _shape_key(self):
        return _floats(self.width) + _floats(self.height)

    @property
    def extent(self):
        return (self.height ** 2 + self.width ** 2) ** 0.5 / 2.0

    @property
    def nsides(self) -> int:
        return 4


class RegPolyAperture(DynamicAperture):
    """layers/apertures.py:475-555: vertices on radius rmax at angles
    linspace(-pi, pi, nsides, endpoint=False)."""

    _kind = _lib.DLX_AP_REG_POLYGON

    def __init__(self, nsides: int, rmax, transformation=None, occulting: bool = False,
                 softening=1.0, normalise: bool = False):
        super().__init__(transformation, occulting, softening, normalise)
        self._nsides = int(nsides)
        self.rmax = _param(rmax)

    def _fill(self, d):
        d.size[0] = _floats(self.rmax)[0]
        d.nsides = self._nsides

    def _sizes(self):
        return ((0, self.rmax),)

    def _shape_key(self):
        return _floats(self.rmax) + (self._nsides,)

    @property
    def extent(self):
        return self.rmax

    @property
    def nsides(self) -> int:
        return self._nsides


class Spider(DynamicAperture):
    """One-sided arms of a given width at the given angles (degrees);
    occulting by default (layers/apertures.py:558-640)."""

    _kind = _lib.DLX_AP_SPIDER

    def __init__(self, width, angles, transformation=None, occulting: bool = True,
                 softening=1.0, normalise: bool = False):
        super().__init__(transformation, occulting, softening, normalise)
        self.width = _param(width)
        self.angles = _param(angles)

    def _fill(self, d):
        angles = _floats(self.angles)
        if not 1 <= len(angles) <= 16:
            raise ValueError(f"a spider needs 1 to 16 arms, got {len(angles)}")
        d.size[0] = _floats(self.width)[0]
        d.n_angles = len(angles)
        for k, a in enumerate(angles):
            d.angles_deg[k] = a

    def _sizes(self):
        if torch.is_grad_enabled() and _tracked(self.angles):
            raise NotImplementedError("gradients with respect to spider angles are not "
                                      "implemented")
        return ((0, self.width),)

    def _shape_key(self):
        return _floats(self.width) + (None,) + _floats(self.angles)

    @property
    def extent(self):
        raise TypeError("Spiders do not have an extent.")

    @property
    def nsides(self):
        raise TypeError("Spiders do not have a number of sides.")


class AberratedAperture(ApertureLayer):
    """A dynamic aperture with a Zernike (or polike) basis on its own
    normalised coordinates (layers/apertures.py:643-800)."""

    def __init__(self, aperture, noll_inds, coefficients=None, effect: str = "opd"):
        if not isinstance(aperture, DynamicAperture):
            raise TypeError(
                "AberratedApertures cannot contain Static, "
                "Compound or Multi Apertures. AberratedApertures can be "
                "placed in Compound or Multi Apertures, which can then be "
                "promoted to Static."
            )
        super().__init__(normalise=aperture.normalise)
        if effect not in ("opd", "phase", "amplitude"):
            raise ValueError("effect must be 'opd', 'phase', or 'amplitude'.")
        self.effect = effect
        if aperture.occulting:
            raise TypeError("AberratedApertures cannot be occulting.")
        if isinstance(aperture, Spider):
            raise TypeError("AberratedApertures cannot be spiders.")
        self.aperture = aperture
        self.noll_inds = [int(j) for j in np.atleast_1d(np.asarray(noll_inds))]
        for j in self.noll_inds:
            if j < 1:
                raise ValueError("Zernike indices must be greater than zero.")
        if coefficients is None:
            coefficients = np.zeros(len(self.noll_inds), np.float32)
        self.aberrations = DynamicZernikeBasis(js=self.noll_inds, coefficients=coefficients,
                                               nsides=aperture.nsides)

    @property
    def coefficients(self):
        """The coefficients of the Zernike parameterisation."""
        return self.aberrations.coefficients

    @coefficients.setter
    def coefficients(self, value):
        self.aberrations = self.aberrations.set("coefficients", value)

    def calculate(self):
        raise NotImplementedError(
            "Aberrated Apertures cannot use the .calculate() method because "
            "they need coords to be generated on. please use "
            ".eval_basis(coords) method instead."
        )

    def transmission(self, coords, pixel_scale):
        return self.aperture.transmission(coords, pixel_scale)

    def _descs(self, transforms, group=None):
        return self.aperture._descs(transforms, group)

    def _param_key(self):
        return self.aperture._param_key()

    def _grid_basis(self, npix, pixel_scale, centre=0.0, outer=()):
        if torch.is_grad_enabled() and (
                any(_tracked(v) for _, v in self.aperture._sizes())
                or any(_tracked(getattr(t, f, None)) for t in
                       tuple(outer) + (self.aperture.transformation,) if t is not None
                       for _, f, _ in CoordTransform._SLOTS)):
            raise NotImplementedError("gradients with respect to the geometry of an "
                                      "aberrated aperture's basis are not implemented")
        chain = tuple(outer) + ((self.aperture.transformation,)
                                if self.aperture.transformation is not None else ())
        explicit = not all(isinstance(t, CoordTransform) for t in chain)
        tfs = () if explicit else tuple(t._struct() for t in chain)
        key = ("b", npix, float(pixel_scale), float(centre), self.aperture._param_key(),
               tuple(self.noll_inds), tuple(bytes(t) for t in tfs),
               tuple(t._key() for t in chain) if explicit else None)
        cached = self.__dict__.get("_basis")
        if cached is not None and cached[0] == key:
            return cached[1]
        coords = _chain_coords(chain, npix, float(pixel_scale), float(centre),
                               runtime.device()) if explicit else None
        basis = runtime.aperture_basis(npix, pixel_scale, self.noll_inds, tfs,
                                       _floats(self.aperture.extent)[0], self.aperture.nsides,
                                       centre, coords=coords)
        self.__dict__["_basis"] = (key, basis)
        return basis

    def calc_basis(self, coords):
        npix, centre = _grid_of(coords)
        dx = float(coords[0, 0, 1] - coords[0, 0, 0])
        return self._grid_basis(npix, dx, centre)

    def eval_basis(self, coords):
        return runtime.eval_basis(self.calc_basis(coords), self.coefficients)

    def _aberration(self, npix, pixel_scale, centre=0.0, outer=()):
        return runtime.eval_basis(self._grid_basis(npix, pixel_scale, centre, outer),
                                  self.coefficients)

    def __call__(self, wavefront):
        dx = float(np.mean(wavefront.pixel_scale))
        centre = float(np.asarray(wavefront.center).reshape(-1)[0])
        wavefront *= self.aperture._grid_transmission(wavefront.npixels, dx, centre)
        if self.normalise:
            wavefront = wavefront.normalise()
        ab = self._aberration(wavefront.npixels, dx, centre)
        if self.effect == "phase":
            return wavefront.add_phase(ab)
        if self.effect == "opd":
            return wavefront.add_opd(ab)
        return wavefront * (1 + ab)

    def _lower(self, plan) -> bool:
        plan.mul_amp(self.aperture._grid_transmission(plan.n, plan.dx))
        if self.normalise:
            plan.normalise()
        ab = self._aberration(plan.n, plan.dx)
        if self.effect == "phase":
            plan.add_phase(ab)
        elif self.effect == "opd":
            plan.add_opd(ab)
        else:
            plan.mul_amp(1 + ab)
        return True


class CompositeAperture(BaseDynamicAperture):
    """Several apertures under one transformation (layers/apertures.py:803-1002)."""

    def __init__(self, apertures, transformation=None, normalise: bool = False):
        from ..optical_systems import list2dictionary
        self.apertures = list2dictionary(apertures, False, ApertureLayer)
        super().__init__(transformation=transformation, normalise=normalise)

    def __getattr__(self, key):
        aps = self.__dict__.get("apertures")
        if aps is not None and key in aps:
            return aps[key]
        raise AttributeError(
            f"{type(self).__name__} has no attribute {key!r} "
            f"(apertures: {list(aps) if aps is not None else []})")

    def _aberrated_apertures(self):
        out = []
        for ap in self.apertures.values():
            if isinstance(ap, AberratedAperture):
                out.append(ap)
            elif isinstance(ap, CompositeAperture) and ap._aberrated_apertures():
                out.append(ap)
        return out

    @property
    def coefficients(self):
        return [ap.coefficients for ap in self._aberrated_apertures()]

    def _param_key(self):
        return (type(self).__name__, self._tkey(),
                tuple(ap._param_key() for ap in self.apertures.values()))

    def _aberration(self, npix, pixel_scale, centre=0.0, outer=()):
        """Sum of the aberrated members' OPDs.  As in the reference
        (CompositeAperture.calc_basis, layers/apertures.py:911-929), the basis
        coordinates carry only the outermost composite's transformation and
        the member aperture's own, not those of intermediate composites."""
        outer = tuple(outer) or ((self.transformation,)
                                 if self.transformation is not None else ())
        total = None
        for ap in self._aberrated_apertures():
            ab = ap._aberration(npix, pixel_scale, centre, outer)
            total = ab if total is None else total + ab
        return total

    def __call__(self, wavefront):
        dx = float(np.mean(wavefront.pixel_scale))
        centre = float(np.asarray(wavefront.center).reshape(-1)[0])
        wavefront *= self._grid_transmission(wavefront.npixels, dx, centre)
        if self._aberrated_apertures():
            wavefront = wavefront.add_opd(self._aberration(wavefront.npixels, dx, centre))
        if self.normalise:
            wavefront = wavefront.normalise()
        return wavefront

    def _lower(self, plan) -> bool:
        plan.mul_amp(self._grid_transmission(plan.n, plan.dx))
        if self._aberrated_apertures():
            plan.add_opd(self._aberration(plan.n, plan.dx))
        if self.normalise:
            plan.normalise()
        return True


class CompoundAperture(CompositeAperture):
    """Product of its apertures (layers/apertures.py:1005-1069)."""

    def __init__(self, apertures, transformation=None, normalise: bool = False):
        super().__init__(apertures, transformation, normalise)
        if sum(isinstance(ap, AberratedAperture) for ap in self.apertures.values()) > 1:
            raise TypeError("CompoundAperture can only have a single AberratedAperture.")

    def _descs(self, transforms, group=None):
        tfs = self._transforms(transforms)
        key = id(self) if group is None else group
        out = []
        for ap in self.apertures.values():
            if isinstance(ap, MultiAperture):
                raise TypeError("CompoundAperture cannot contain a MultiAperture.")
            out.extend(ap._descs(tfs, key))
        return out


class MultiAperture(CompositeAperture):
    """Sum of its apertures (layers/apertures.py:1072-1117)."""

    def _descs(self, transforms, group=None):
        tfs = self._transforms(transforms)
        out = []
        for ap in self.apertures.values():
            out.extend(ap._descs(tfs))   # each member is a group of its own
        return out
