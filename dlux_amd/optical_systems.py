"""Optical systems (optical_systems.py:30-775) with fused polychromatic propagation.

``OpticalSystem.propagate`` keeps the reference's signature, validation and
return modes.  When every layer before the focal-plane MFT only multiplies
the field by a lambda-independent amplitude, adds a lambda-independent OPD or
phase, tilts it or normalises it (TransmissiveLayer, AberratedLayer,
BasisLayer, Optic, BasisOptic, Tilt, Normalise), the chain collapses to

    E_l = s * A * exp(i (2 pi / l) (OPD + tilt . r) + i Phi),

because |exp(i.)| = 1 makes every normalisation factor lambda-independent.
The system then computes sum_l w_l |MFT_l(E_l)|^2 in one dlx_psf_poly call
(pupil synthesis, per-wavelength DFT matrices, two MFMA contractions and the
weighted |E|^2 reduction).  Any other layer makes it fall back to applying the
layers one by one on a wavelength-vectorised Wavefront (still on the GPU).
"""

from __future__ import annotations

from collections import OrderedDict

import numpy as np
import torch

from . import runtime
from .base import Base
from .layers.optical_layers import OpticalLayer
from .layers.propagators import MFT as MFTLayer
from .psfs import PSF
from .utils import units
from .wavefronts import Wavefront

__all__ = ["BaseOpticalSystem", "OpticalSystem", "ParametricOpticalSystem",
           "LayeredOpticalSystem", "AngularOpticalSystem", "CartesianOpticalSystem",
           "list2dictionary"]


def list2dictionary(list_in, ordered: bool = True, allowed_types=()):
    """(name, item) list -> dict with de-duplicated names (utils/helpers.py:93-160)."""
    names, items, repeats = [], [], []
    for item in list_in:
        if isinstance(item, tuple):
            name, item = item
        else:
            name = item.__class__.__name__
        if allowed_types != () and not isinstance(item, allowed_types):
            raise TypeError(f"Item {name} is not an allowed type, got {type(item)}")
        if name in names:
            repeats.append(name)
        names.append(name)
        items.append(item)
    for r in set(repeats):
        idx = 0
        for j in range(len(names)):
            if names[j] == r:
                names[j] = f"{names[j]}_{idx}"
                idx += 1
    out = OrderedDict() if ordered else {}
    for name, item in zip(names, items):
        if " " in name:
            raise ValueError(f"Names cannot contain spaces, got {name}")
        out[name] = item
    return out


def _tensor_key(x):
    """Identity + in-place version of a tensor (or the value itself)."""
    if isinstance(x, torch.Tensor):
        return (id(x), x._version)
    return (id(x),)


class PupilPlan:
    """Lambda-independent pupil E = s A exp(i(k (OPD + tilt.r) + Phi)), on the device.

    Layers record their effect through mul_amp / normalise (the amplitude
    pipeline, which the power normalisation depends on) and add_opd /
    add_phase / add_tilt.  The amplitude pipeline depends only on
    transmission arrays, so ``finalise`` caches its result across calls;
    OPD and phase terms (e.g. a BasisLayer's coefficients) are evaluated on
    every call.
    """

    def __init__(self, n: int, dx: float, dev):
        self.n, self.dx, self.dev = n, dx, dev
        self.amp_ops = []
        self.opd_terms = []
        self.phase_terms = []
        self.tilt = np.zeros(2)
        self.amp = None
        self.opd = None
        self.phase = None
        self.scale = 1.0 / n ** 2          # reference initial phasor ones/N^2

    def _t(self, x):
        t = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, np.float32))
        t = t.to(device=self.dev, dtype=torch.float32)
        if t.ndim == 0:
            t = t.expand(self.n, self.n)
        if tuple(t.shape) != (self.n, self.n):
            raise ValueError(f"cannot broadcast shape {tuple(t.shape)} to {(self.n, self.n)}")
        return t

    def mul_amp(self, t):
        if t is not None:
            self.amp_ops.append(("mul", t))

    def normalise(self, mode: str = "power", value: float = 1.0):
        self.amp_ops.append(("norm", mode, float(value)))

    def add_opd(self, o):
        if o is not None:
            self.opd_terms.append(o)

    def add_phase(self, p):
        if p is not None:
            self.phase_terms.append(p)

    def add_tilt(self, a):
        self.tilt = self.tilt + np.asarray(a, float)

    def _amp_key(self):
        return tuple(("mul", _tensor_key(op[1])) if op[0] == "mul" else op
                     for op in self.amp_ops)

    def finalise(self, amp_cache: dict) -> "PupilPlan":
        # The cache entry keeps its key tensors alive, so their ids stay unique;
        # copies of a system (Base.set) share the cache and hit it while their
        # transmission arrays are the same, unmodified tensors.
        key = self._amp_key()
        # an amplitude that carries an autograd graph is rebuilt every call
        tracked = torch.is_grad_enabled() and any(
            op[0] == "mul" and isinstance(op[1], torch.Tensor) and op[1].requires_grad
            for op in self.amp_ops)
        if not tracked and key in amp_cache:
            self.amp, self.scale, _ = amp_cache[key]
        else:
            amp, scale = None, 1.0 / self.n ** 2
            for op in self.amp_ops:
                if op[0] == "mul":
                    t = self._t(op[1])
                    amp = t if amp is None else amp * t
                else:
                    # |E| = s |A| at every wavelength: fold the normalisation into
                    # the amplitude on the device (no host synchronisation).
                    a = amp if amp is not None else torch.ones(
                        (self.n, self.n), dtype=torch.float32, device=self.dev)
                    e2 = (a.double() * scale) ** 2
                    ref = torch.sum(e2) if op[1] == "power" else torch.max(e2)
                    amp = (a.double() * (scale * torch.sqrt(op[2] / ref))).float()
                    scale = 1.0
            if not tracked:
                amp_cache.clear()
                refs = [op[1] for op in self.amp_ops if op[0] == "mul"]
                amp_cache[key] = (amp, scale, refs)
            self.amp, self.scale = amp, scale
        for attr, terms in (("opd", self.opd_terms), ("phase", self.phase_terms)):
            acc = None
            for t in terms:
                t = self._t(t)
                acc = t if acc is None else acc + t
            setattr(self, attr, acc)
        return self


class BaseOpticalSystem(Base):
    """Abstract optical system (optical_systems.py:30-135)."""

    def propagate_mono(self, wavelength, offset=None, return_wf: bool = False):
        raise NotImplementedError

    def propagate(self, wavelengths, offset=None, weights=None, return_wf=False,
                  return_psf=False):
        raise NotImplementedError

    def model(self, source, return_wf: bool = False, return_psf: bool = False):
        raise NotImplementedError


class OpticalSystem(BaseOpticalSystem):
    """propagate / model shared by all systems (optical_systems.py:138-231)."""

    def propagate(self, wavelengths, offset=None, weights=None, return_wf: bool = False,
                  return_psf: bool = False):
        if return_wf and return_psf:
            raise ValueError(
                "Cannot return both Wavefront and PSF objects. Choose one: "
                "return_wf=True for Wavefront, or return_psf=True for PSF."
            )
        wavelengths = np.atleast_1d(np.asarray(wavelengths, float))
        if weights is None:
            weights = np.ones_like(wavelengths) / len(wavelengths)
        else:
            weights = np.atleast_1d(np.asarray(weights, float))
        if weights.shape != wavelengths.shape:
            raise ValueError(
                f"Wavelength and weight shape mismatch: "
                f"wavelengths {wavelengths.shape} vs weights {weights.shape}. "
                f"Must have same length and dimensions."
            )
        offset = np.zeros(2) if offset is None else np.asarray(offset, float)
        if offset.shape != (2,):
            raise ValueError(
                f"offset must be [x, y] array of shape (2,), "
                f"got shape {offset.shape}. "
                "Pass offset as [on_axis_x, on_axis_y] angles in radians."
            )

        if not return_wf:
            fused = self._fused_psf(wavelengths, weights, offset)
            if fused is not None:
                psf, p_out = fused
                return PSF(psf, p_out) if return_psf else psf

        # Generic path: the whole wavelength batch as one vectorised Wavefront.
        wf = self.propagate_mono(wavelengths, offset, return_wf=True)
        sw = torch.as_tensor(np.sqrt(weights).astype(np.float32), device=wf.phasor.device)
        wf = wf.set(phasor=wf.phasor * sw.reshape(-1, 1, 1))
        if return_wf:
            return wf
        psf = runtime.intensity(wf.phasor, sum_batch=True)
        if return_psf:
            return PSF(psf, np.mean(wf.pixel_scale))
        return psf

    def _fused_psf(self, wavelengths, weights, offset, detector_ops=None):
        return None   # overridden where the layers lower to one fused call

    def model(self, source, return_wf: bool = False, return_psf: bool = False):
        return source.model(self, return_wf, return_psf)


class ParametricOpticalSystem(OpticalSystem):
    """psf_npixels / oversample / psf_pixel_scale (optical_systems.py:234-292)."""

    def __init__(self, psf_npixels: int, psf_pixel_scale, oversample: int = 1, **kwargs):
        self.psf_npixels = int(psf_npixels)
        self.oversample = int(oversample)
        self.psf_pixel_scale = np.asarray(psf_pixel_scale, float)
        super().__init__(**kwargs)

    @property
    def fov(self):
        return self.psf_npixels * self.psf_pixel_scale


class LayeredOpticalSystem(OpticalSystem):
    """A chain of OpticalLayers (optical_systems.py:298-507)."""

    def __init__(self, wf_npixels: int, diameter, layers, **kwargs):
        self.wf_npixels = int(wf_npixels)
        self.diameter = np.asarray(diameter, float)
        self.layers = list2dictionary(layers, True, OpticalLayer)
        if kwargs:
            super().__init__(**kwargs)

    def __getattr__(self, key):
        layers = self.__dict__.get("layers")
        if layers is None:
            raise AttributeError(key)
        if key in layers:
            return layers[key]
        for layer in layers.values():
            if hasattr(layer, key):
                return getattr(layer, key)
        raise AttributeError(
            f"{type(self).__name__} has no attribute {key!r} (layers: {list(layers)})")

    def initialise_wavefront(self, wavelength, offset=None) -> Wavefront:
        """Wavefront(lambda, N, D) tilted by the source offset (optical_systems.py:363-389)."""
        if offset is None:
            offset = np.zeros(2)
        wf = Wavefront(wavelength, self.wf_npixels, float(self.diameter))
        return wf.tilt(offset)

    def propagate_mono(self, wavelength, offset=None, return_wf: bool = False):
        wf = self.initialise_wavefront(wavelength, offset)
        for layer in list(self.layers.values()):
            wf = layer(wf)
        return wf if return_wf else wf.psf

    def debug_propagate_mono(self, wavelength, offset=None):
        outputs = {}
        wf = self.initialise_wavefront(wavelength, offset)
        outputs["initial_wavefront"] = wf
        for name, layer in self.layers.items():
            wf = layer(wf)
            outputs[name] = wf
        return wf, outputs

    def insert_layer(self, layer, index: int):
        items = list(self.layers.items())
        items.insert(index, layer if isinstance(layer, tuple) else
                     (layer.__class__.__name__, layer))
        return self.set("layers", list2dictionary(items, True, OpticalLayer))

    def remove_layer(self, key: str):
        layers = OrderedDict(self.layers)
        layers.pop(key)
        return self.set("layers", layers)

    # ---- fused lowering -------------------------------------------------
    def _lower(self, layers, offset):
        """(PupilPlan, total tilt) for `layers`, or None if one is not lowerable."""
        n = self.wf_npixels
        plan = PupilPlan(n, float(self.diameter) / n, runtime.device())
        for layer in layers:
            if not layer._lower(plan):
                return None
        plan.finalise(self.__dict__.setdefault("_plan_cache", {}))
        return plan, plan.tilt + offset

    def _fused_psf(self, wavelengths, weights, offset, detector_ops=None):
        """(PSF, p_out) from one dlx_psf_poly call, or None when the layers do not
        lower; with `detector_ops`, the detector image from one
        dlx_psf_poly_detector call instead."""
        layers = list(self.layers.values())
        if not layers or not isinstance(layers[-1], MFTLayer):
            return None
        prop = layers[-1]
        lowered = self._lower(layers[:-1], offset)
        if lowered is None:
            return None
        plan, tilt = lowered
        kw = dict(amp=plan.amp, opd=plan.opd, phase=plan.phase, scale=plan.scale, tilt=tilt,
                  focal_length=prop.focal_length, inverse=prop.inverse)
        if detector_ops is not None:
            img = runtime.psf_poly_detector(plan.n, prop.npixels, wavelengths, weights, plan.dx,
                                            float(prop.pixel_scale), detector_ops, **kw)
            return img, float(prop.pixel_scale)
        psf = runtime.psf_poly(plan.n, prop.npixels, wavelengths, weights, plan.dx,
                               float(prop.pixel_scale), **kw)
        return psf, float(prop.pixel_scale)


class ParametricLayeredOpticalSystem(ParametricOpticalSystem, LayeredOpticalSystem):
    """Layered system with a focal plane (optical_systems.py:510-594)."""

    def __init__(self, wf_npixels, diameter, layers, psf_npixels, psf_pixel_scale,
                 oversample=1):
        LayeredOpticalSystem.__init__(self, wf_npixels, diameter, layers)
        self.psf_npixels = int(psf_npixels)
        self.oversample = int(oversample)
        self.psf_pixel_scale = np.asarray(psf_pixel_scale, float)

    def _focus_params(self):  # -> (npixels, pixel_scale, focal_length)
        raise NotImplementedError

    def to_focus(self, wavefront: Wavefront) -> Wavefront:
        n, p, f = self._focus_params()
        return wavefront.propagate(n, p, f)

    def propagate_mono(self, wavelength, offset=None, return_wf: bool = False):
        wf = LayeredOpticalSystem.propagate_mono(self, wavelength, offset, return_wf=True)
        wf = self.to_focus(wf)
        return wf if return_wf else wf.psf

    def debug_propagate_mono(self, wavelength, offset=None):
        wf, outputs = LayeredOpticalSystem.debug_propagate_mono(self, wavelength, offset)
        wf = self.to_focus(wf)
        outputs["final_wavefront"] = wf
        return wf, outputs

    def _fused_psf(self, wavelengths, weights, offset, detector_ops=None):
        lowered = self._lower(list(self.layers.values()), offset)
        if lowered is None:
            return None
        plan, tilt = lowered
        n_out, p_out, f = self._focus_params()
        kw = dict(amp=plan.amp, opd=plan.opd, phase=plan.phase, scale=plan.scale, tilt=tilt,
                  focal_length=f)
        if detector_ops is not None:
            img = runtime.psf_poly_detector(plan.n, n_out, wavelengths, weights, plan.dx, p_out,
                                            detector_ops, **kw)
            return img, p_out
        psf = runtime.psf_poly(plan.n, n_out, wavelengths, weights, plan.dx, p_out, **kw)
        return psf, p_out


    def _sources_psf(self, terms, n_outputs: int = 1):
        """Point-source images as ONE uneven batched MFT call over every (star,
        wavelength) instance.  `terms` lists (wavelengths (L,), positions (S, 2)
        in radians, weights (S, L), output index); image o is
        sum over the terms sent to o of sum_l w[s, l] |MFT_l(P_l)|^2, each
        star's position entering as an MFT shift of position / p_out output
        pixels (equal in intensity to the pupil tilt of sources.py:316-327).
        The pupil fields are synthesised once per distinct wavelength.
        Returns (list of n_outputs images, p_out), or None when the layers do
        not lower, gradients are being recorded, or the FP32 form is selected
        (the batched launches run the split-BF16x3 contractions)."""
        if runtime.get_precision() != "bf16x3":
            return None
        lowered = self._lower(list(self.layers.values()), np.zeros(2))
        if lowered is None:
            return None
        plan, tilt = lowered
        if torch.is_grad_enabled() and any(
                t is not None and isinstance(t, torch.Tensor) and t.requires_grad
                for t in (plan.amp, plan.opd, plan.phase)):
            return None
        n_out, p_out, f = self._focus_params()
        if f is not None:
            return None
        wls, pos, wts, outs = [], [], [], []
        for wavelengths, positions, weights, out_index in terms:
            wl = np.atleast_1d(np.asarray(wavelengths, float))
            p = np.asarray(positions, float).reshape(-1, 2)
            w = np.asarray(weights, float).reshape(len(p), len(wl))
            out_index = int(out_index)
            if not 0 <= out_index < n_outputs:
                raise ValueError(f"output index {out_index} outside [0, {n_outputs})")
            wls.append(np.tile(wl, len(p)))
            pos.append(np.repeat(p, len(wl), axis=0))
            wts.append(w.reshape(-1))
            outs.append(np.full(w.size, out_index, np.int64))
        if not wls:
            return None
        wl_all = np.concatenate(wls)
        unique_wl, slice_index = np.unique(wl_all, return_inverse=True)
        fields = runtime.pupil_field(plan.n, unique_wl, plan.dx, tilt=tilt, amp=plan.amp,
                                     opd=plan.opd, phase=plan.phase, scale=plan.scale)
        pos_all = np.concatenate(pos)
        batch = runtime.MftBatch.from_arrays(
            [fields], np.zeros(len(wl_all), np.int64), wl_all, plan.dx, int(n_out),
            float(p_out), np.concatenate(wts), np.concatenate(outs),
            shift_x=pos_all[:, 0] / p_out, shift_y=pos_all[:, 1] / p_out,
            slice_index=slice_index, n_outputs=n_outputs)
        images = batch()
        dev = fields.device
        images = [torch.zeros((int(n_out), int(n_out)), device=dev) if im is None else im
                  for im in images]
        return images, p_out


class AngularOpticalSystem(ParametricLayeredOpticalSystem):
    """Focal plane sampled in arcseconds (optical_systems.py:597-680)."""

    def _focus_params(self):
        true_pixel_scale = self.psf_pixel_scale / self.oversample
        return (self.psf_npixels * self.oversample, float(units.arcsec2rad(true_pixel_scale)),
                None)


class CartesianOpticalSystem(ParametricLayeredOpticalSystem):
    """Focal plane sampled in microns (optical_systems.py:683-775).

    Like the reference, to_focus passes p = 1e-6 * psf_pixel_scale / oversample
    to the MFT without the focal length (optical_systems.py:772-775).
    """

    def __init__(self, wf_npixels, diameter, layers, focal_length, psf_npixels,
                 psf_pixel_scale, oversample=1):
        self.focal_length = np.asarray(focal_length, float)
        super().__init__(wf_npixels, diameter, layers, psf_npixels, psf_pixel_scale,
                         oversample)

    def _focus_params(self):
        true_pixel_scale = self.psf_pixel_scale / self.oversample
        return self.psf_npixels * self.oversample, float(1e-6 * true_pixel_scale), None
