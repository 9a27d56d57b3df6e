"""Wavefront state and propagation (wavefronts.py:17-922) on the HIP path.

A ``Wavefront`` holds a complex64 device phasor -- (N, N), or (B, N, N) for a
wavefront vectorised over B wavelengths, the shape dLux obtains from
``eqx.filter_vmap`` -- and host-side FP64 physical parameters (wavelength,
pixel scale, centre), which are what the C ABI takes.  Every method returns a
new Wavefront.
"""

from __future__ import annotations

import numpy as np
import torch

from . import runtime
from .base import Base
from .coordinates import CoordSpec
from .psfs import PSF
from .utils import array_ops, fourier, units
from .utils.propagation import FFT as _FFT

__all__ = ["Wavefront"]


def _batched(x) -> bool:
    return np.ndim(x) > 0


class Wavefront(Base):
    """Complex field plus sampling; see wavefronts.py:17-122."""

    def __init__(self, wavelength, npixels: int, diameter=None, pixel_scale=None, center=None):
        if diameter is None and pixel_scale is None:
            raise ValueError("Provide one: diameter or pixel_scale.")
        if diameter is not None and pixel_scale is not None:
            raise ValueError(
                "Cannot specify both 'diameter' and 'pixel_scale' - they are "
                "interdependent (diameter = pixel_scale × npixels). Choose one: "
                "use 'diameter' for wavefront diameter, or 'pixel_scale' for "
                "wavefront sampling."
            )
        self.wavelength = np.asarray(wavelength, float)
        if diameter is not None:
            ps = np.asarray(diameter / npixels, float)
        else:
            ps = np.asarray(pixel_scale, float)
        if self.wavelength.ndim == 1 and ps.ndim == 0:
            ps = np.full(self.wavelength.shape, float(ps))
        self.pixel_scale = ps
        if center is not None:
            self.center = np.asarray(center, float)
            if self.center.shape != (1,):
                raise ValueError("center must have shape (1,).")
        else:
            self.center = np.zeros(1, float)
        shape = (npixels, npixels) if self.wavelength.ndim == 0 else (
            self.wavelength.size, npixels, npixels)
        self.phasor = torch.full(shape, 1.0 / npixels ** 2, dtype=torch.complex64,
                                 device=runtime.device())

    @classmethod
    def from_phasor(cls, phasor, wavelength, pixel_scale=None, diameter=None, center=None):
        """wavefronts.py:124-160."""
        ph = runtime.as_complex(phasor)
        wf = cls.__new__(cls)
        wf.wavelength = np.asarray(wavelength, float)
        n = ph.shape[-1]
        if diameter is None and pixel_scale is None:
            raise ValueError("Provide one: diameter or pixel_scale.")
        if diameter is not None and pixel_scale is not None:
            raise ValueError("Cannot specify both 'diameter' and 'pixel_scale'.")
        wf.pixel_scale = np.asarray(diameter / n if diameter is not None else pixel_scale, float)
        wf.center = np.zeros(1, float) if center is None else np.asarray(center, float)
        wf.phasor = ph
        return wf

    # ------------------------------------------------------------------ props
    @property
    def npixels(self) -> int:
        return self.phasor.shape[-1]

    @property
    def diameter(self):
        return self.npixels * self.pixel_scale

    @property
    def real(self):
        return self.phasor.real

    @property
    def imaginary(self):
        return self.phasor.imag

    @property
    def amplitude(self):
        return torch.abs(self.phasor)

    @property
    def phase(self):
        return torch.angle(self.phasor)

    @property
    def complex(self):
        return torch.stack([self.phasor.real, self.phasor.imag], dim=0)

    @property
    def polar(self):
        return torch.stack([self.amplitude, self.phase], dim=0)

    @property
    def psf(self) -> torch.Tensor:
        """|E|^2 (dlx_intensity)."""
        out = runtime.intensity(self.phasor)
        return out if self.phasor.ndim == 3 else out[0]

    def to_psf(self) -> PSF:
        return PSF(self.psf, self.pixel_scale)

    @property
    def wavenumber(self):
        return 2 * np.pi / self.wavelength

    @property
    def ndim(self) -> int:
        return self.pixel_scale.ndim

    @property
    def power(self):
        """sum |E|^2 (dlx_power); one value per instance when vectorised."""
        return runtime.power(self.phasor)

    @property
    def spec(self) -> CoordSpec:
        return CoordSpec(self.npixels, self.pixel_scale, self.center)

    @property
    def xs(self) -> np.ndarray:
        d = float(np.ravel(self.pixel_scale)[0])
        return CoordSpec(self.npixels, d, self.center).xs

    def coordinates(self, scale=1.0, polar: bool = False) -> torch.Tensor:
        xs = torch.as_tensor((self.xs * scale).astype(np.float32), device=self.phasor.device)
        X, Y = torch.meshgrid(xs, xs, indexing="xy")
        coords = torch.stack([X, Y])
        if polar:
            return torch.stack([torch.hypot(X, Y), torch.atan2(Y, X)])
        return coords

    # ------------------------------------------------------------ internals
    def _nb(self) -> int:
        return self.phasor.shape[0] if self.phasor.ndim == 3 else 1

    def _wl_list(self) -> np.ndarray:
        return np.broadcast_to(np.atleast_1d(self.wavelength), (self._nb(),))

    def _apply(self, amp=None, opd=None, phase=None, tilt=(0.0, 0.0), scale=1.0) -> "Wavefront":
        """phasor *= scale * amp * exp(i(k (opd + tilt) + phase)) via dlx_pupil_field."""
        n = self.npixels
        ph3 = self.phasor if self.phasor.ndim == 3 else self.phasor[None]
        d = float(np.ravel(self.pixel_scale)[0])
        if np.ndim(self.pixel_scale) and not np.allclose(self.pixel_scale, d, rtol=0, atol=0):
            outs = [self._select(i)._apply(amp, opd, phase, tilt, scale).phasor
                    for i in range(self._nb())]
            return self.set(phasor=torch.stack(outs))
        args = []
        for a in (amp, opd, phase):
            if a is None:
                args.append(None)
                continue
            t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))
            if t.is_complex():
                raise TypeError("amplitude / OPD / phase must be real")
            if t.ndim == 0:
                t = t.expand(n, n)
            if t.ndim == 3 and t.shape[0] == self._nb() and self._nb() > 1:
                outs = []
                for i in range(self._nb()):
                    sub = [None if x is None else (x[i] if getattr(x, "ndim", 0) == 3 else x)
                           for x in (amp, opd, phase)]
                    outs.append(self._select(i)._apply(*sub, tilt=tilt, scale=scale).phasor)
                return self.set(phasor=torch.stack(outs))
            if t.ndim == 3 and t.shape[0] == 1:
                t = t[0]
            if tuple(t.shape) != (n, n):
                raise ValueError(f"cannot broadcast shape {tuple(t.shape)} to {(n, n)}")
            args.append(t)
        out = runtime.pupil_field(n, self._wl_list(), d, center=float(self.center[0]), tilt=tilt,
                                  amp=args[0], opd=args[1], phase=args[2], scale=scale,
                                  field_in=ph3)
        return self.set(phasor=out if self.phasor.ndim == 3 else out[0])

    def _select(self, i: int) -> "Wavefront":
        return self.set(phasor=self.phasor[i],
                        wavelength=np.asarray(np.atleast_1d(self.wavelength)[i]),
                        pixel_scale=np.asarray(np.atleast_1d(self.pixel_scale)[i]))

    # --------------------------------------------------------------- methods
    def add_phase(self, phase) -> "Wavefront":
        """phasor * exp(1j * phase).  wavefronts.py:332-349."""
        if phase is None:
            return self
        return self._apply(phase=phase)

    def add_opd(self, opd) -> "Wavefront":
        """phasor * exp(1j * k * opd).  wavefronts.py:351-368."""
        if opd is None:
            return self
        return self._apply(opd=opd)

    def tilt(self, angles, unit: str = "rad") -> "Wavefront":
        """OPD = theta_x X + theta_y Y.  wavefronts.py:370-395."""
        angles = np.asarray(angles, dtype=float)
        if angles.shape != (2,):
            raise ValueError("angles must be a 1d array of shape (2,).")
        f = units.unit_factor_to_rad(unit)
        return self._apply(tilt=(angles[0] * f, angles[1] * f))

    def normalise(self, mode: str = "power", value: float = 1.0) -> "Wavefront":
        """wavefronts.py:397-424."""
        if mode == "power":
            scale = torch.sqrt(value / runtime.power(self.phasor))
        elif mode == "peak":
            peak = runtime.intensity(self.phasor).flatten(-2).amax(-1)
            scale = torch.sqrt(value / peak)
        else:
            raise ValueError("mode must be 'power' or 'peak'")
        if self.phasor.ndim == 3:
            scale = scale.reshape(-1, 1, 1)
        return self.set(phasor=self.phasor * scale)

    def flip(self, axis) -> "Wavefront":
        axes = (axis,) if isinstance(axis, int) else tuple(axis)
        off = self.phasor.ndim - 2
        return self.set(phasor=torch.flip(self.phasor, [a + off if a >= 0 else a for a in axes]))

    def resize(self, npixels: int) -> "Wavefront":
        """Zero pad / crop, centre preserving.  wavefronts.py:568-582."""
        return self.set(phasor=array_ops.resize(self.phasor, npixels, 0.0))

    def set_spec(self, spec: CoordSpec) -> "Wavefront":
        ps = None if spec.d is None else np.asarray(spec.d, float)
        c = None if spec.c is None else np.asarray(spec.c, float)
        return self.set(pixel_scale=ps, center=c)

    def propagate(self, npixels: int, pixel_scale, focal_length=None,
                  inverse: bool = False) -> "Wavefront":
        """MFT to an (npixels, npixels) plane.  wavefronts.py:729-772."""
        ph = runtime.mft(self.phasor, self._wl_list(),
                         np.broadcast_to(np.atleast_1d(self.pixel_scale), (self._nb(),)),
                         int(npixels), pixel_scale, focal_length=focal_length, inverse=inverse)
        ps = np.array(pixel_scale, float)
        if self.ndim and ps.ndim == 0:
            ps = np.full(self.pixel_scale.shape, float(ps))
        return self.set(phasor=ph, pixel_scale=ps)

    def propagate_MFT(self, spec_out: CoordSpec, focal_length=None, inverse=None) -> "Wavefront":
        """wavefronts.py:774-805."""
        return self.propagate(spec_out.n, spec_out.d, focal_length, bool(inverse))

    def propagate_FFT(self, pad: int = 2, focal_length=None, spec_out: CoordSpec = None,
                      inverse: bool = False) -> "Wavefront":
        """Padded centred FFT with optional output re-centring.  wavefronts.py:654-727."""
        n_in = self.npixels
        d_in = float(np.ravel(self.pixel_scale)[0])
        n_out = n_in * pad
        if spec_out is not None:
            if spec_out.d is not None:
                raise ValueError("Output spec cannot specify d; FFT output d is fixed.")
            if spec_out.n is not None:
                raise ValueError(
                    "Output spec cannot specify n; FFT output n is determined by the ",
                    "pad parameter.",
                )
        # every wavelength's transform in one batched call; the ramps (when
        # re-centring) differ per wavelength and multiply by broadcasting
        wls = np.asarray(self._wl_list(), float)
        vec = self.phasor.ndim == 3
        specs = [fourier.fft_spec(n_out, d_in, float(wl), focal_length) for wl in wls]
        ph = self.phasor
        out_ramp = None
        if spec_out is not None:
            xs_in = CoordSpec(n_in, d_in, self.center).xs
            rin, rout = [], []
            for wl, (d_fft, c_fft) in zip(wls, specs):
                rin.append(fourier.fft_phase_ramp_1d(xs_in, float(wl), c_fft - spec_out.c,
                                                     focal_length, inverse))
                so = spec_out.set(n=n_out, d=d_fft)
                shift = fourier.fft_spec(so.n, so.d, float(wl), focal_length)[1]
                rout.append(fourier.fft_phase_ramp_1d(so.xs, float(wl), shift, focal_length,
                                                      inverse))
            rin, rout = torch.stack(rin), torch.stack(rout)   # (B, n), (B, n_out)
            in_ramp = rin[:, None, :] * rin[:, :, None]
            out_ramp = rout[:, None, :] * rout[:, :, None]
            if not vec:
                in_ramp, out_ramp = in_ramp[0], out_ramp[0]
            ph = ph * in_ramp
            c_out = np.atleast_1d(spec_out.c).astype(float)
        else:
            c_out = np.atleast_1d(np.asarray(specs[0][1], float))
        out, ps = _FFT(ph, wls if vec else float(wls[0]), d_in, focal_length=focal_length,
                       pad=pad, inverse=inverse)
        if out_ramp is not None:
            out = out * out_ramp
        return self.set(phasor=out, pixel_scale=np.asarray(ps, float), center=c_out)

    # --- resampling (wavefronts.py:442-566) ---------------------------------
    def _fields(self, complex: bool):
        """(2, ...) float planes: real and imaginary, or amplitude and phase."""
        ph = self.phasor
        if complex:
            return torch.stack([ph.real, ph.imag]).float()
        return torch.stack([torch.abs(ph), torch.angle(ph)]).float()

    def _from_fields(self, f, complex: bool):
        if complex:
            return torch.complex(f[0], f[1])
        return f[0] * torch.exp(1j * f[1].to(torch.float32))

    def scale_to(self, npixels: int, pixel_scale, complex: bool = True) -> "Wavefront":
        """Resampled to npixels at pixel_scale (wavefronts.py:442-483)."""
        from .utils.interpolation import scale
        pixel_scale = np.asarray(pixel_scale, float)
        ratio = float(np.ravel(pixel_scale)[0]) / float(np.ravel(self.pixel_scale)[0])
        f = scale(self._fields(complex), int(npixels), ratio)
        return self.set(phasor=self._from_fields(f, complex), pixel_scale=pixel_scale)

    def interpolate(self, transformation, method: str = "linear", complex: bool = True,
                    fill: float = 0.0) -> "Wavefront":
        """Sampled at transformation(self.coordinates()) (wavefronts.py:485-524)."""
        from .coordinates import BaseCoordTransform
        if not isinstance(transformation, BaseCoordTransform):
            raise TypeError("transformation must be a BaseCoordTransform.")
        from .coordinates import CoordTransform, _chain_coords, _untracked
        _untracked(transformation, "Wavefront.interpolate")
        d = float(np.ravel(self.pixel_scale)[0])
        c = float(np.ravel(self.center)[0])
        if isinstance(transformation, CoordTransform):
            f = runtime.interp_transform(self._fields(complex), d, c, self.npixels, d, c,
                                         [transformation._struct()], method, fill)
        else:
            q = _chain_coords((transformation,), self.npixels, d, c, self.phasor.device)
            f = runtime.interp_points(self._fields(complex), d, c, q[0], q[1], method, fill)
        return self.set(phasor=self._from_fields(f, complex))

    def rotate(self, angle, method: str = "linear", complex: bool = True) -> "Wavefront":
        """Rotated by `angle` radians about the grid centre (wavefronts.py:526-566)."""
        from .utils.interpolation import rotate
        f = rotate(self._fields(complex), angle, method)
        return self.set(phasor=self._from_fields(f, complex))

    # placeholders, as in the reference (wavefronts.py:814-832)
    def propagate_fresnel(self):
        """LCT-based MFT Fresnel propagation (not implemented in the reference either)."""
        raise NotImplementedError()

    def propagate_fresnel_fft(self):
        """LCT-based FFT Fresnel propagation (not implemented in the reference either)."""
        raise NotImplementedError()

    def propagate_fraunhofer(self):
        """Fraunhofer propagation via MFT (not implemented in the reference either)."""
        raise NotImplementedError()

    def propagate_fraunhofer_fft(self):
        """Fraunhofer propagation via FFT (not implemented in the reference either)."""
        raise NotImplementedError()

    def propagate_ASM(self, distance: float, npad: int) -> "Wavefront":
        """Angular-spectrum propagation to `distance`, padded to npad (no crop)."""
        d = float(np.ravel(self.pixel_scale)[0])
        return self.set(phasor=runtime.asm_prop(self.phasor, self._wl_list(), d,
                                                float(distance), int(npad)))

    # ----------------------------------------------------------- arithmetic
    def _op(self, other, op: str) -> "Wavefront":
        if other is None:
            return self
        if isinstance(other, Wavefront):
            other = other.phasor
        if not isinstance(other, (torch.Tensor, np.ndarray, float, int, complex)):
            raise TypeError(
                f"Unsupported type for {op}: {type(other)}. Must be an array, "
                "Wavefront, or None."
            )
        real = (isinstance(other, (float, int)) or
                (not isinstance(other, complex) and not torch.is_tensor(other)
                 and not np.iscomplexobj(other)) or
                (torch.is_tensor(other) and not other.is_complex()))
        if op == "multiply" and real and np.ndim(other) >= 2:
            return self._apply(amp=other)
        o = other if torch.is_tensor(other) else torch.as_tensor(np.asarray(other))
        o = o.to(self.phasor.device)
        if op == "add":
            return self.set(phasor=self.phasor + o)
        if op == "subtract":
            return self.set(phasor=self.phasor - o)
        if op == "multiply":
            return self.set(phasor=self.phasor * o)
        if op == "divide":
            return self.set(phasor=self.phasor / o)
        raise ValueError(f"Unsupported operation '{op}'.")

    def __add__(self, other):
        return self._op(other, "add")

    def __sub__(self, other):
        return self._op(other, "subtract")

    def __mul__(self, other):
        return self._op(other, "multiply")

    def __truediv__(self, other):
        return self._op(other, "divide")

    __iadd__ = __add__
    __isub__ = __sub__
    __imul__ = __mul__
    __itruediv__ = __truediv__
