"""Device plumbing and thin typed wrappers over the C ABI.

torch provides device memory, the caching allocator and the current HIP
stream; every computation is a call into ``libdlux_amd.so``.  Inputs are
converted to contiguous device tensors (complex64 / float32) before the call.
"""

from __future__ import annotations

import contextlib
import ctypes
import os

import numpy as np
import torch

from . import _lib

_PRECISIONS = {"fp32": _lib.DLX_PREC_FP32, "bf16x3": _lib.DLX_PREC_BF16X3}
_state = {"precision": os.environ.get("DLUX_AMD_PRECISION", "fp32")}


def set_precision(name: str) -> None:
    """Select the MFT arithmetic form: 'fp32' (default) or 'bf16x3'."""
    if name not in _PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {name!r}")
    _state["precision"] = name


def get_precision() -> str:
    return _state["precision"]


@contextlib.contextmanager
def precision(name: str):
    """Context manager that temporarily selects the MFT arithmetic form."""
    old = get_precision()
    set_precision(name)
    try:
        yield
    finally:
        set_precision(old)


def _prec_code(prec: str | None) -> int:
    name = prec or get_precision()
    if name not in _PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {name!r}")
    return _PRECISIONS[name]


def device() -> torch.device:
    """The GPU the library runs on; raises if there is none (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise RuntimeError("dlux_amd requires a ROCm GPU (gfx950); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def as_complex(x, dev=None) -> torch.Tensor:
    """Contiguous complex64 device tensor."""
    dev = dev or device()
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev)
        if not t.is_complex():
            t = t.to(torch.float32).to(torch.complex64)
        return t.to(torch.complex64).contiguous()
    return torch.as_tensor(np.asarray(x, np.complex64), device=dev).contiguous()


def as_float(x, dev=None) -> torch.Tensor:
    """Contiguous float32 device tensor."""
    dev = dev or device()
    if isinstance(x, torch.Tensor):
        if x.is_complex():
            raise TypeError("expected a real array, got a complex one")
        return x.to(device=dev, dtype=torch.float32).contiguous()
    return torch.as_tensor(np.asarray(x, np.float32), device=dev).contiguous()


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def workspace(nbytes: int, dev=None) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev or device())


def geom(n_in: int, n_out: int, focal_length=None, inverse=False, focal_shift=0.0):
    g = _lib.MftGeom()
    g.n_in = int(n_in)
    g.n_out = int(n_out)
    g.inverse = int(bool(inverse))
    g.has_focal = 0 if focal_length is None else 1
    g.focal_length = 0.0 if focal_length is None else float(focal_length)
    g.focal_shift = float(focal_shift)
    return g


def _bcast(v, n: int) -> np.ndarray:
    """A length-n float64 vector from a scalar or n values."""
    a = np.asarray(v, np.float64).reshape(-1)
    if a.size == 1:
        a = np.repeat(a, n)
    if a.size != n:
        raise ValueError(f"expected a scalar or {n} values, got shape {np.shape(v)}")
    return a


# ---------------------------------------------------------------------------
# entry-point wrappers
# ---------------------------------------------------------------------------


def _untracked_inputs(name: str, *tensors) -> None:
    """Raise when an input of a non-differentiable entry point requires grad
    while gradients are recorded (instead of returning a result without a
    gradient path)."""
    if not torch.is_grad_enabled():
        return
    for t in tensors:
        if isinstance(t, torch.Tensor) and t.requires_grad:
            raise NotImplementedError(
                f"{name} is not differentiable; gradients flow through the fused paths "
                "(psf_poly, soft apertures, basis evaluation, detector, resampling, intensity)")


def transfer_matrix_T(wavelengths, n_in, pixel_scales_in, n_out, pixel_scales_out,
                      shifts=0.0, focal_length=None, focal_shift=0.0, inverse=False):
    """(B, N_out, N_in) complex64: transposed transfer matrices for B instances."""
    wl = np.atleast_1d(np.asarray(wavelengths, np.float64))
    b = wl.size
    lib = _lib.load()
    out = torch.empty((b, int(n_out), int(n_in)), dtype=torch.complex64, device=device())
    g = geom(n_in, n_out, focal_length, inverse, focal_shift)
    st = lib.dlx_transfer_matrix(g, b, _lib.dptr(wl), _lib.dptr(_bcast(pixel_scales_in, b)),
                                 _lib.dptr(_bcast(pixel_scales_out, b)),
                                 _lib.dptr(_bcast(shifts, b)), out.data_ptr(), stream_handle())
    _lib.check(st, "dlx_transfer_matrix")
    return out


def mft(fields: torch.Tensor, wavelengths, pixel_scales_in, n_out: int, pixel_scales_out,
        focal_length=None, shift_x=0.0, shift_y=0.0, inverse=False, field_scale=None,
        precision: str | None = None) -> torch.Tensor:
    """Batched MFT of (B, N, N) complex64 device fields -> (B, n_out, n_out)."""
    _untracked_inputs("mft", fields)
    f = as_complex(fields)
    squeeze = f.ndim == 2
    if squeeze:
        f = f[None]
    b, n, n2 = f.shape
    if n != n2:
        raise ValueError(f"MFT input must be square, got {tuple(f.shape[-2:])}")
    lib = _lib.load()
    g = geom(n, n_out, focal_length, inverse)
    prec = _prec_code(precision)
    ws = workspace(lib.dlx_mft_workspace_size(g, b, prec), f.device)
    out = torch.empty((b, int(n_out), int(n_out)), dtype=torch.complex64, device=f.device)
    fs = None if field_scale is None else _lib.dptr(_bcast(field_scale, b))
    st = lib.dlx_mft(g, b, f.data_ptr(), out.data_ptr(), _lib.dptr(_bcast(wavelengths, b)),
                     _lib.dptr(_bcast(pixel_scales_in, b)),
                     _lib.dptr(_bcast(pixel_scales_out, b)), _lib.dptr(_bcast(shift_x, b)),
                     _lib.dptr(_bcast(shift_y, b)), fs, prec, ws.data_ptr(), ws.numel(),
                     stream_handle())
    _lib.check(st, "dlx_mft")
    return out[0] if squeeze else out


def pupil_field(n: int, wavelengths, pixel_scale: float, center=0.0, tilt=(0.0, 0.0),
                amp=None, opd=None, phase=None, scale=1.0, field_in=None) -> torch.Tensor:
    """(B, n, n) complex64: s * amp * exp(i(k_b (opd + tilt) + phase)) [* field_in]."""
    _untracked_inputs("pupil_field", amp, opd, phase, field_in)
    wl = np.atleast_1d(np.asarray(wavelengths, np.float64))
    b = wl.size
    dev = device()
    lib = _lib.load()
    amp_t = None if amp is None else as_float(amp, dev)
    opd_t = None if opd is None else as_float(opd, dev)
    ph_t = None if phase is None else as_float(phase, dev)
    for name, t in (("amp", amp_t), ("opd", opd_t), ("phase", ph_t)):
        if t is not None and tuple(t.shape) != (n, n):
            raise ValueError(f"{name} must have shape {(n, n)}, got {tuple(t.shape)}")
    fin = None
    if field_in is not None:
        fin = as_complex(field_in, dev)
        if fin.ndim == 2:
            fin = fin.expand(b, n, n).contiguous()
        if tuple(fin.shape) != (b, n, n):
            raise ValueError(f"field must have shape {(b, n, n)}, got {tuple(fin.shape)}")
    out = torch.empty((b, n, n), dtype=torch.complex64, device=dev)
    st = lib.dlx_pupil_field(int(n), b, _lib.dptr(wl), float(pixel_scale), float(center),
                             float(tilt[0]), float(tilt[1]), _ptr(amp_t), _ptr(opd_t),
                             _ptr(ph_t), float(scale), _ptr(fin), out.data_ptr(),
                             stream_handle())
    _lib.check(st, "dlx_pupil_field")
    return out


def _eval_basis_impl(bt: torch.Tensor, ct: torch.Tensor, base_t, nb: int, npix: int,
                     pix_shape) -> torch.Tensor:
    out = torch.empty(pix_shape, dtype=torch.float32, device=bt.device)
    st = _lib.load().dlx_eval_basis(nb, npix, 1, bt.data_ptr(), ct.reshape(-1).data_ptr(),
                                    _ptr(base_t), out.data_ptr(), stream_handle())
    _lib.check(st, "dlx_eval_basis")
    return out


def project_basis(basis, grad) -> torch.Tensor:
    """Transpose of eval_basis: out[..., j] = sum_p basis[j, p] grad[..., p].

    basis is (nbasis, *pix), grad is (*pix) or (nsets, *pix); the result is
    (nbasis,) or (nsets, nbasis), float32, summed in a fixed order.
    """
    dev = device()
    bt = as_float(basis, dev)
    nbasis = int(bt.shape[0])
    npix = int(np.prod(bt.shape[1:]))
    gt = as_float(grad, dev)
    if gt.numel() % max(npix, 1) != 0:
        raise ValueError(f"grad of shape {tuple(gt.shape)} does not match basis pixels {npix}")
    nsets = gt.numel() // npix if npix else 1
    lib = _lib.load()
    need = lib.dlx_project_basis_workspace_size(nbasis, npix, nsets)
    ws = workspace(need, dev)
    out = torch.empty((nsets, nbasis), dtype=torch.float32, device=dev)
    st = lib.dlx_project_basis(nbasis, npix, nsets, bt.data_ptr(), gt.data_ptr(),
                               out.data_ptr(), ws.data_ptr(), ws.numel(), stream_handle())
    _lib.check(st, "dlx_project_basis")
    return out[0] if gt.numel() == npix else out


class _EvalBasisFn(torch.autograd.Function):
    """eval_basis with the coefficient (and base) gradient on the device."""

    @staticmethod
    def forward(ctx, ct, bt, base_t, nb, npix, pix_shape):
        ctx.save_for_backward(bt)
        ctx.cshape = tuple(ct.shape)
        ctx.nb, ctx.npix = nb, npix
        return _eval_basis_impl(bt, ct, base_t, nb, npix, pix_shape)

    @staticmethod
    def backward(ctx, grad):
        (bt,) = ctx.saved_tensors
        if ctx.needs_input_grad[1]:
            raise NotImplementedError("eval_basis: gradient with respect to the basis is not "
                                      "implemented on the device path")
        g_c = None
        if ctx.needs_input_grad[0]:
            g_c = project_basis(bt.reshape(ctx.nb, ctx.npix), grad.reshape(-1))
            g_c = g_c.reshape(ctx.cshape)
        g_base = grad if ctx.needs_input_grad[2] else None
        return g_c, None, g_base, None, None, None


def eval_basis(basis, coeffs, base=None) -> torch.Tensor:
    """sum_j coeffs[..., j] basis[j] on the device; coeffs (nsets, nb) or (nb,).

    Differentiable with respect to coeffs and base (torch autograd); the
    backward is dlx_project_basis.
    """
    dev = device()
    bt = as_float(basis, dev)
    ct = as_float(coeffs, dev)
    lead = ct.shape
    nb = int(np.prod(lead)) if ct.ndim else 1
    if tuple(bt.shape[: ct.ndim]) != tuple(lead):
        raise ValueError(
            "The leading basis dimensions must match the coefficient shape, "
            f"received {tuple(bt.shape)} and {tuple(lead)}."
        )
    pix_shape = tuple(bt.shape[ct.ndim:])
    npix = int(np.prod(pix_shape))
    base_t = None if base is None else as_float(base, dev)
    if torch.is_grad_enabled() and any(
            t is not None and t.requires_grad for t in (ct, bt, base_t)):
        return _EvalBasisFn.apply(ct, bt, base_t, nb, npix, pix_shape)
    return _eval_basis_impl(bt, ct, base_t, nb, npix, pix_shape)


def power(fields: torch.Tensor) -> torch.Tensor:
    """Per-instance sum |E|^2 of (B, ...) complex64 fields -> (B,) float32."""
    _untracked_inputs("power", fields)
    f = as_complex(fields)
    b = f.shape[0] if f.ndim == 3 else 1
    npix = f.numel() // b
    out = torch.empty(b, dtype=torch.float32, device=f.device)
    st = _lib.load().dlx_power(npix, b, f.data_ptr(), out.data_ptr(), stream_handle())
    _lib.check(st, "dlx_power")
    return out if f.ndim == 3 else out[0]


def _intensity(f: torch.Tensor, sum_batch: bool) -> torch.Tensor:
    if f.ndim == 2:
        f = f[None]
    b = f.shape[0]
    npix = f[0].numel()
    f = f.contiguous()
    shape = tuple(f.shape[1:]) if sum_batch else tuple(f.shape)
    out = torch.empty(shape, dtype=torch.float32, device=f.device)
    st = _lib.load().dlx_intensity(npix, b, f.data_ptr(), int(sum_batch), out.data_ptr(),
                                   stream_handle())
    _lib.check(st, "dlx_intensity")
    return out


class _IntensityFn(torch.autograd.Function):
    """|E|^2 (summed over the leading axis or not); dL/dE = 2 E dL/dI (torch's
    convention for a real function of a complex input)."""

    @staticmethod
    def forward(ctx, f, sum_batch):
        ctx.save_for_backward(f)
        ctx.sum_batch = sum_batch
        return _intensity(f, sum_batch)

    @staticmethod
    def backward(ctx, g):
        (f,) = ctx.saved_tensors
        f3 = f if f.ndim == 3 else f[None]
        g3 = g[None] if ctx.sum_batch else g.reshape(f3.shape)
        return (2.0 * f3 * g3.to(f3.real.dtype)).reshape(f.shape), None


def intensity(fields: torch.Tensor, sum_batch: bool = False) -> torch.Tensor:
    """|E|^2 of complex64 fields (optionally summed over the leading axis).
    Differentiable with respect to the fields."""
    f = as_complex(fields)
    if isinstance(fields, torch.Tensor) and fields.requires_grad and torch.is_grad_enabled():
        return _IntensityFn.apply(f, bool(sum_batch))
    return _intensity(f, sum_batch)


def psf_poly_workspace(n_in, n_out, nwl, precision=None) -> int:
    g = geom(n_in, n_out)
    return int(_lib.load().dlx_psf_poly_workspace_size(g, int(nwl), _prec_code(precision)))


def _poly_args(n_in, n_out, wavelengths, weights, focal_length, inverse, precision):
    wl = np.atleast_1d(np.asarray(wavelengths, np.float64))
    w = np.atleast_1d(np.asarray(weights, np.float64))
    if w.shape != wl.shape:
        raise ValueError(f"Wavelength and weight shape mismatch: {wl.shape} vs {w.shape}")
    return wl, w, geom(n_in, n_out, focal_length, inverse), _prec_code(precision)


def _pupil_inputs(n_in, amp, opd, phase, dev):
    ts = []
    for name, x in (("amp", amp), ("opd", opd), ("phase", phase)):
        t = None if x is None else as_float(x, dev)
        if t is not None and tuple(t.shape) != (n_in, n_in):
            raise ValueError(f"{name} must have shape {(n_in, n_in)}, got {tuple(t.shape)}")
        ts.append(t)
    return ts


def _psf_poly_impl(p: dict, amp_t, opd_t, ph_t, out=None, ws=None) -> torch.Tensor:
    lib = _lib.load()
    dev = p["dev"]
    need = lib.dlx_psf_poly_workspace_size(p["g"], p["wl"].size, p["prec"])
    if ws is None or ws.numel() < need:
        ws = workspace(need, dev)
    if out is None:
        out = torch.empty((p["n_out"], p["n_out"]), dtype=torch.float32, device=dev)
    st = lib.dlx_psf_poly(p["g"], p["wl"].size, _lib.dptr(p["wl"]), _lib.dptr(p["w"]),
                          p["dx"], p["pout"], p["center"], p["tilt"][0], p["tilt"][1],
                          _ptr(amp_t), _ptr(opd_t), _ptr(ph_t), p["scale"], out.data_ptr(),
                          p["prec"], ws.data_ptr(), ws.numel(), stream_handle())
    _lib.check(st, "dlx_psf_poly")
    return out


def _psf_poly_vjp_impl(p: dict, amp_t, opd_t, ph_t, dpsf, need_opd=True, need_phase=False,
                      need_amp=False, ws=None):
    lib = _lib.load()
    dev = p["dev"]
    n = p["n_in"]
    need = lib.dlx_psf_poly_vjp_workspace_size(p["g"], p["wl"].size, p["prec"])
    if ws is None or ws.numel() < need:
        ws = workspace(need, dev)
    g_opd = torch.empty((n, n), dtype=torch.float32, device=dev) if need_opd else None
    g_ph = torch.empty((n, n), dtype=torch.float32, device=dev) if need_phase else None
    g_amp = torch.empty((n, n), dtype=torch.float32, device=dev) if need_amp else None
    d = as_float(dpsf, dev)
    if tuple(d.shape) != (p["n_out"], p["n_out"]):
        raise ValueError(f"dpsf must have shape {(p['n_out'], p['n_out'])}, "
                         f"got {tuple(d.shape)}")
    st = lib.dlx_psf_poly_vjp(p["g"], p["wl"].size, _lib.dptr(p["wl"]), _lib.dptr(p["w"]),
                              p["dx"], p["pout"], p["center"], p["tilt"][0], p["tilt"][1],
                              _ptr(amp_t), _ptr(opd_t), _ptr(ph_t), p["scale"], d.data_ptr(),
                              _ptr(g_opd), _ptr(g_ph), _ptr(g_amp), p["prec"], ws.data_ptr(),
                              ws.numel(), stream_handle())
    _lib.check(st, "dlx_psf_poly_vjp")
    if need_amp:
        return g_opd, g_ph, g_amp
    return g_opd, g_ph


class _PsfPolyFn(torch.autograd.Function):
    """Fused polychromatic PSF with the hand-written adjoint as its backward."""

    @staticmethod
    def forward(ctx, opd_t, ph_t, amp_t, p):
        ctx.save_for_backward(opd_t, ph_t, amp_t)
        ctx.p = p
        return _psf_poly_impl(p, amp_t, opd_t, ph_t)

    @staticmethod
    def backward(ctx, dpsf):
        opd_t, ph_t, amp_t = ctx.saved_tensors
        need_amp = bool(ctx.needs_input_grad[2])
        out = _psf_poly_vjp_impl(ctx.p, amp_t, opd_t, ph_t, dpsf.contiguous(),
                                 need_opd=ctx.needs_input_grad[0],
                                 need_phase=ctx.needs_input_grad[1], need_amp=need_amp)
        g_amp = None
        if need_amp:
            g_opd, g_ph, g_amp = out   # dL/d amp at fixed scale (P already carries it)
        else:
            g_opd, g_ph = out
        return g_opd, g_ph, g_amp, None


def psf_poly(n_in: int, n_out: int, wavelengths, weights, pixel_scale_in: float,
             pixel_scale_out: float, amp=None, opd=None, phase=None, scale=1.0,
             tilt=(0.0, 0.0), center=0.0, focal_length=None, inverse=False,
             precision: str | None = None, out: torch.Tensor | None = None,
             ws: torch.Tensor | None = None) -> torch.Tensor:
    """Fused polychromatic PSF sum_l w_l |MFT_l(P_l)|^2 -> (n_out, n_out) float32.

    Differentiable (torch autograd) with respect to opd, phase and amp: the
    backward is dlx_psf_poly_vjp (dL/d amp at the fixed scale, from the pupil
    phasor, so it is exact where amp == 0 too).
    """
    wl, w, g, prec = _poly_args(n_in, n_out, wavelengths, weights, focal_length, inverse,
                                precision)
    dev = device()
    amp_t, opd_t, ph_t = _pupil_inputs(n_in, amp, opd, phase, dev)
    p = dict(dev=dev, g=g, wl=wl, w=w, prec=prec, n_in=int(n_in), n_out=int(n_out),
             dx=float(pixel_scale_in), pout=float(pixel_scale_out), center=float(center),
             tilt=(float(tilt[0]), float(tilt[1])), scale=float(scale))
    if torch.is_grad_enabled() and any(t is not None and t.requires_grad
                                       for t in (amp_t, opd_t, ph_t)):
        if out is not None:
            raise ValueError("psf_poly: out= cannot be used when gradients are recorded")
        return _PsfPolyFn.apply(opd_t, ph_t, amp_t, p)
    return _psf_poly_impl(p, amp_t, opd_t, ph_t, out=out, ws=ws)


def psf_poly_vjp(n_in: int, n_out: int, wavelengths, weights, pixel_scale_in: float,
                 pixel_scale_out: float, dpsf, amp=None, opd=None, phase=None, scale=1.0,
                 tilt=(0.0, 0.0), center=0.0, focal_length=None, inverse=False,
                 precision: str | None = None, need_opd: bool = True,
                 need_phase: bool = False, need_amp: bool = False,
                 ws: torch.Tensor | None = None):
    """(dL/dOPD, dL/dphase) of the fused PSF for dL/dPSF = dpsf (either may be
    None), and dL/d amp (at the fixed scale) third when ``need_amp``.  ``ws``:
    an optional caller workspace (used when large enough)."""
    wl, w, g, prec = _poly_args(n_in, n_out, wavelengths, weights, focal_length, inverse,
                                precision)
    dev = device()
    amp_t, opd_t, ph_t = _pupil_inputs(n_in, amp, opd, phase, dev)
    p = dict(dev=dev, g=g, wl=wl, w=w, prec=prec, n_in=int(n_in), n_out=int(n_out),
             dx=float(pixel_scale_in), pout=float(pixel_scale_out), center=float(center),
             tilt=(float(tilt[0]), float(tilt[1])), scale=float(scale))
    return _psf_poly_vjp_impl(p, amp_t, opd_t, ph_t, dpsf, need_opd, need_phase, need_amp,
                              ws=ws)


class MftBatch:
    """A prepared uneven batch: the descriptor array, outputs and workspace are
    built once; each call runs dlx_mft_batched_psf on the current stream.

    ``instances`` as for mft_batched_psf.
    """

    def __init__(self, instances, n_outputs: int | None = None, outputs=None):
        self.dev = device()
        self.lib = _lib.load()
        n = len(instances)
        self.n = n
        self.arr = (_lib.MftInstance * max(n, 1))()
        self._keep = []
        nout_of = {}
        for i, d in enumerate(instances):
            _untracked_inputs("mft_batched_psf", d["field"])
            f = as_complex(d["field"], self.dev)
            if f.ndim != 2 or f.shape[0] != f.shape[1]:
                raise ValueError(f"instance {i}: field must be square, got {tuple(f.shape)}")
            self._keep.append(f)
            e = self.arr[i]
            e.field = f.data_ptr()
            e.n_in = int(f.shape[0])
            e.n_out = int(d["n_out"])
            e.out_index = int(d.get("out_index", 0))
            fl = d.get("focal_length")
            e.inverse = int(bool(d.get("inverse", False)))
            e.has_focal = 0 if fl is None else 1
            e.focal_length = 0.0 if fl is None else float(fl)
            e.focal_shift = float(d.get("focal_shift", 0.0))
            e.wavelength = float(d["wavelength"])
            e.pixel_scale_in = float(d["pixel_scale_in"])
            e.pixel_scale_out = float(d["pixel_scale_out"])
            sh = d.get("shift", (0.0, 0.0))
            e.shift_x, e.shift_y = float(sh[0]), float(sh[1])
            e.weight = float(d.get("weight", 1.0))
            nout_of.setdefault(e.out_index, e.n_out)
        if n_outputs is None:
            n_outputs = (max(nout_of) + 1) if nout_of else 0
        self.n_outputs = int(n_outputs)
        if outputs is None:
            outputs = [torch.empty((nout_of[o], nout_of[o]), dtype=torch.float32,
                                   device=self.dev) if o in nout_of else None
                       for o in range(self.n_outputs)]
        self.outputs = outputs
        self.ptrs = (ctypes.c_void_p * max(self.n_outputs, 1))(
            *[(0 if t is None else t.data_ptr()) for t in outputs])
        self.ws_bytes = int(self.lib.dlx_mft_batched_workspace_size(n, self.arr,
                                                                    self.n_outputs))
        self.ws = workspace(self.ws_bytes, self.dev)

    @classmethod
    def from_arrays(cls, fields, field_index, wavelength, pixel_scale_in, n_out,
                    pixel_scale_out, weight, out_index, shift_x=0.0, shift_y=0.0,
                    slice_index=0, n_outputs: int | None = None,
                    ws: torch.Tensor | None = None):
        """Vectorised construction: `fields` is a list of complex64 device
        tensors, each (N, N) or a stack (B, N, N); instance i uses slice
        slice_index[i] of fields[field_index[i]].  Every other argument is a
        scalar or one value per instance.  The descriptor array is filled from a
        numpy structured array with the C layout, without per-instance Python.
        ``ws``: an optional caller workspace (used when large enough)."""
        self = cls.__new__(cls)
        self.dev = device()
        self.lib = _lib.load()
        _untracked_inputs("mft_batched_psf", *fields)
        self._keep = [as_complex(f, self.dev) for f in fields]
        for f in self._keep:
            if f.ndim not in (2, 3) or f.shape[-1] != f.shape[-2]:
                raise ValueError(f"fields must be (N, N) or (B, N, N), got {tuple(f.shape)}")
        fi = np.asarray(field_index, np.int64).reshape(-1)
        n = fi.size
        self.n = n
        si = np.broadcast_to(np.asarray(slice_index, np.int64), (n,))
        depth = np.array([f.shape[0] if f.ndim == 3 else 1 for f in self._keep], np.int64)
        if n and (np.any(fi < 0) or np.any(fi >= len(self._keep)) or np.any(si < 0) or
                  np.any(si >= depth[fi])):
            raise ValueError("field_index / slice_index outside the given fields")
        rec = np.zeros(max(n, 1), dtype=_lib.INSTANCE_DTYPE)[:n]
        ptr = np.array([f.data_ptr() for f in self._keep], np.uint64)
        side = np.array([f.shape[-1] for f in self._keep], np.int64)
        rec["field"] = ptr[fi] + (si * side[fi] * side[fi] * 8).astype(np.uint64)
        rec["n_in"] = side[fi].astype(np.int32)
        rec["n_out"] = np.broadcast_to(np.asarray(n_out, np.int32), (n,))
        rec["out_index"] = np.broadcast_to(np.asarray(out_index, np.int32), (n,))
        for name, v in (("wavelength", wavelength), ("pixel_scale_in", pixel_scale_in),
                        ("pixel_scale_out", pixel_scale_out), ("shift_x", shift_x),
                        ("shift_y", shift_y), ("weight", weight)):
            rec[name] = np.broadcast_to(np.asarray(v, np.float64), (n,))
        self._rec = np.ascontiguousarray(rec)
        self.arr = self._rec.ctypes.data_as(ctypes.POINTER(_lib.MftInstance))
        if n_outputs is None:
            n_outputs = int(rec["out_index"].max()) + 1 if n else 0
        self.n_outputs = int(n_outputs)
        sizes = {}
        for o, k in zip(rec["out_index"].tolist(), rec["n_out"].tolist()):
            sizes.setdefault(o, k)
        self.outputs = [torch.empty((sizes[o], sizes[o]), dtype=torch.float32, device=self.dev)
                        if o in sizes else None for o in range(self.n_outputs)]
        self.ptrs = (ctypes.c_void_p * max(self.n_outputs, 1))(
            *[(0 if t is None else t.data_ptr()) for t in self.outputs])
        self.ws_bytes = int(self.lib.dlx_mft_batched_workspace_size(n, self.arr,
                                                                    self.n_outputs))
        self.ws = ws if ws is not None and ws.numel() >= self.ws_bytes else \
            workspace(self.ws_bytes, self.dev)
        return self

    def __call__(self) -> list:
        st = self.lib.dlx_mft_batched_psf(self.n, self.arr, self.n_outputs, self.ptrs,
                                          self.ws.data_ptr(), self.ws.numel(), stream_handle())
        _lib.check(st, "dlx_mft_batched_psf")
        return self.outputs


def mft_batched_psf(instances, n_outputs: int | None = None, outputs=None,
                    ws: torch.Tensor | None = None) -> list:
    """Uneven batched PSFs (dlx_mft_batched_psf).

    ``instances`` is a sequence of dicts with keys ``field`` (complex64
    n_in x n_in device tensor; instances may share one), ``wavelength``,
    ``pixel_scale_in``, ``n_out``, ``pixel_scale_out`` and optionally
    ``shift`` ((sx, sy) output pixels), ``weight`` (1), ``out_index`` (0),
    ``focal_length`` (None), ``inverse`` (False).  Returns the list of output
    images: output o is the sum over its instances of weight * |MFT(field)|^2
    (None for an output no instance maps to, unless ``outputs`` is given).
    """
    return MftBatch(instances, n_outputs, outputs)()


# ---------------------------------------------------------------------------
# Dynamic apertures (dlx_soft_apertures / dlx_aperture_basis)
# ---------------------------------------------------------------------------


def soft_apertures(npix: int, pixel_scale: float, apertures, center: float = 0.0,
                   out: torch.Tensor | None = None) -> torch.Tensor:
    """Transmission (npix x npix float32) of a list of ``_lib.SoftAperture``
    descriptors: sum over groups of the product over each group."""
    n = len(apertures)
    dev = device()
    lib = _lib.load()
    arr = (_lib.SoftAperture * max(n, 1))(*apertures)
    if out is None:
        out = torch.empty((npix, npix), dtype=torch.float32, device=dev)
    ws = workspace(lib.dlx_soft_apertures_workspace_size(int(npix), n), dev)
    st = lib.dlx_soft_apertures(int(npix), float(pixel_scale), float(center), n, arr,
                                out.data_ptr(), ws.data_ptr(), ws.numel(), stream_handle())
    _lib.check(st, "dlx_soft_apertures")
    return out


def soft_apertures_vjp(npix: int, pixel_scale: float, apertures, dout: torch.Tensor,
                       center: float = 0.0) -> torch.Tensor:
    """(n, 9) float32 dL/dtheta of each aperture's geometry parameters (size
    0, size 1, and translation x, y, shear x, y, compression x, y, rotation of
    the last transform of its chain) given dL/dtransmission
    (dlx_soft_apertures_vjp)."""
    n = len(apertures)
    dev = device()
    lib = _lib.load()
    dout = dout.detach()
    if dout.dtype != torch.float32 or not dout.is_contiguous() or dout.device != dev:
        dout = dout.to(device=dev, dtype=torch.float32).contiguous()
    if tuple(dout.shape) != (npix, npix):
        raise ValueError(f"dout must be ({npix}, {npix}), got {tuple(dout.shape)}")
    arr = (_lib.SoftAperture * max(n, 1))(*apertures)
    grads = torch.empty((max(n, 1), 9), dtype=torch.float32, device=dev)
    ws = workspace(lib.dlx_soft_apertures_vjp_workspace_size(int(npix), n), dev)
    st = lib.dlx_soft_apertures_vjp(int(npix), float(pixel_scale), float(center), n, arr,
                                    dout.data_ptr(), grads.data_ptr(), ws.data_ptr(),
                                    ws.numel(), stream_handle())
    _lib.check(st, "dlx_soft_apertures_vjp")
    return grads[:n]


class _SoftAperturesFn(torch.autograd.Function):
    """Transmission of soft apertures as a function of their tracked geometry
    tensors; the backward is dlx_soft_apertures_vjp."""

    @staticmethod
    def forward(ctx, npix, pixel_scale, center, apertures, smap, keep, *params):
        # keep: device tensors the descriptors point into (coordinate planes),
        # held until the backward has run
        ctx.meta = (npix, pixel_scale, center, apertures, smap, keep)
        ctx.shapes = [(p.shape, p.dtype, p.device) for p in params]
        return soft_apertures(npix, pixel_scale, apertures, center)

    @staticmethod
    def backward(ctx, gout):
        npix, pixel_scale, center, apertures, smap, _ = ctx.meta
        g = soft_apertures_vjp(npix, pixel_scale, apertures, gout, center).double().cpu()
        grads = [None] * len(ctx.shapes)
        for a, slot, i, elem in smap:
            shape, dtype, dev = ctx.shapes[i]
            part = torch.zeros(shape, dtype=torch.float64)
            if elem is None:
                part.fill_(float(g[a, slot]))
            else:
                part.reshape(-1)[elem] = g[a, slot]
            grads[i] = part if grads[i] is None else grads[i] + part
        grads = [None if q is None else q.to(device=ctx.shapes[i][2], dtype=ctx.shapes[i][1])
                 for i, q in enumerate(grads)]
        return (None, None, None, None, None, None) + tuple(grads)


def soft_apertures_tracked(npix: int, pixel_scale: float, apertures, center, params, smap,
                           keep=()):
    """soft_apertures whose result carries gradients to the geometry tensors
    ``params``; ``smap`` lists (aperture, parameter slot, tensor index,
    element or None) for each slot that feeds a tensor; ``keep`` holds the
    device buffers the descriptors' ``coords`` point to until the backward."""
    return _SoftAperturesFn.apply(int(npix), float(pixel_scale), float(center),
                                  list(apertures), list(smap), tuple(keep), *params)


def aperture_basis(npix: int, pixel_scale: float, js, transforms=(), extent: float = 1.0,
                   nsides: int = 0, center: float = 0.0, coords=None) -> torch.Tensor:
    """(len(js), npix, npix) Zernike / polike basis on transformed coordinates
    divided by `extent` (dlx_aperture_basis); `coords` (2, npix, npix) float64
    device planes replace the transforms when given."""
    js = [int(j) for j in np.atleast_1d(np.asarray(js))]
    dev = device()
    tfs = (_lib.CoordTransformC * max(len(transforms), 1))(*transforms)
    jarr = (ctypes.c_int32 * max(len(js), 1))(*js)
    out = torch.empty((len(js), npix, npix), dtype=torch.float32, device=dev)
    st = _lib.load().dlx_aperture_basis(int(npix), float(pixel_scale), float(center),
                                        len(transforms), tfs,
                                        None if coords is None else coords.data_ptr(),
                                        float(extent), int(nsides), len(js), jarr,
                                        out.data_ptr(), stream_handle())
    _lib.check(st, "dlx_aperture_basis")
    return out


# ---------------------------------------------------------------------------
# Resampling (dlx_interp_transform / dlx_interp_points)
# ---------------------------------------------------------------------------

_INTERP_METHODS = {"nearest": 0, "linear": 1}


def _interp_method(method: str) -> int:
    if method not in _INTERP_METHODS:
        raise NotImplementedError(
            f"interpolation method {method!r}: 'linear' and 'nearest' are implemented")
    return _INTERP_METHODS[method]


def _interp_prepare(images, dev):
    x = as_float(images, dev).contiguous()
    if x.ndim < 2 or x.shape[-1] != x.shape[-2]:
        raise ValueError(f"images must be (..., n, n), got {tuple(x.shape)}")
    lead = tuple(x.shape[:-2])
    return x, lead, (int(np.prod(lead)) if lead else 1)


def _interp_transform_fwd(x, lead, batch, pixel_scale_in, center_in, n_out, pixel_scale_out,
                          center_out, transforms, method, fill):
    tfs = (_lib.CoordTransformC * max(len(transforms), 1))(*transforms)
    out = torch.empty(lead + (int(n_out), int(n_out)), dtype=torch.float32, device=x.device)
    st = _lib.load().dlx_interp_transform(int(x.shape[-1]), float(pixel_scale_in),
                                          float(center_in), int(n_out), float(pixel_scale_out),
                                          float(center_out), len(transforms), tfs,
                                          _interp_method(method), float(fill), batch,
                                          x.data_ptr(), out.data_ptr(), stream_handle())
    _lib.check(st, "dlx_interp_transform")
    return out


def _interp_points_fwd(x, lead, batch, pixel_scale_in, center_in, qx, qy, method, fill):
    out = torch.empty(lead + tuple(qx.shape), dtype=torch.float32, device=x.device)
    st = _lib.load().dlx_interp_points(int(x.shape[-1]), float(pixel_scale_in),
                                       float(center_in), qx.numel(), qx.data_ptr(),
                                       qy.data_ptr(), _interp_method(method), float(fill),
                                       batch, x.data_ptr(), out.data_ptr(), stream_handle())
    _lib.check(st, "dlx_interp_points")
    return out


class _InterpFn(torch.autograd.Function):
    """Resampling as a function of the images; the backward is the transpose of
    the sampling (dlx_interp_transform_vjp / dlx_interp_points_vjp)."""

    @staticmethod
    def forward(ctx, images, spec):
        x, lead, batch = _interp_prepare(images, images.device)
        ctx.meta = (tuple(x.shape), lead, batch, spec)
        if spec[0] == "transform":
            return _interp_transform_fwd(x, lead, batch, *spec[1:])
        ctx.save_for_backward(spec[3], spec[4])
        return _interp_points_fwd(x, lead, batch, spec[1], spec[2], spec[3], spec[4], *spec[5:])

    @staticmethod
    def backward(ctx, gout):
        shape, lead, batch, spec = ctx.meta
        n_in = shape[-1]
        g = gout.to(torch.float32).contiguous()
        din = torch.empty(shape, dtype=torch.float32, device=g.device)
        ws = torch.empty(max(batch * n_in * n_in, 1), dtype=torch.float64, device=g.device)
        lib = _lib.load()
        if spec[0] == "transform":
            _, d_in, c_in, n_out, d_out, c_out, transforms, method, _ = spec
            tfs = (_lib.CoordTransformC * max(len(transforms), 1))(*transforms)
            st = lib.dlx_interp_transform_vjp(n_in, float(d_in), float(c_in), int(n_out),
                                              float(d_out), float(c_out), len(transforms), tfs,
                                              _interp_method(method), batch, g.data_ptr(),
                                              din.data_ptr(), ws.data_ptr(), stream_handle())
            _lib.check(st, "dlx_interp_transform_vjp")
        else:
            qx, qy = ctx.saved_tensors
            _, d_in, c_in, _, _, method, _ = spec
            st = lib.dlx_interp_points_vjp(n_in, float(d_in), float(c_in), qx.numel(),
                                           qx.data_ptr(), qy.data_ptr(), _interp_method(method),
                                           batch, g.data_ptr(), din.data_ptr(), ws.data_ptr(),
                                           stream_handle())
            _lib.check(st, "dlx_interp_points_vjp")
        return din, None


def _tracks(images) -> bool:
    return isinstance(images, torch.Tensor) and images.requires_grad and torch.is_grad_enabled()


def interp_transform(images, pixel_scale_in: float, center_in: float, n_out: int,
                     pixel_scale_out: float, center_out: float, transforms=(),
                     method: str = "linear", fill: float = 0.0) -> torch.Tensor:
    """Float32 images (..., n_in, n_in) resampled onto an n_out grid whose pixel
    coordinates go through the ``_lib.CoordTransformC`` chain first.
    Differentiable with respect to the images (dlx_interp_transform_vjp)."""
    dev = device()
    _interp_method(method)
    if _tracks(images):
        return _InterpFn.apply(images.to(dev), ("transform", pixel_scale_in, center_in, n_out,
                                                pixel_scale_out, center_out, tuple(transforms),
                                                method, fill))
    x, lead, batch = _interp_prepare(images, dev)
    return _interp_transform_fwd(x, lead, batch, pixel_scale_in, center_in, n_out,
                                 pixel_scale_out, center_out, transforms, method, fill)


def interp_points(images, pixel_scale_in: float, center_in: float, xs, ys,
                  method: str = "linear", fill: float = 0.0) -> torch.Tensor:
    """Float32 images (..., n, n) sampled at the points (xs, ys) (any common
    shape) -> (..., *xs.shape).  Differentiable with respect to the images
    (dlx_interp_points_vjp); the points are constants."""
    dev = device()
    _interp_method(method)
    qx = torch.as_tensor(xs, device=dev).detach().to(torch.float64).contiguous()
    qy = torch.as_tensor(ys, device=dev).detach().to(torch.float64).contiguous()
    if qx.shape != qy.shape:
        raise ValueError("xs and ys must have the same shape")
    if _tracks(images):
        return _InterpFn.apply(images.to(dev), ("points", pixel_scale_in, center_in, qx, qy,
                                                method, fill))
    x, lead, batch = _interp_prepare(images, dev)
    return _interp_points_fwd(x, lead, batch, pixel_scale_in, center_in, qx, qy, method, fill)


# ---------------------------------------------------------------------------
# Detector chain (dlx_detector_apply / dlx_detector_vjp)
# ---------------------------------------------------------------------------

DETECTOR_KINDS = {
    "pixel_response": _lib.DLX_DET_PIXEL_RESPONSE,
    "convolve": _lib.DLX_DET_CONVOLVE,
    "downsample": _lib.DLX_DET_DOWNSAMPLE,
    "saturate": _lib.DLX_DET_SATURATE,
    "add": _lib.DLX_DET_ADD_CONSTANT,
    "convolve_separable": _lib.DLX_DET_CONVOLVE_SEPARABLE,
}


def _detector_ops(ops, dev):
    """[(kind name, array or scalar)] -> (dlx_detector_op array, tensors kept alive).
    pixel_response / convolve take a 2-D array, convolve_separable a pair
    (column factor, row factor) of 1-D arrays of one length, downsample an int
    factor, saturate / add a scalar."""
    arr = (_lib.DetectorOp * max(len(ops), 1))()
    keep = []
    for i, (kind, arg) in enumerate(ops):
        if kind not in DETECTOR_KINDS:
            raise ValueError(f"unknown detector op {kind!r}")
        op = arr[i]
        op.kind = DETECTOR_KINDS[kind]
        if kind in ("pixel_response", "convolve"):
            t = as_float(arg, dev).contiguous()
            if t.ndim != 2 or t.shape[0] != t.shape[1]:
                raise ValueError(f"{kind} needs a square 2d array, got {tuple(t.shape)}")
            keep.append(t)
            op.size = int(t.shape[0])
            op.data = t.data_ptr()
        elif kind == "convolve_separable":
            wy, wx = (as_float(a, dev).reshape(-1) for a in arg)
            if wy.numel() != wx.numel():
                raise ValueError("convolve_separable factors must have the same length")
            if wy.is_contiguous() and wx.is_contiguous() and \
                    wx.data_ptr() == wy.data_ptr() + 4 * wy.numel():
                keep += [wy, wx]   # already one buffer (column factor, then row factor)
                op.data = wy.data_ptr()
            else:
                t = torch.cat([wy, wx]).contiguous()
                keep.append(t)
                op.data = t.data_ptr()
            op.size = int(wy.numel())
        elif kind == "downsample":
            op.size = int(arg)
        else:
            op.value = float(arg.detach().float().cpu()) if isinstance(arg, torch.Tensor) \
                else float(np.asarray(arg, np.float32))
    return arr, keep


def detector_out_side(n: int, ops) -> int:
    """Side of the image the chain produces from an n x n input (validates it)."""
    arr, _keep = _detector_ops(ops, device())
    side = ctypes.c_int32(0)
    _lib.check(_lib.load().dlx_detector_out_side(int(n), arr, len(ops), ctypes.byref(side)),
               "dlx_detector_out_side")
    return int(side.value)


def _detector_impl(images, ops) -> torch.Tensor:
    dev = device()
    x = as_float(images, dev).contiguous()
    if x.ndim < 2 or x.shape[-1] != x.shape[-2]:
        raise ValueError(f"detector input must be (..., n, n), got {tuple(x.shape)}")
    n = int(x.shape[-1])
    lead = tuple(x.shape[:-2])
    batch = int(np.prod(lead)) if lead else 1
    lib = _lib.load()
    arr, keep = _detector_ops(ops, dev)
    side = ctypes.c_int32(0)
    _lib.check(lib.dlx_detector_out_side(n, arr, len(ops), ctypes.byref(side)),
               "dlx_detector_out_side")
    out = torch.empty(lead + (side.value, side.value), dtype=torch.float32, device=dev)
    ws = workspace(lib.dlx_detector_workspace_size(n, batch, arr, len(ops), 0), dev)
    st = lib.dlx_detector_apply(n, batch, x.data_ptr(), arr, len(ops), out.data_ptr(),
                                ws.data_ptr(), ws.numel(), stream_handle())
    _lib.check(st, "dlx_detector_apply")
    del keep
    return out


def _detector_vjp_impl(images, ops, dout, param_ops=()):
    """d/dimages (and, for the op indices in param_ops, d/d each op's map or
    value) of the detector chain for the cotangent dout."""
    dev = device()
    x = as_float(images, dev).contiguous()
    g = as_float(dout, dev).contiguous()
    n = int(x.shape[-1])
    lead = tuple(x.shape[:-2])
    batch = int(np.prod(lead)) if lead else 1
    lib = _lib.load()
    arr, keep = _detector_ops(ops, dev)
    side = ctypes.c_int32(0)
    _lib.check(lib.dlx_detector_out_side(n, arr, len(ops), ctypes.byref(side)),
               "dlx_detector_out_side")
    if tuple(g.shape) != lead + (side.value, side.value):
        raise ValueError(f"cotangent shape {tuple(g.shape)} != output shape "
                         f"{lead + (side.value, side.value)}")
    din = torch.empty_like(x)
    if not param_ops:
        ws = workspace(lib.dlx_detector_workspace_size(n, batch, arr, len(ops), 1), dev)
        st = lib.dlx_detector_vjp(n, batch, x.data_ptr(), arr, len(ops), g.data_ptr(),
                                  din.data_ptr(), ws.data_ptr(), ws.numel(), stream_handle())
        _lib.check(st, "dlx_detector_vjp")
        del keep
        return din
    grads = {}
    ptrs = (ctypes.c_void_p * max(len(ops), 1))()
    for k in param_ops:
        kind = ops[k][0]
        if kind in ("pixel_response", "convolve"):
            t = torch.empty((arr[k].size, arr[k].size), dtype=torch.float32, device=dev)
        elif kind == "convolve_separable":
            t = torch.empty((2 * arr[k].size,), dtype=torch.float32, device=dev)
        elif kind in ("saturate", "add"):
            t = torch.empty((), dtype=torch.float32, device=dev)
        else:
            raise NotImplementedError(f"a {kind} op has no parameter gradient")
        grads[k] = t
        ptrs[k] = t.data_ptr()
    ws = workspace(lib.dlx_detector_vjp_params_workspace_size(n, batch, arr, len(ops)), dev)
    st = lib.dlx_detector_vjp_params(n, batch, x.data_ptr(), arr, len(ops), g.data_ptr(),
                                     din.data_ptr(), ptrs, ws.data_ptr(), ws.numel(),
                                     stream_handle())
    _lib.check(st, "dlx_detector_vjp_params")
    del keep
    return din, grads


def _tracked_param(arg) -> bool:
    return isinstance(arg, torch.Tensor) and arg.requires_grad


class _DetectorFn(torch.autograd.Function):
    """The detector chain as a function of the images and of the tracked op
    parameters (response maps, convolution kernels or their separable factors,
    saturation levels, constants) listed in pmap as (op index, position in
    the op's tuple or None); the backward is dlx_detector_vjp(_params)."""

    @staticmethod
    def forward(ctx, images, ops, pmap, *params):
        ctx.ops = ops
        ctx.pmap = pmap
        ctx.save_for_backward(images)
        return _detector_impl(images, ops)

    @staticmethod
    def backward(ctx, g):
        (images,) = ctx.saved_tensors
        need = ctx.needs_input_grad[3:]
        want = sorted({k for (k, _), nd in zip(ctx.pmap, need) if nd})
        nones = [None] * len(ctx.pmap)
        if not want:
            return _detector_vjp_impl(images, ctx.ops, g), None, None, *nones
        din, grads = _detector_vjp_impl(images, ctx.ops, g, want)
        out = []
        for (k, j), nd in zip(ctx.pmap, need):
            if not nd:
                out.append(None)
                continue
            gk = grads[k]
            arg = ctx.ops[k][1]
            ref = arg if j is None else arg[j]
            if j is not None:   # separable factors: the column then the row values
                half = gk.numel() // 2
                gk = gk[:half] if j == 0 else gk[half:]
            out.append(gk.to(device=ref.device, dtype=ref.dtype).reshape(ref.shape))
        return (din if ctx.needs_input_grad[0] else None), None, None, *out


def detector(images, ops) -> torch.Tensor:
    """The detector chain `ops` applied to float32 images (..., n, n), as one
    dlx_detector_apply call; differentiable in `images` and in the ops'
    arrays and values given as tensors that require grad -- response maps,
    convolution kernels (or separable factors), saturation levels, constants
    (dlx_detector_vjp / dlx_detector_vjp_params)."""
    pmap, params = [], []
    if torch.is_grad_enabled():
        for k, (kind, arg) in enumerate(ops):
            parts = list(enumerate(arg)) if isinstance(arg, (tuple, list)) else [(None, arg)]
            for j, a in parts:
                if _tracked_param(a):
                    if kind == "downsample":
                        raise NotImplementedError("a downsample factor has no gradient")
                    pmap.append((k, j))
                    params.append(a)
    track_images = isinstance(images, torch.Tensor) and images.requires_grad
    if torch.is_grad_enabled() and (track_images or pmap):
        return _DetectorFn.apply(images, list(ops), pmap, *params)
    return _detector_impl(images, ops)


def psf_poly_detector(n_in: int, n_out: int, wavelengths, weights, pixel_scale_in: float,
                      pixel_scale_out: float, ops, amp=None, opd=None, phase=None, scale=1.0,
                      tilt=(0.0, 0.0), center=0.0, focal_length=None, inverse=False,
                      precision: str | None = None) -> torch.Tensor:
    """The fused polychromatic PSF (as psf_poly) through the detector chain
    `ops` (as detector), with the wavelength sum fused into the chain's first
    pass (dlx_psf_poly_detector): equal to detector(psf_poly(...), ops).
    When gradients are recorded for the pupil inputs or the ops' parameters it
    runs as that composition, whose two parts are differentiable."""
    wl, w, g, prec = _poly_args(n_in, n_out, wavelengths, weights, focal_length, inverse,
                                precision)
    dev = device()
    amp_t, opd_t, ph_t = _pupil_inputs(n_in, amp, opd, phase, dev)
    if torch.is_grad_enabled():
        tracked = any(t is not None and t.requires_grad for t in (amp_t, opd_t, ph_t))
        for _kind, arg in ops:
            parts = arg if isinstance(arg, (tuple, list)) else (arg,)
            tracked = tracked or any(_tracked_param(a) for a in parts)
        if tracked:
            psf = psf_poly(n_in, n_out, wl, w, pixel_scale_in, pixel_scale_out, amp=amp_t,
                           opd=opd_t, phase=ph_t, scale=scale, tilt=tilt, center=center,
                           focal_length=focal_length, inverse=inverse, precision=precision)
            return detector(psf, ops)
    lib = _lib.load()
    arr, keep = _detector_ops(ops, dev)
    side = ctypes.c_int32(0)
    _lib.check(lib.dlx_detector_out_side(int(n_out), arr, len(ops), ctypes.byref(side)),
               "dlx_detector_out_side")
    out = torch.empty((side.value, side.value), dtype=torch.float32, device=dev)
    ws = workspace(lib.dlx_psf_poly_detector_workspace_size(g, wl.size, prec, arr, len(ops)),
                   dev)
    st = lib.dlx_psf_poly_detector(g, wl.size, _lib.dptr(wl), _lib.dptr(w), float(pixel_scale_in),
                                   float(pixel_scale_out), float(center), float(tilt[0]),
                                   float(tilt[1]), _ptr(amp_t), _ptr(opd_t), _ptr(ph_t),
                                   float(scale), arr, len(ops), out.data_ptr(), prec,
                                   ws.data_ptr(), ws.numel(), stream_handle())
    _lib.check(st, "dlx_psf_poly_detector")
    del keep
    return out


def detector_vjp(images, ops, dout) -> torch.Tensor:
    """(d detector(images) / d images)^T dout (dlx_detector_vjp)."""
    return _detector_vjp_impl(images, ops, dout)


def _pow2(m: int) -> bool:
    return m >= 2 and (m & (m - 1)) == 0 and m <= 8192


def fft2(fields, m: int, inverse: bool = False) -> torch.Tensor:
    """Centred zero-padded 2-D FFT (dLux normalisation) of (B, N, N) -> (B, m, m).

    Power-of-two m runs the LDS FFT kernel (dlx_fft2).  Any other size runs
    the same transform exactly as a matrix DFT through dlx_mft: with
    wavelength 1, dx = 1/N, p_out = N/m and a half-pixel shift for even N, the
    MFT's transfer matrices are the centred DFT matrices and its nfringes
    normalisation is 1/m.
    """
    _untracked_inputs("fft2", fields)
    f = as_complex(fields)
    squeeze = f.ndim == 2
    if squeeze:
        f = f[None]
    b, n, _ = f.shape
    m = int(m)
    if (m - n) % 2 or m < n:
        raise ValueError(f"padded size {m} must be >= {n} with even difference")
    if _pow2(m):
        lib = _lib.load()
        ws = workspace(lib.dlx_fft2_workspace_size(n, m, b), f.device)
        out = torch.empty((b, m, m), dtype=torch.complex64, device=f.device)
        st = lib.dlx_fft2(n, m, b, int(bool(inverse)), f.data_ptr(), out.data_ptr(),
                          ws.data_ptr(), ws.numel(), stream_handle())
        _lib.check(st, "dlx_fft2")
    else:
        s = 0.5 if n % 2 == 0 else 0.0
        out = mft(f, 1.0, 1.0 / n, m, n / m, shift_x=s, shift_y=s, inverse=inverse,
                  precision="fp32")
    return out[0] if squeeze else out


def asm_prop(fields, wavelengths, pixel_scale: float, distance: float, m: int) -> torch.Tensor:
    """Angular-spectrum propagation of (B, N, N) fields padded to (B, m, m)."""
    _untracked_inputs("asm_prop", fields)
    f = as_complex(fields)
    squeeze = f.ndim == 2
    if squeeze:
        f = f[None]
    b, n, _ = f.shape
    wl = _bcast(wavelengths, b)
    lib = _lib.load()
    ws = workspace(lib.dlx_asm_workspace_size(n, int(m), b), f.device)
    out = torch.empty((b, int(m), int(m)), dtype=torch.complex64, device=f.device)
    st = lib.dlx_asm_prop(n, int(m), b, _lib.dptr(wl), float(pixel_scale), float(distance),
                          f.data_ptr(), out.data_ptr(), ws.data_ptr(), ws.numel(),
                          stream_handle())
    _lib.check(st, "dlx_asm_prop")
    return out[0] if squeeze else out


def lct(fields, wavelengths, abcd, pixel_scale_in: float, center_in: float, n_out: int,
        pixel_scale_out: float, center_out: float = 0.0,
        precision: str | None = None) -> torch.Tensor:
    """Collins-integral LCT of (B, N, N) fields to (B, n_out, n_out) on the
    given output grid (dlx_lct)."""
    _untracked_inputs("lct", fields)
    f = as_complex(fields)
    squeeze = f.ndim == 2
    if squeeze:
        f = f[None]
    b, n, _ = f.shape
    wl = _bcast(wavelengths, b)
    m = np.asarray(abcd, float).reshape(-1)
    if m.size != 4:
        raise ValueError("abcd must be a 2 x 2 matrix")
    prec = _prec_code(precision)
    lib = _lib.load()
    ws = workspace(lib.dlx_lct_workspace_size(n, int(n_out), b, prec), f.device)
    out = torch.empty((b, int(n_out), int(n_out)), dtype=torch.complex64, device=f.device)
    st = lib.dlx_lct(n, int(n_out), b, _lib.dptr(m), _lib.dptr(wl), float(pixel_scale_in),
                     float(center_in), float(pixel_scale_out), float(center_out), f.data_ptr(),
                     out.data_ptr(), prec, ws.data_ptr(), ws.numel(), stream_handle())
    _lib.check(st, "dlx_lct")
    return out[0] if squeeze else out


def lct_fft(fields, wavelengths, abcd, pixel_scale_in: float, center_in: float, m: int,
            center_out: float = 0.0):
    """FFT-grid LCT of (B, N, N) fields to (B, m, m) (dlx_lct_fft); returns
    (fields, output pixel scale per field)."""
    _untracked_inputs("lct_fft", fields)
    f = as_complex(fields)
    squeeze = f.ndim == 2
    if squeeze:
        f = f[None]
    b, n, _ = f.shape
    wl = _bcast(wavelengths, b)
    mat = np.asarray(abcd, float).reshape(-1)
    if mat.size != 4:
        raise ValueError("abcd must be a 2 x 2 matrix")
    lib = _lib.load()
    ws = workspace(lib.dlx_lct_fft_workspace_size(n, int(m), b), f.device)
    out = torch.empty((b, int(m), int(m)), dtype=torch.complex64, device=f.device)
    ps = _lib.dptr([0.0] * b)
    st = lib.dlx_lct_fft(n, int(m), b, _lib.dptr(mat), _lib.dptr(wl), float(pixel_scale_in),
                         float(center_in), float(center_out), f.data_ptr(), out.data_ptr(), ps,
                         ws.data_ptr(), ws.numel(), stream_handle())
    _lib.check(st, "dlx_lct_fft")
    pss = np.array(ps[:b], float)
    return (out[0], pss[0]) if squeeze else (out, pss)


def profile_enable(enable: bool = True) -> None:
    """Start (and reset) or stop live event timing of the MFT contraction launches."""
    _lib.check(_lib.load().dlx_profile_enable(int(bool(enable))), "dlx_profile_enable")


def profile_read(stage: int) -> tuple[float, int]:
    """(total ms, launch count) of contraction stage 1 or 2 since profile_enable."""
    import ctypes
    t = ctypes.c_double()
    c = ctypes.c_int32()
    _lib.check(_lib.load().dlx_profile_read(int(stage), ctypes.byref(t), ctypes.byref(c)),
               "dlx_profile_read")
    return float(t.value), int(c.value)


def batched_graphs(enable: bool | None = None) -> bool:
    """Graph replay of the uneven batched launches (dlx_batched_graphs):
    set it when `enable` is given; returns the previous setting."""
    prev = _lib.load().dlx_batched_graphs(-1 if enable is None else int(bool(enable)))
    if prev < 0:
        _lib.check(prev, "dlx_batched_graphs")
    return bool(prev)


def batched_graph_stats() -> dict:
    """Graphs captured, graph launches and graphs held by the uneven batched
    path (dlx_batched_graph_stats)."""
    import ctypes
    c, l, e = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int32()
    _lib.check(_lib.load().dlx_batched_graph_stats(ctypes.byref(c), ctypes.byref(l),
                                                   ctypes.byref(e)), "dlx_batched_graph_stats")
    return {"captures": int(c.value), "launches": int(l.value), "entries": int(e.value)}


def cache_clear() -> None:
    """Free every cached transfer matrix (after their pending uses)."""
    _lib.check(_lib.load().dlx_cache_clear(), "dlx_cache_clear")


def cache_set_budget(nbytes: int) -> None:
    """Device memory the transfer-matrix cache may hold (0 disables it)."""
    _lib.check(_lib.load().dlx_cache_set_budget(int(nbytes)), "dlx_cache_set_budget")


def mfma_rate(random_operands: bool = True) -> float:
    """bf16 MFMA TFLOP/s this GPU sustains from registers (dlx_mfma_rate), with
    full-entropy operands (as the split-BF16x3 contractions see) or low-entropy
    ones (the nominal rate)."""
    dev = device()
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    ws = workspace(2 * cus * 256 * 4, dev)
    out = ctypes.c_double(0.0)
    _lib.check(_lib.load().dlx_mfma_rate(int(bool(random_operands)), ctypes.byref(out),
                                         ws.data_ptr(), ws.numel(), stream_handle()),
               "dlx_mfma_rate")
    return float(out.value)


class cache_bypass:
    """Context manager: calls made from this thread inside it skip the
    transfer-matrix cache (dlx_cache_bypass), generating their matrices into
    their workspace -- no allocation, eviction or wait."""

    def __enter__(self):
        _lib.check(_lib.load().dlx_cache_bypass(1), "dlx_cache_bypass")
        return self

    def __exit__(self, *exc):
        _lib.check(_lib.load().dlx_cache_bypass(0), "dlx_cache_bypass")
        return False


def cache_stats() -> dict:
    """{'hits', 'misses', 'bytes', 'entries'} of the transfer-matrix cache."""
    import ctypes
    h, m = ctypes.c_uint64(), ctypes.c_uint64()
    b, e = ctypes.c_size_t(), ctypes.c_int32()
    _lib.check(_lib.load().dlx_cache_stats(ctypes.byref(h), ctypes.byref(m), ctypes.byref(b),
                                           ctypes.byref(e)), "dlx_cache_stats")
    return {"hits": h.value, "misses": m.value, "bytes": b.value, "entries": e.value}


class graph_step:
    """A fixed-shape step -- forward model, loss and backward, as fn() -- captured
    once in a HIP graph and replayed, so a training or fitting loop costs one
    graph launch of host time per step instead of the Python and dispatch work
    of every call.  The role jax.jit plays for dLux's gradient loops.

    fn reads its inputs from tensors that the caller updates in place
    (``t.copy_(new)``, ``with torch.no_grad(): p.copy_(...)``) and zeroes
    gradients in place (``p.grad.zero_()`` when ``p.grad`` exists), never
    ``p.grad = None``.  Calling the object replays the graph and returns fn's
    outputs from the capture: static tensors overwritten by every replay.
    fn is run ``warmup`` times on a side stream first (the transfer matrices
    are cached then, and parameter gradients allocated)."""

    def __init__(self, fn, warmup: int = 3):
        self.graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(max(int(warmup), 1)):
                fn()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        with torch.cuda.graph(self.graph):
            self.outputs = fn()

    def __call__(self):
        self.graph.replay()
        return self.outputs
