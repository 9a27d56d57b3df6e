"""Device plumbing and thin typed wrappers over the C ABI.

torch provides device memory, the caching allocator and the current HIP
stream; every computation is a call into ``libdlux_amd.so``.  Inputs are
converted to contiguous device tensors (complex64 / float32) before the call.
"""

from __future__ import annotations

import contextlib
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


def eval_basis(basis, coeffs, base=None) -> torch.Tensor:
    """sum_j coeffs[..., j] basis[j] on the device; coeffs (nsets, nb) or (nb,)."""
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
    pix_shape = bt.shape[ct.ndim:]
    npix = int(np.prod(pix_shape))
    lib = _lib.load()
    out = torch.empty(pix_shape, dtype=torch.float32, device=dev)
    base_t = None if base is None else as_float(base, dev)
    st = lib.dlx_eval_basis(nb, npix, 1, bt.data_ptr(), ct.reshape(-1).data_ptr(),
                            _ptr(base_t), out.data_ptr(), stream_handle())
    _lib.check(st, "dlx_eval_basis")
    return out


def power(fields: torch.Tensor) -> torch.Tensor:
    """Per-instance sum |E|^2 of (B, ...) complex64 fields -> (B,) float32."""
    f = as_complex(fields)
    b = f.shape[0] if f.ndim == 3 else 1
    npix = f.numel() // b
    out = torch.empty(b, dtype=torch.float32, device=f.device)
    st = _lib.load().dlx_power(npix, b, f.data_ptr(), out.data_ptr(), stream_handle())
    _lib.check(st, "dlx_power")
    return out if f.ndim == 3 else out[0]


def intensity(fields: torch.Tensor, sum_batch: bool = False) -> torch.Tensor:
    """|E|^2 of complex64 fields (optionally summed over the leading axis)."""
    f = as_complex(fields)
    if f.ndim == 2:
        f = f[None]
    b = f.shape[0]
    npix = f[0].numel()
    shape = tuple(f.shape[1:]) if sum_batch else tuple(f.shape)
    out = torch.empty(shape, dtype=torch.float32, device=f.device)
    st = _lib.load().dlx_intensity(npix, b, f.data_ptr(), int(sum_batch), out.data_ptr(),
                                   stream_handle())
    _lib.check(st, "dlx_intensity")
    return out


def psf_poly_workspace(n_in, n_out, nwl, precision=None) -> int:
    g = geom(n_in, n_out)
    return int(_lib.load().dlx_psf_poly_workspace_size(g, int(nwl), _prec_code(precision)))


def psf_poly(n_in: int, n_out: int, wavelengths, weights, pixel_scale_in: float,
             pixel_scale_out: float, amp=None, opd=None, phase=None, scale=1.0,
             tilt=(0.0, 0.0), center=0.0, focal_length=None, inverse=False,
             precision: str | None = None, out: torch.Tensor | None = None,
             ws: torch.Tensor | None = None) -> torch.Tensor:
    """Fused polychromatic PSF sum_l w_l |MFT_l(P_l)|^2 -> (n_out, n_out) float32."""
    wl = np.atleast_1d(np.asarray(wavelengths, np.float64))
    w = np.atleast_1d(np.asarray(weights, np.float64))
    if w.shape != wl.shape:
        raise ValueError(f"Wavelength and weight shape mismatch: {wl.shape} vs {w.shape}")
    dev = device()
    lib = _lib.load()
    g = geom(n_in, n_out, focal_length, inverse)
    prec = _prec_code(precision)
    amp_t = None if amp is None else as_float(amp, dev)
    opd_t = None if opd is None else as_float(opd, dev)
    ph_t = None if phase is None else as_float(phase, dev)
    for name, t in (("amp", amp_t), ("opd", opd_t), ("phase", ph_t)):
        if t is not None and tuple(t.shape) != (n_in, n_in):
            raise ValueError(f"{name} must have shape {(n_in, n_in)}, got {tuple(t.shape)}")
    need = lib.dlx_psf_poly_workspace_size(g, wl.size, prec)
    if ws is None or ws.numel() < need:
        ws = workspace(need, dev)
    if out is None:
        out = torch.empty((int(n_out), int(n_out)), dtype=torch.float32, device=dev)
    st = lib.dlx_psf_poly(g, wl.size, _lib.dptr(wl), _lib.dptr(w), float(pixel_scale_in),
                          float(pixel_scale_out), float(center), float(tilt[0]),
                          float(tilt[1]), _ptr(amp_t), _ptr(opd_t), _ptr(ph_t), float(scale),
                          out.data_ptr(), prec, ws.data_ptr(), ws.numel(), stream_handle())
    _lib.check(st, "dlx_psf_poly")
    return out


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


def cache_clear() -> None:
    """Free every cached transfer matrix (after their pending uses)."""
    _lib.check(_lib.load().dlx_cache_clear(), "dlx_cache_clear")


def cache_set_budget(nbytes: int) -> None:
    """Device memory the transfer-matrix cache may hold (0 disables it)."""
    _lib.check(_lib.load().dlx_cache_set_budget(int(nbytes)), "dlx_cache_set_budget")


def cache_stats() -> dict:
    """{'hits', 'misses', 'bytes', 'entries'} of the transfer-matrix cache."""
    import ctypes
    h, m = ctypes.c_uint64(), ctypes.c_uint64()
    b, e = ctypes.c_size_t(), ctypes.c_int32()
    _lib.check(_lib.load().dlx_cache_stats(ctypes.byref(h), ctypes.byref(m), ctypes.byref(b),
                                           ctypes.byref(e)), "dlx_cache_stats")
    return {"hits": h.value, "misses": m.value, "bytes": b.value, "entries": e.value}
