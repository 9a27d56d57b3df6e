"""Basis evaluation and small numeric helpers (utils/math.py:137-249)."""

from __future__ import annotations

from math import lgamma, exp

import numpy as np
import torch

from .. import runtime

__all__ = ["eval_basis", "solve_basis", "factorial", "nandiv", "triangular_number",
           "gaussian"]


def eval_basis(basis, coefficients) -> torch.Tensor:
    """tensordot(basis, coefficients) over the coefficient dims, on the device.

    utils/math.py:177-196 (HIP kernel dlx_eval_basis).
    """
    return runtime.eval_basis(basis, coefficients)


def solve_basis(array, basis):
    """Least-squares coefficients of ``array`` over ``basis`` (its trailing
    dimensions match array's), shape basis.shape[:-array.ndim]
    (utils/math.py:199-225).  A host FP64 solve: fitting, not the hot path."""
    a = array.detach().double().cpu().numpy() if isinstance(array, torch.Tensor) \
        else np.asarray(array, float)
    b = basis.detach().double().cpu().numpy() if isinstance(basis, torch.Tensor) \
        else np.asarray(basis, float)
    out_shape = b.shape[-a.ndim:] if a.ndim else ()
    if out_shape != a.shape:
        raise ValueError(
            "The trailing dimensions of basis must match the shape of array, "
            f"received {out_shape} and {a.shape}."
        )
    matrix = b.reshape((-1, a.size)).T
    lead = b.shape[:-a.ndim] if a.ndim else b.shape
    return np.linalg.lstsq(matrix, a.ravel(), rcond=None)[0].reshape(lead)


def factorial(n: float) -> float:
    """n! as exp(lgamma(n + 1)), 1 at 0.  utils/math.py:137-157."""
    n = float(n)
    if n == 0:
        return 1.0
    if n < 0 and n.is_integer():
        return float("inf")
    return exp(lgamma(n + 1.0))


def triangular_number(n: int):
    return n * (n + 1) / 2


def nandiv(a, b, fill=float("inf")):
    """a / b with `fill` where b == 0.  utils/math.py:228-249."""
    a = torch.as_tensor(a)
    b = torch.as_tensor(b, device=a.device)
    safe = torch.where(b == 0, torch.ones_like(b), b)
    return torch.where(b == 0, torch.full_like(a / safe, fill), a / safe)


def gaussian(mean=0.0, std=1.0, npixels=64, extent: float = 5.0) -> np.ndarray:
    """Normalised separable Gaussian (utils/math.py:20-67): per axis the normal
    pdf (loc=mean, scale=std) sampled on linspace(-extent, extent, n), the
    outer product in 'xy' order, divided by its sum.  A small host array (a
    detector kernel), returned as float32."""
    npix = (int(npixels),) if np.ndim(npixels) == 0 else tuple(int(n) for n in npixels)
    ndim = max(len(npix), np.size(mean), np.size(std))
    mean = np.broadcast_to(np.asarray(mean, np.float64), (ndim,))
    std = np.broadcast_to(np.asarray(std, np.float64), (ndim,))
    if len(npix) != ndim:
        npix = npix * ndim
    one_d = [np.exp(-0.5 * ((np.linspace(-extent, extent, n) - m) / s) ** 2)
             / (s * np.sqrt(2.0 * np.pi)) for n, m, s in zip(npix, mean, std)]
    kernel = np.array(np.meshgrid(*one_d, indexing="xy")).prod(0)
    return (kernel / np.sum(kernel)).astype(np.float32)
