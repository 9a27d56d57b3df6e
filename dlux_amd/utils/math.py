"""Basis evaluation and small numeric helpers (utils/math.py:137-249)."""

from __future__ import annotations

from math import lgamma, exp

import torch

from .. import runtime

__all__ = ["eval_basis", "factorial", "nandiv", "triangular_number"]


def eval_basis(basis, coefficients) -> torch.Tensor:
    """tensordot(basis, coefficients) over the coefficient dims, on the device.

    utils/math.py:177-196 (HIP kernel dlx_eval_basis).
    """
    return runtime.eval_basis(basis, coefficients)


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


def nandiv(a, b, fill=0.0):
    """a / b with `fill` where b == 0.  utils/math.py:228-249."""
    a = torch.as_tensor(a)
    b = torch.as_tensor(b, device=a.device)
    safe = torch.where(b == 0, torch.ones_like(b), b)
    return torch.where(b == 0, torch.full_like(a / safe, fill), a / safe)
