"""Coefficient-weighted basis parameterisations (parametric.py:28-158).

The bases are device tensors; ``evaluate`` contracts them with the
coefficients on the device (dlx_eval_basis, differentiable in the
coefficients through torch autograd).  Coefficients given as torch tensors
are kept as given, so gradients reach them.
"""

from __future__ import annotations

import numpy as np
import torch

from . import runtime
from .base import Base
from .utils.math import solve_basis

__all__ = ["Parametric", "ParametricBasis", "ExplicitBasis", "ImplicitBasis", "CoordBasis"]


def _shape(v) -> tuple:
    return tuple(v.shape)


def _as_coefficients(v):
    if isinstance(v, torch.Tensor):
        return v
    return np.asarray(v, dtype=float)


class Parametric(Base):
    """Abstract parameterisation that evaluates to a value (parametric.py:28-37)."""

    def evaluate(self, **kwargs):  # pragma: no cover - abstract
        raise NotImplementedError


class ParametricBasis(Parametric):
    """Coefficients over a basis (parametric.py:40-83)."""

    @property
    def coeffs(self):
        return self.coefficients

    @property
    def c(self):
        return self.coefficients

    @property
    def alpha(self):
        return self.coefficients

    @property
    def coefficient_shape(self) -> tuple:
        return _shape(self.coefficients)

    def _set_coefficients(self, coefficients, coefficient_shape) -> None:
        coefficients = _as_coefficients(coefficients)
        if _shape(coefficients) != tuple(coefficient_shape):
            raise ValueError(
                "Coefficient shape must match the leading basis dimensions. "
                f"Expected {tuple(coefficient_shape)}, got {_shape(coefficients)}."
            )
        self.coefficients = coefficients

    def solve_basis(self, value, **kwargs):  # pragma: no cover - abstract
        raise NotImplementedError


class ExplicitBasis(ParametricBasis):
    """A parameterisation over an explicitly sampled basis (parametric.py:86-121)."""

    def __init__(self, basis, coefficients=None, coefficient_shape=None):
        self.basis = runtime.as_float(basis)
        if coefficients is None:
            if coefficient_shape is None:
                raise ValueError("Provide either coefficients or coefficient_shape.")
            coefficients = np.zeros(coefficient_shape)
        else:
            coefficients = _as_coefficients(coefficients)
            if coefficient_shape is None:
                coefficient_shape = _shape(coefficients)
        coefficient_shape = tuple(coefficient_shape)
        if tuple(self.basis.shape[:len(coefficient_shape)]) != coefficient_shape:
            raise ValueError("The leading basis dimensions must match the coefficient shape.")
        self._set_coefficients(coefficients, coefficient_shape)

    def evaluate(self, **kwargs) -> torch.Tensor:
        return runtime.eval_basis(self.basis, self.coefficients)

    def solve_basis(self, value, **kwargs):
        return solve_basis(value, self.basis)


class ImplicitBasis(ParametricBasis):
    """A basis generated at evaluation time (parametric.py:124-139)."""

    def calculate_basis(self, **kwargs):  # pragma: no cover - abstract
        raise NotImplementedError

    def evaluate(self, **kwargs) -> torch.Tensor:
        return runtime.eval_basis(self.calculate_basis(**kwargs), self.coefficients)

    def solve_basis(self, value, **kwargs):
        return solve_basis(value, self.calculate_basis(**kwargs))


class CoordBasis(ImplicitBasis):
    """An implicit basis evaluated at Cartesian coordinates (parametric.py:142-158)."""

    @staticmethod
    def get_coordinates(*, wavefront=None, coordinates=None):
        if coordinates is not None:
            return coordinates
        if wavefront is None:
            raise ValueError("Provide either wavefront or coordinates.")
        return wavefront.coordinates()
