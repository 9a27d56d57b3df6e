"""Immutable model objects with path-based updates.

dLux's classes are zodiax pytrees: every update (``set``, ``multiply``,
``add``, ``divide``) returns a new object and leaves the original untouched
(e.g. wavefronts.py:772, optical_systems.py:489-507).  ``Base`` gives the same
semantics to plain Python objects whose array leaves live on the GPU.  Paths
may be dotted (``"layers.aperture.transmission"``) and descend through
attributes and dict entries.
"""

from __future__ import annotations

import copy
from collections import OrderedDict
from typing import Any


def _split(path: str) -> list[str]:
    return path.split(".") if isinstance(path, str) else list(path)


def _get(obj: Any, key: str) -> Any:
    if isinstance(obj, dict):
        return obj[key]
    return getattr(obj, key)


def _replace(obj: Any, keys: list[str], value: Any) -> Any:
    """Return a shallow copy of obj with the leaf at keys replaced."""
    head, rest = keys[0], keys[1:]
    new_val = value if not rest else _replace(_get(obj, head), rest, value)
    if isinstance(obj, dict):
        out = OrderedDict(obj) if isinstance(obj, OrderedDict) else dict(obj)
        out[head] = new_val
        return out
    out = copy.copy(obj)
    object.__setattr__(out, head, new_val)
    return out


class Base:
    """Base class: immutable updates via ``set/get/multiply/add/divide``."""

    def get(self, path: str) -> Any:
        obj = self
        for k in _split(path):
            obj = _get(obj, k)
        return obj

    def set(self, path=None, value=None, **kwargs):
        """``obj.set("a.b", v)``, ``obj.set(["a", "b"], [v1, v2])`` or ``obj.set(a=v)``."""
        updates: list[tuple[str, Any]] = []
        if path is not None:
            if isinstance(path, (list, tuple)) and not isinstance(path, str):
                if not isinstance(value, (list, tuple)) or len(value) != len(path):
                    raise ValueError("paths and values must have the same length")
                updates.extend(zip(path, value))
            else:
                updates.append((path, value))
        updates.extend(kwargs.items())
        out = self
        for p, v in updates:
            out = _replace(out, _split(p), v)
        return out

    def multiply(self, path: str, value) -> "Base":
        return self.set(path, self.get(path) * value)

    def add(self, path: str, value) -> "Base":
        return self.set(path, self.get(path) + value)

    def divide(self, path: str, value) -> "Base":
        return self.set(path, self.get(path) / value)

    def __repr__(self) -> str:
        fields = []
        for k, v in vars(self).items():
            if k.startswith("_"):
                continue
            shape = getattr(v, "shape", None)
            if shape is not None and len(shape) > 0:
                fields.append(f"{k}={getattr(v, 'dtype', '')}{list(shape)}")
            else:
                fields.append(f"{k}={v!r}")
        return f"{type(self).__name__}({', '.join(fields)})"
