"""Angular unit conversions (host scalars / arrays).  Mirrors utils/units.py:24-312."""

from __future__ import annotations

from functools import lru_cache

import numpy as np

_BASE_TO_RAD = {
    "rad": 1.0,
    "deg": np.pi / 180.0,
    "arcmin": np.pi / (180.0 * 60.0),
    "arcsec": np.pi / (180.0 * 3600.0),
}
_ALIASES = {
    "radian": "rad", "radians": "rad", "degree": "deg", "degrees": "deg",
    "am": "arcmin", "arcmins": "arcmin", "arcminute": "arcmin", "arcminutes": "arcmin",
    "as": "arcsec", "arcsecs": "arcsec", "arcsecond": "arcsec", "arcseconds": "arcsec",
}
_PREFIX = {"G": 1e9, "M": 1e6, "k": 1e3, "": 1.0, "m": 1e-3, "u": 1e-6, "n": 1e-9}


def _canon(unit: str) -> str:
    if not isinstance(unit, str):
        raise TypeError("unit must be a string.")
    u = unit.strip()
    if not u:
        raise ValueError("unit cannot be empty.")
    return _ALIASES.get(u.lower(), u)


@lru_cache(maxsize=None)
def unit_factor_to_rad(unit: str) -> float:
    """Factor f with value_in_unit * f = value_in_radians (utils/units.py:71-90)."""
    u = _canon(unit)
    if u in _BASE_TO_RAD:
        return _BASE_TO_RAD[u]
    p, base = u[0], u[1:]
    if p in _PREFIX and base in _BASE_TO_RAD:
        return _PREFIX[p] * _BASE_TO_RAD[base]
    raise ValueError(f"Unknown angular unit '{unit}'.")


def convert(value, unit_in: str, unit_out: str, scale: float = 1.0):
    """value in unit_in expressed in (scale * unit_out)  (utils/units.py:93-107)."""
    return value * (unit_factor_to_rad(unit_in) / unit_factor_to_rad(unit_out)) / scale


def rad2arcsec(v): return convert(v, "rad", "arcsec")
def rad2deg(v): return convert(v, "rad", "deg")
def rad2arcmin(v): return convert(v, "rad", "arcmin")
def deg2rad(v): return convert(v, "deg", "rad")
def deg2arcmin(v): return convert(v, "deg", "arcmin")
def deg2arcsec(v): return convert(v, "deg", "arcsec")
def arcmin2rad(v): return convert(v, "arcmin", "rad")
def arcmin2deg(v): return convert(v, "arcmin", "deg")
def arcmin2arcsec(v): return convert(v, "arcmin", "arcsec")
def arcsec2rad(v): return convert(v, "arcsec", "rad")
def arcsec2deg(v): return convert(v, "arcsec", "deg")
def arcsec2arcmin(v): return convert(v, "arcsec", "arcmin")
