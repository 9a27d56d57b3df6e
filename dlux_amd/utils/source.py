"""Binary-source helpers (utils/source.py:10-52)."""

from __future__ import annotations

import numpy as np

__all__ = ["fluxes_from_contrast", "positions_from_sep"]


def fluxes_from_contrast(mean_flux, contrast) -> np.ndarray:
    """(flux_1, flux_2) of a binary with the given mean flux and contrast
    flux_1 / flux_2 (utils/source.py:10-26): both average to mean_flux."""
    mean_flux = np.asarray(mean_flux, float)
    contrast = np.asarray(contrast, float)
    return 2.0 * np.array([contrast * mean_flux, mean_flux]) / (1.0 + contrast)


def positions_from_sep(position, separation, position_angle) -> np.ndarray:
    """((x1, y1), (x2, y2)) of a binary centred on `position`, each star half
    the separation away along the position angle, measured from +y towards +x
    (utils/source.py:29-52)."""
    position = np.asarray(position, float)
    half = np.asarray(separation, float) / 2.0
    phi = np.asarray(position_angle, float)
    offset = np.array([half * np.sin(phi), half * np.cos(phi)])
    return np.array([position + offset, position - offset])
