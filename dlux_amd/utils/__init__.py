"""``dlux_amd.utils`` -- the hot-path subset of ``dLux.utils`` (utils/__init__.py:1-59)."""

from . import units, coordinates, array_ops, propagation, fourier, zernikes, math
from . import geometry, apertures, source, interpolation
from .units import *  # noqa: F401,F403
from .coordinates import nd_coords, pixel_coords, cart2polar, polar2cart, rotate_coords
from .array_ops import pad_to, crop_to, resize, downsample
from .propagation import FFT, MFT, transfer_matrix, calc_nfringes
from .fourier import fft_spec, fft_phase_ramp
from .zernikes import (noll_indices, zernike_factors, zernike, zernike_basis, zernike_name,
                       radial_orders_to_indices, polike, polike_basis)
from .math import eval_basis, solve_basis, factorial, nandiv, triangular_number, gaussian
from .geometry import combine, circle, spider
from .apertures import circular_aperture
from .source import fluxes_from_contrast, positions_from_sep
from .interpolation import interp, scale, rotate
