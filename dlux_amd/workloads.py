"""Seeded synthetic workloads the hot path is measured and tested on.

Common geometry: D = 1 m, hard circular aperture of radius N_in/2 px (pixel
centres, r < R), Zernikes zernike_basis(js, pixel_coords(N_in, D), D) with
coefficients in metres, flat weights, wavelengths np.linspace(lo, hi, n), and
"2x Nyquist" output sampling p_out = lambda_ref / (4 D) with lambda_ref = 1 um
(2.5e-7 rad).  RNG: numpy default_rng(seed).

  mono    256^2 pupil, Noll 2..11 ~ N(0, 20 nm), seed 0, 1.0 um, 128^2 output
  poly    512^2 pupil + 3 spider vanes 8 px wide at 0/120/240 deg, Noll 2..37
          ~ N(0, 30 nm), seed 1, 64 wavelengths on 0.8-1.2 um, 256^2 output
  grad    512^2 pupil, Noll 2..37 ~ N(0, 50 nm), seed 2, 16 wavelengths on
          0.9-1.1 um, 128^2 output; data = Poisson(1e6-photon PSF of seed-3
          coefficients, seed 4)
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

D = 1.0
LAMBDA_REF = 1.0e-6
P_OUT = LAMBDA_REF / (4 * D)  # 2x Nyquist at lambda_ref, radians per pixel


@dataclass(frozen=True)
class PolyConfig:
    name: str
    n_in: int
    n_out: int
    js: tuple
    coeff_std: float
    seed: int
    wl_lo: float
    wl_hi: float
    n_wl: int
    spider_px: float = 0.0
    spider_angles: tuple = ()
    p_out: float = P_OUT
    extra: dict = field(default_factory=dict)

    @property
    def wavelengths(self) -> np.ndarray:
        if self.n_wl == 1:
            return np.array([self.wl_lo])
        return np.linspace(self.wl_lo, self.wl_hi, self.n_wl)

    @property
    def weights(self) -> np.ndarray:
        return np.ones(self.n_wl) / self.n_wl

    def coefficients(self, seed: int | None = None) -> np.ndarray:
        rng = np.random.default_rng(self.seed if seed is None else seed)
        return rng.normal(0.0, self.coeff_std, len(self.js))

    @property
    def mft_flops(self) -> float:
        """Algorithmic MFT FLOPs per PSF: 8 (N_out N_in^2 + N_out^2 N_in) per wavelength."""
        return 8.0 * (self.n_out * self.n_in ** 2 + self.n_out ** 2 * self.n_in) * self.n_wl


MONO = PolyConfig("mono", 256, 128, tuple(range(2, 12)), 20e-9, 0, 1.0e-6, 1.0e-6, 1)
POLY = PolyConfig("poly", 512, 256, tuple(range(2, 38)), 30e-9, 1, 0.8e-6, 1.2e-6, 64,
                  spider_px=8.0, spider_angles=(0.0, 120.0, 240.0))
GRAD = PolyConfig("grad", 512, 128, tuple(range(2, 38)), 50e-9, 2, 0.9e-6, 1.1e-6, 16,
                  extra={"target_seed": 3, "noise_seed": 4, "photons": 1e6})


def build_inputs(cfg: PolyConfig):
    """(mask, basis) device tensors from the product's own builders."""
    from .utils import apertures, coordinates, zernikes
    width = cfg.spider_px * D / cfg.n_in if cfg.spider_px else None
    angles = list(cfg.spider_angles) if cfg.spider_px else None
    mask = apertures.circular_aperture(cfg.n_in, D, oversample=1, spider_width=width,
                                       spider_angles=angles)
    basis = zernikes.zernike_basis(cfg.js, coordinates.pixel_coords(cfg.n_in, diameter=D), D)
    return mask, basis


def build_system(cfg: PolyConfig, coeffs=None, mask=None, basis=None):
    """AngularOpticalSystem: TransmissiveLayer(mask, normalise) -> BasisLayer -> MFT focus."""
    from . import layers
    from .optical_systems import AngularOpticalSystem
    from .utils import units
    if mask is None or basis is None:
        mask, basis = build_inputs(cfg)
    coeffs = cfg.coefficients() if coeffs is None else coeffs
    return AngularOpticalSystem(
        wf_npixels=cfg.n_in, diameter=D,
        layers=[("aperture", layers.TransmissiveLayer(mask, normalise=True)),
                ("aberrations", layers.BasisLayer(basis, coeffs))],
        psf_npixels=cfg.n_out, psf_pixel_scale=float(units.rad2arcsec(cfg.p_out)),
        oversample=1)
