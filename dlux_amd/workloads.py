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
  uneven  256 point sources (seed 5), each with its own complex64 pupil field
          (N_in drawn from {256, 512, 1024}; circular aperture, Noll 2..11
          ~ N(0, 30 nm) phase at 1 um), output size from {64, 128, 256}, tilt
          U(+-2 arcsec) per axis (an MFT shift of tilt / p_out pixels), flux
          log-uniform on 1e3-1e6, 16 wavelengths on 0.8-1.2 um with flat
          weights: 4096 (source x wavelength) instances, split-BF16x3
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


@dataclass(frozen=True)
class BatchConfig:
    name: str = "uneven"
    n_sources: int = 256
    n_wl: int = 16
    n_in_choices: tuple = (256, 512, 1024)
    n_out_choices: tuple = (64, 128, 256)
    tilt_arcsec: float = 2.0
    flux_lo: float = 1e3
    flux_hi: float = 1e6
    wl_lo: float = 0.8e-6
    wl_hi: float = 1.2e-6
    js: tuple = tuple(range(2, 12))
    coeff_std: float = 30e-9
    seed: int = 5
    p_out: float = P_OUT

    @property
    def wavelengths(self) -> np.ndarray:
        return np.linspace(self.wl_lo, self.wl_hi, self.n_wl)

    def sources(self) -> list:
        """Per-source draws: n_in, n_out, tilt (rad, x/y), flux, phase coefficients."""
        rng = np.random.default_rng(self.seed)
        out = []
        for _ in range(self.n_sources):
            n_in = int(rng.choice(self.n_in_choices))
            n_out = int(rng.choice(self.n_out_choices))
            tilt = rng.uniform(-self.tilt_arcsec, self.tilt_arcsec, 2) * np.pi / (180 * 3600)
            flux = float(np.exp(rng.uniform(np.log(self.flux_lo), np.log(self.flux_hi))))
            coeffs = rng.normal(0.0, self.coeff_std, len(self.js))
            out.append(dict(n_in=n_in, n_out=n_out, tilt=tilt, flux=flux, coeffs=coeffs))
        return out

    def flops(self, sources) -> float:
        """Algorithmic MFT FLOPs of one batch: sum over instances of F(N_in, N_out)."""
        return float(sum(8.0 * (s["n_out"] * s["n_in"] ** 2 + s["n_out"] ** 2 * s["n_in"])
                         for s in sources) * self.n_wl)


BATCH = BatchConfig()


def build_batch(cfg: BatchConfig, sources=None):
    """(instances, sources): device pupil fields and per-instance descriptors for
    runtime.mft_batched_psf (output index = source index)."""
    import torch
    from . import runtime
    from .utils import apertures, coordinates, zernikes
    sources = cfg.sources() if sources is None else sources
    dev = runtime.device()
    geo = {}
    instances = []
    for si, src in enumerate(sources):
        n = src["n_in"]
        if n not in geo:
            mask = apertures.circular_aperture(n, D, oversample=1)
            basis = zernikes.zernike_basis(cfg.js, coordinates.pixel_coords(n, diameter=D), D)
            geo[n] = (torch.as_tensor(mask, device=dev).float(),
                      torch.as_tensor(basis, device=dev).float())
        mask, basis = geo[n]
        c = torch.as_tensor(src["coeffs"], dtype=torch.float32, device=dev)
        opd = runtime.eval_basis(basis, c)
        phase = (2 * np.pi / LAMBDA_REF) * opd
        field = mask * torch.exp(1j * phase.to(torch.float32))
        field = (field / torch.sqrt(torch.sum(mask ** 2))).to(torch.complex64).contiguous()
        for wl in cfg.wavelengths:
            instances.append(dict(field=field, wavelength=float(wl), pixel_scale_in=D / n,
                                  n_out=src["n_out"], pixel_scale_out=cfg.p_out,
                                  shift=(float(src["tilt"][0] / cfg.p_out),
                                         float(src["tilt"][1] / cfg.p_out)),
                                  weight=src["flux"] / cfg.n_wl, out_index=si))
    return instances, sources
