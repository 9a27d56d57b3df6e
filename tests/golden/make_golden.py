"""Write the small FP64 golden fixtures in this directory.

The reference holds no golden vectors for this path, so these come from the
FP64 oracle (oracle/dlux_oracle.py), whose own pins are in test_oracle.py.
They freeze the oracle's outputs on seeded inputs (a change to the oracle
that alters them fails test_oracle.py) and give the GPU tests fixed
expected values.  Every fixture stores its inputs, parameters and seed.

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import dlux_oracle as orc  # noqa: E402


def _field(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / n


def main():
    # MFT of the reference tests' fixture geometry (32 -> 16, wl 1, dx 0.1, p 0.05)
    # in every mode (tests/utils/test_propagation.py:95-176 use this geometry)
    E = _field(32, 7)
    modes = [(None, False, (0.0, 0.0)), (2.0, False, (1.0, -2.0)), (2.0, True, (1.0, -2.0)),
             (None, True, (0.5, 0.5)), (None, False, (3.0, -1.5))]
    outs = []
    for f, inv, sh in modes:
        outs.append(orc.mft(E, 1.0, 0.1, 16, 0.05, focal_length=f, shift=sh, inverse=inv))
    np.savez_compressed(HERE / "mft_32to16.npz", field=E, out=np.stack(outs),
             focal=np.array([np.nan if f is None else f for f, _, _ in modes]),
             inverse=np.array([inv for _, inv, _ in modes]),
             shift=np.array([sh for _, _, sh in modes]), seed=7)

    # config-1 style monochromatic PSF at reduced size: 64 -> 32, Noll 2..11 ~ N(0, 20 nm)
    mask, basis = orc.circular_aperture(64, 1.0, oversample=1, zernike_nolls=range(2, 12))
    c = np.random.default_rng(0).normal(0, 20e-9, basis.shape[0])
    psf = orc.poly_psf(mask, basis, c, [1e-6], [1.0], 1.0, 32, 2.5e-7)
    np.savez_compressed(HERE / "psf_mono_64to32.npz", mask=mask, basis=basis, coeffs=c, psf=psf,
             wavelength=1e-6, p_out=2.5e-7, seed=0)

    # three-wavelength polychromatic sum with uneven weights and spiders: 48 -> 24
    mask, basis = orc.circular_aperture(48, 1.0, oversample=3, spider_width=0.04,
                                        spider_angles=[0, 120, 240],
                                        zernike_nolls=range(2, 8))
    c = np.random.default_rng(1).normal(0, 30e-9, basis.shape[0])
    wls = np.array([0.9e-6, 1.0e-6, 1.1e-6])
    w = np.array([0.2, 0.5, 0.3])
    psf = orc.poly_psf(mask, basis, c, wls, w, 1.0, 24, 2.5e-7)
    np.savez_compressed(HERE / "psf_poly3_48to24.npz", mask=mask, basis=basis, coeffs=c, psf=psf,
             wavelengths=wls, weights=w, p_out=2.5e-7, seed=1)

    # Zernike basis, Noll 1..10, on 16^2 pixel coordinates of a unit disk
    zb = orc.zernike_basis(range(1, 11), orc.pixel_coords(16, 1.0), 1.0)
    np.savez_compressed(HERE / "zernike_16.npz", js=np.arange(1, 11), basis=zb)

    # centred FFT 16 -> 32 (pad 2), forward and inverse
    E = _field(16, 3)
    fwd = orc.fft(E, 1.0, 1.0, pad=2)[0]
    inv = orc.fft(E, 1.0, 1.0, pad=2, inverse=True)[0]
    np.savez_compressed(HERE / "fft_16to32.npz", field=E, forward=fwd, inverse=inv, seed=3)

    # gradient of the Gaussian loss: 32^2 pupil, Noll 2..6, two wavelengths -> 16^2
    rng = np.random.default_rng(8)
    mask, basis = orc.circular_aperture(32, 1.0, oversample=1, zernike_nolls=range(2, 7))
    c = rng.normal(0, 60e-9, basis.shape[0])
    wls, w = np.array([0.9e-6, 1.1e-6]), np.array([0.4, 0.6])
    data = rng.poisson(1e4 * orc.poly_psf(mask, basis, c * 0.5, wls, w, 1.0, 16, 3e-7))
    loss, grad = orc.poly_psf_grad(mask, basis, c, wls, w, 1.0, 16, 3e-7, data, 1e4)
    np.savez_compressed(HERE / "grad_32to16.npz", mask=mask, basis=basis, coeffs=c, wavelengths=wls,
             weights=w, data=data.astype(np.float64), flux=1e4, p_out=3e-7, loss=loss,
             grad=grad, seed=8)
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


if __name__ == "__main__":
    main()
