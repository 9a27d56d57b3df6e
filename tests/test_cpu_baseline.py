"""The FP32 CPU restatement timed as the CPU baseline computes what the FP64
oracle computes, to FP32 accuracy (CPU only).

Bounds: PSFs 1e-5 of the peak (the hot path's bound; FP32 reaches ~5e-7),
single-instance intensities 1e-5 of the peak, coefficient gradients 1e-3 of
the largest component (FP32 sums over 512^2 pixels).
"""

import numpy as np

from oracle import dlux_fp32 as f32
from oracle import dlux_oracle as orc


def _err(x, ref):
    return float(np.max(np.abs(np.asarray(x, np.float64) - ref)) / np.max(np.abs(ref)))


def test_fp32_psf_matches_fp64_oracle():
    mask, basis = orc.circular_aperture(128, 1.0, oversample=1, spider_width=0.03,
                                        spider_angles=[0, 120, 240], zernike_nolls=range(2, 16))
    c = np.random.default_rng(1).normal(0, 30e-9, basis.shape[0])
    wls = np.linspace(0.8e-6, 1.2e-6, 6)
    w = np.ones(6) / 6
    for offset in ((0.0, 0.0), (1.3e-6, -0.4e-6)):
        ref = orc.poly_psf(mask, basis, c, wls, w, 1.0, 64, 2.5e-7, offset=offset)
        got = f32.poly_psf(mask.astype(np.float32), basis.astype(np.float32), c, wls, w, 1.0,
                           64, 2.5e-7, offset=offset)
        assert got.dtype == np.float32
        assert _err(got, ref) < 1e-5


def test_fp32_gradient_matches_fp64_oracle():
    rng = np.random.default_rng(2)
    mask, basis = orc.circular_aperture(96, 1.0, oversample=1, zernike_nolls=range(2, 12))
    c = rng.normal(0, 50e-9, basis.shape[0])
    wls = np.linspace(0.9e-6, 1.1e-6, 4)
    w = np.ones(4) / 4
    data = rng.poisson(1e6 * orc.poly_psf(mask, basis, c * 0.7, wls, w, 1.0, 48, 2.5e-7))
    loss_ref, g_ref = orc.poly_psf_grad(mask, basis, c, wls, w, 1.0, 48, 2.5e-7, data, 1e6)
    loss, g = f32.poly_psf_grad(mask.astype(np.float32), basis.astype(np.float32), c, wls, w,
                                1.0, 48, 2.5e-7, data, 1e6)
    assert abs(float(loss) - loss_ref) < 1e-4 * loss_ref
    assert _err(g, g_ref) < 1e-3


def test_fp32_single_instance_with_shift():
    rng = np.random.default_rng(3)
    n = 64
    f = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / n
    ref = np.abs(orc.mft(f, 1.1e-6, 1.0 / n, 40, 2.5e-7, shift=(1.7, -2.25))) ** 2
    got = f32.mft_intensity(f, 1.1e-6, 1.0 / n, 40, 2.5e-7, shift=(1.7, -2.25))
    assert _err(got, ref) < 1e-5
