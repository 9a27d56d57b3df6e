"""Stage 1 of the split-BF16x3 PSF with the pupil generated in its loader
(cgemm_pupil_kernel): no pupil array is written, the loader forms the same
split4 phasors as the pupil kernel (one shared definition), so the PSF must
be bit-identical to the two-kernel path (DLUX_AMD_PUPIL_PROLOGUE=0) for both
tile widths, with and without amplitude / OPD, with and without tilt, and equal to the
FP64 oracle within 1e-5 of the peak.
"""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

PSF_TOL = 1e-5


@pytest.fixture(scope="module")
def rt():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dlux_amd import runtime
    return runtime


def _pupil(n, seed):
    mask, basis = orc.circular_aperture(n, 1.0, oversample=1, spider_width=0.03,
                                        spider_angles=[0, 120, 240], zernike_nolls=range(2, 12))
    c = np.random.default_rng(seed).normal(0, 40e-9, basis.shape[0])
    opd = np.tensordot(np.nan_to_num(basis), c, axes=(0, 0))
    return mask, opd


# (n_in, n_out, wavelengths): 128 x 128 tiles (512 -> 256, 1024 -> 256), 128 x 64
# tiles (512 -> 128), and shapes the kernel does not cover (too few tiles, or
# more than two tile columns), where the pupil kernel runs
SHAPES = [(512, 256, 64), (1024, 256, 32), (512, 128, 64), (1024, 256, 16), (256, 256, 64),
          (256, 128, 16)]


# with a tilt the loader forms tx X + ty Y + OPD in FP64; with none it takes the
# OPD as the path, which must give the same bits as the pupil kernel's sum
@pytest.mark.parametrize("tilted", [True, False])
@pytest.mark.parametrize("n_in,n_out,nwl", SHAPES)
def test_prologue_bit_identical_to_pupil_kernel(rt, monkeypatch, n_in, n_out, nwl, tilted):
    print(f"[pupil prologue] {n_in}->{n_out} x{nwl} tilted={tilted}", flush=True)
    mask, opd = _pupil(n_in, n_in + nwl)
    p_out = 2.5e-7 * 512 / n_in
    wls = np.linspace(0.8e-6, 1.2e-6, nwl)
    w = np.random.default_rng(nwl).uniform(0.5, 1.5, nwl)
    scale = 1.0 / np.sqrt(np.sum(mask ** 2))
    cases = [dict(amp=mask.astype(np.float32), opd=opd.astype(np.float32)),
             dict(amp=None, opd=opd.astype(np.float32)),
             dict(amp=mask.astype(np.float32), opd=None)]
    for i, kw in enumerate(cases):
        args = (n_in, n_out, wls, w, 1.0 / n_in, p_out)
        tilt = (1.3 * p_out, -0.6 * p_out) if tilted else (0.0, 0.0)
        common = dict(scale=scale, tilt=tilt, precision="bf16x3", **kw)
        monkeypatch.delenv("DLUX_AMD_PUPIL_PROLOGUE", raising=False)
        fused = rt.psf_poly(*args, **common)
        monkeypatch.setenv("DLUX_AMD_PUPIL_PROLOGUE", "0")
        two = rt.psf_poly(*args, **common)
        monkeypatch.delenv("DLUX_AMD_PUPIL_PROLOGUE")
        torch.cuda.synchronize()
        assert torch.isfinite(fused).all()
        assert torch.equal(fused, two), f"case {i}"


def test_prologue_against_oracle_with_tilt(rt):
    n_in, n_out, nwl, p_out = 512, 128, 64, 5e-7
    mask, opd = _pupil(n_in, 11)
    wls = np.linspace(0.85e-6, 1.15e-6, nwl)
    w = np.random.default_rng(4).uniform(0.5, 1.5, nwl)
    w /= w.sum()
    scale = 1.0 / np.sqrt(np.sum(mask ** 2))
    tilt = (2.2 * p_out, 0.9 * p_out)
    got = rt.psf_poly(n_in, n_out, wls, w, 1.0 / n_in, p_out, amp=mask.astype(np.float32),
                      opd=opd.astype(np.float32), scale=scale, tilt=tilt, precision="bf16x3")
    ref = orc.poly_psf(mask, opd[None], np.ones(1), wls, w, 1.0, n_out, p_out, offset=tilt)
    err = float(np.max(np.abs(got.double().cpu().numpy() - ref)) / np.max(ref))
    print(f"  err vs oracle {err:.2e}")
    assert err < PSF_TOL
