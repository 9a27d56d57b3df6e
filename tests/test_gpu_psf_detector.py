"""The detector fused into the wavelength sum (dlx_psf_poly_detector): every
wavelength's w_l |E_l|^2 plane goes straight into the detector chain, whose
first pass loads their sum, so the PSF is never written.

Checks: identical bits to dlx_psf_poly followed by dlx_detector_apply (the
planes are summed in the same wavelength order) for every kind of first
pass (convolve + downsample, convolve, downsample, elementwise), one and two
wavelength chunks, both arithmetic forms, small frames (planes summed before
the chain) and large ones (summed by the first pass's loads); the FP64 oracle (poly_psf then the
detector restatement) within 1e-5 of the peak; Telescope.model of a point
source through a LayeredDetector taking the fused call, equal to the
unfused model, and differentiable (it runs as the composition then).
"""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

PSF_TOL = 1e-5


@pytest.fixture(scope="module")
def dl():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dlux_amd
    return dlux_amd


def _err(x, ref):
    x = x.detach().double().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


def _pupil(n, seed):
    mask, basis = orc.circular_aperture(n, 1.0, oversample=1, spider_width=0.03,
                                        spider_angles=[0, 120, 240], zernike_nolls=range(2, 12))
    c = np.random.default_rng(seed).normal(0, 40e-9, basis.shape[0])
    opd = np.tensordot(np.nan_to_num(basis), c, axes=(0, 0))
    return mask, opd


def _chains(n_out, rng):
    jit = orc.jitter_kernel(0.7, 7, 3)
    f = jit.sum(0)   # jit = outer(f, f)
    return {
        "jitter-response-down-sat-add": [
            ("convolve_separable", (f, f)),
            ("pixel_response", rng.uniform(0.8, 1.2, (n_out, n_out))),
            ("downsample", 4), ("saturate", 3e-3), ("add", 1e-5)],
        "response-convolve": [("pixel_response", rng.uniform(0.8, 1.2, (n_out, n_out))),
                              ("convolve", jit)],
        "down-response": [("downsample", 2),
                          ("pixel_response", rng.uniform(0.8, 1.2, (n_out // 2, n_out // 2)))],
        "elementwise": [("pixel_response", rng.uniform(0.8, 1.2, (n_out, n_out))),
                        ("add", 2e-6)],
    }


def _device(chain):
    out = []
    for kind, arg in chain:
        if isinstance(arg, tuple):
            arg = tuple(torch.as_tensor(a, dtype=torch.float32, device="cuda") for a in arg)
        elif isinstance(arg, np.ndarray):
            arg = torch.as_tensor(arg, dtype=torch.float32, device="cuda")
        out.append((kind, arg))
    return out


# 64^2 frames: every chain's first launch is small, so the planes are summed by
# the pixel-parallel plane sum first; 1024^2 frames: the first launch has >= 512
# workgroups and sums the planes as it loads them
@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("n,n_out,nwl", [(96, 64, 7), (96, 64, 70), (128, 1024, 7)])
def test_fused_equals_psf_then_detector_and_oracle(dl, prec, n, n_out, nwl):
    from dlux_amd import runtime
    p_out = 2.5e-7
    mask, opd = _pupil(n, 3)
    wls = np.linspace(0.85e-6, 1.15e-6, nwl)
    w = np.random.default_rng(nwl).uniform(0.5, 1.5, nwl)
    w /= w.sum()
    scale = 1.0 / np.sqrt(np.sum(mask ** 2))
    kw = dict(amp=mask.astype(np.float32), opd=opd.astype(np.float32), scale=scale,
              tilt=(0.7 * p_out, -1.2 * p_out), precision=prec)
    # the OPD enters the oracle as a one-mode basis with coefficient 1
    ref_psf = orc.poly_psf(mask, opd[None], np.ones(1), wls, w, 1.0, n_out, p_out,
                           offset=(0.7 * p_out, -1.2 * p_out))
    for name, chain in _chains(n_out, np.random.default_rng(5)).items():
        print(f"[psf+detector] prec={prec} {n}->{n_out} nwl={nwl} chain={name}", flush=True)
        ops = _device(chain)
        fused = runtime.psf_poly_detector(n, n_out, wls, w, 1.0 / n, p_out, ops, **kw)
        psf = runtime.psf_poly(n, n_out, wls, w, 1.0 / n, p_out, **kw)
        two = runtime.detector(psf, ops)
        assert fused.shape == two.shape
        assert torch.equal(fused, two), name
        ref = orc.detector(ref_psf, chain)
        err = _err(fused, ref)
        print(f"  err vs oracle {err:.2e}")
        assert err < PSF_TOL, name


def test_telescope_point_source_takes_the_fused_call(dl, monkeypatch):
    from dlux_amd import layers, runtime, workloads
    cfg = workloads.MONO
    optics = workloads.build_system(cfg)
    wls = [0.9e-6, 1.0e-6, 1.1e-6]
    src = dl.PointSource(wls, position=[1.3 * cfg.p_out, -0.4 * cfg.p_out], flux=2.5e4,
                         weights=[1.0, 2.0, 1.5])
    resp = np.random.default_rng(2).uniform(0.9, 1.1, (cfg.n_out, cfg.n_out))
    det = dl.LayeredDetector([layers.ApplyJitter(0.8, 7, 3), layers.ApplyPixelResponse(resp),
                              layers.Downsample(4), layers.AddConstant(3.0)])
    tel = dl.Telescope(optics, src, det)
    calls = []
    real = runtime.psf_poly_detector

    def counting(*a, **k):
        calls.append(1)
        return real(*a, **k)
    monkeypatch.setattr(runtime, "psf_poly_detector", counting)
    for prec in ("fp32", "bf16x3"):
        with runtime.precision(prec):
            got = tel.model(return_psf=True)
            psf = optics.model(src, return_psf=True)
            unfused = det.model(psf, return_psf=True)
        assert torch.equal(got.data, unfused.data)
        assert np.isclose(float(got.pixel_scale), float(unfused.pixel_scale))
    assert len(calls) == 2
    # against the oracle
    mask = orc.circular_aperture(cfg.n_in, 1.0, oversample=1)
    basis = orc.zernike_basis(cfg.js, orc.pixel_coords(cfg.n_in, 1.0), 1.0)
    ww = np.array([1.0, 2.0, 1.5]) / 4.5 * 2.5e4
    img = orc.poly_psf(mask, basis, cfg.coefficients(), wls, ww, 1.0, cfg.n_out, cfg.p_out,
                       offset=(1.3 * cfg.p_out, -0.4 * cfg.p_out))
    ref = orc.detector(img, [("convolve", orc.jitter_kernel(0.8, 7, 3)), ("pixel_response", resp),
                             ("downsample", 4), ("add", 3.0)])
    assert _err(got.data, ref) < PSF_TOL


def test_fused_path_with_gradients_runs_as_the_composition(dl):
    """Coefficients that require grad: the fused entry point runs as
    psf_poly -> detector, whose backward gives the same gradient as the
    unfused Telescope model."""
    from dlux_amd import layers, workloads
    cfg = workloads.MONO
    c = torch.tensor(cfg.coefficients(), dtype=torch.float32, device="cuda", requires_grad=True)
    optics = workloads.build_system(cfg, c)
    src = dl.PointSource([0.95e-6, 1.05e-6], flux=1e4)
    det = dl.LayeredDetector([layers.ApplyJitter(0.6, 5, 3), layers.Downsample(2)])
    out = dl.Telescope(optics, src, det).model()
    (out ** 2).sum().backward()
    g1 = c.grad.clone()
    c.grad = None
    psf = optics.model(src, return_psf=True)
    out2 = det.model(psf)
    (out2 ** 2).sum().backward()
    assert torch.equal(out.detach(), out2.detach())
    assert torch.allclose(g1, c.grad, rtol=1e-6, atol=0)
