"""GPU parity of the hand-written adjoint (dlx_psf_poly_vjp, dlx_project_basis)
against the FP64 oracle's adjoint, which test_oracle pins by central finite
differences.

Error metric: peak-normalised max |g - g_ref| / max |g_ref|.  Tolerances:
FP32 form 1e-4, split-BF16x3 form 3e-4 on dL/dOPD and dL/dphase (the adjoint
chains three contractions and a product with the pupil, each at the forward
form's precision); 1e-4 / 3e-4 on dL/dc through the dLux API.
"""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GRAD_TOL = {"fp32": 1e-4, "bf16x3": 3e-4}


def _peak_err(x, ref):
    x = np.asarray(x)
    ref = np.asarray(ref)
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def rt():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dlux_amd import runtime
    return runtime


def _pupil_case(n, nmodes, seed, std=40e-9):
    mask, basis = orc.circular_aperture(n, 1.0, oversample=2, spider_width=0.03,
                                        spider_angles=[0, 120, 240],
                                        zernike_nolls=range(2, 2 + nmodes))
    c = np.random.default_rng(seed).normal(0, std, basis.shape[0])
    return mask, basis, c


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("n_in,n_out,nwl", [(64, 32, 5), (48, 40, 3), (128, 64, 8)])
def test_psf_poly_vjp_vs_oracle(rt, prec, n_in, n_out, nwl):
    """dL/dOPD and dL/dphase for a random cotangent, every wavelength weighted."""
    print(f"[vjp] prec={prec} n_in={n_in} n_out={n_out} nwl={nwl}", flush=True)
    mask, basis, c = _pupil_case(n_in, 8, n_in + nwl)
    rng = np.random.default_rng(n_out)
    wls = np.linspace(0.85e-6, 1.15e-6, nwl)
    w = rng.uniform(0.5, 1.5, nwl)
    p_out = 1e-6 / (4 * 1.0)
    G = rng.standard_normal((n_out, n_out))
    g_opd_ref, g_ph_ref = orc.poly_psf_vjp(mask, basis, c, wls, w, 1.0, n_out, p_out, G)
    opd = np.tensordot(basis, c, axes=(0, 0))
    norm = 1.0 / np.sqrt(np.sum((mask / n_in ** 2) ** 2))
    g_opd, g_ph = rt.psf_poly_vjp(n_in, n_out, wls, w, 1.0 / n_in, p_out, G, amp=mask, opd=opd,
                                  scale=norm / n_in ** 2, precision=prec, need_phase=True)
    e1 = _peak_err(g_opd.cpu().numpy(), g_opd_ref)
    e2 = _peak_err(g_ph.cpu().numpy(), g_ph_ref)
    print(f"  err opd {e1:.3e} phase {e2:.3e}")
    assert e1 < GRAD_TOL[prec] and e2 < GRAD_TOL[prec]


def test_psf_poly_vjp_chunks_and_phase_only(rt):
    """70 wavelengths (two launch chunks, accumulated) and the phase-only output."""
    print("[vjp-70]", flush=True)
    n_in, n_out, nwl = 32, 24, 70
    mask, basis, c = _pupil_case(n_in, 5, 70)
    wls = np.linspace(0.8e-6, 1.2e-6, nwl)
    w = np.full(nwl, 1.0 / nwl)
    p_out = 2.5e-7
    G = np.random.default_rng(1).standard_normal((n_out, n_out))
    g_opd_ref, g_ph_ref = orc.poly_psf_vjp(mask, basis, c, wls, w, 1.0, n_out, p_out, G)
    opd = np.tensordot(basis, c, axes=(0, 0))
    norm = 1.0 / np.sqrt(np.sum((mask / n_in ** 2) ** 2))
    g_opd, none = rt.psf_poly_vjp(n_in, n_out, wls, w, 1.0 / n_in, p_out, G, amp=mask, opd=opd,
                                  scale=norm / n_in ** 2, precision="fp32")
    assert none is None
    assert _peak_err(g_opd.cpu().numpy(), g_opd_ref) < GRAD_TOL["fp32"]
    none, g_ph = rt.psf_poly_vjp(n_in, n_out, wls, w, 1.0 / n_in, p_out, G, amp=mask, opd=opd,
                                 scale=norm / n_in ** 2, precision="bf16x3", need_opd=False,
                                 need_phase=True)
    assert none is None
    assert _peak_err(g_ph.cpu().numpy(), g_ph_ref) < GRAD_TOL["bf16x3"]


@pytest.mark.parametrize("npix,nsets", [(4096, 1), (1001, 3), (262144, 2), (3, 1)])
def test_project_basis(rt, npix, nsets):
    print(f"[project] npix={npix} nsets={nsets}", flush=True)
    rng = np.random.default_rng(npix)
    basis = rng.standard_normal((7, npix)).astype(np.float32)
    g = rng.standard_normal((nsets, npix)).astype(np.float32)
    out = rt.project_basis(basis, g if nsets > 1 else g[0]).cpu().numpy()
    ref = g.astype(np.float64) @ basis.astype(np.float64).T
    ref = ref if nsets > 1 else ref[0]
    assert out.shape == ref.shape
    # FP32 sums of npix unit-variance products: error well under 1e-5 sqrt(npix)
    assert np.max(np.abs(out - ref)) < 1e-5 * np.sqrt(npix)


def test_eval_basis_autograd(rt):
    print("[eval-basis-autograd]", flush=True)
    rng = np.random.default_rng(2)
    basis = torch.as_tensor(rng.standard_normal((6, 40, 40)).astype(np.float32)).cuda()
    c = torch.as_tensor(rng.standard_normal(6).astype(np.float32)).cuda().requires_grad_(True)
    w = torch.as_tensor(rng.standard_normal((40, 40)).astype(np.float32)).cuda()
    out = rt.eval_basis(basis, c)
    (out * w).sum().backward()
    ref = np.tensordot(basis.cpu().numpy().astype(np.float64),
                       w.cpu().numpy().astype(np.float64), axes=([1, 2], [0, 1]))
    assert np.allclose(c.grad.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_api_gradient_matches_oracle(rt, prec):
    """dLux-style phase retrieval step: coefficients with requires_grad through
    AngularOpticalSystem.propagate, a Gaussian loss, loss.backward()."""
    print(f"[api-grad] prec={prec}", flush=True)
    from dlux_amd import layers
    from dlux_amd.optical_systems import AngularOpticalSystem
    from dlux_amd.utils import units

    n_in, n_out, nwl = 96, 48, 6
    mask, basis, c = _pupil_case(n_in, 10, 5, std=50e-9)
    wls = np.linspace(0.9e-6, 1.1e-6, nwl)
    w = np.ones(nwl) / nwl
    p_out = 2.5e-7
    rng = np.random.default_rng(4)
    truth = c + rng.normal(0, 20e-9, c.shape)
    flux = 1e6
    data = rng.poisson(flux * orc.poly_psf(mask, basis, truth, wls, w, 1.0, n_out, p_out))
    loss_ref, grad_ref = orc.poly_psf_grad(mask, basis, c, wls, w, 1.0, n_out, p_out, data, flux)

    coeffs = torch.tensor(c, dtype=torch.float32, device="cuda", requires_grad=True)
    optics = AngularOpticalSystem(
        wf_npixels=n_in, diameter=1.0,
        layers=[("aperture", layers.TransmissiveLayer(mask.astype(np.float32), normalise=True)),
                ("aberrations", layers.BasisLayer(basis.astype(np.float32), coeffs))],
        psf_npixels=n_out, psf_pixel_scale=float(units.rad2arcsec(p_out)))
    d = torch.as_tensor(data, dtype=torch.float32, device="cuda")
    with rt.precision(prec):
        psf = optics.propagate(wls, weights=w)
    var = torch.clamp(d, min=1.0)
    loss = 0.5 * torch.sum((flux * psf - d) ** 2 / var)
    loss.backward()
    g = coeffs.grad.cpu().numpy().astype(np.float64)
    err = float(np.max(np.abs(g - grad_ref)) / np.max(np.abs(grad_ref)))
    print(f"  loss {loss.item():.6e} ref {loss_ref:.6e}  grad err {err:.3e}")
    assert abs(loss.item() - loss_ref) < 1e-3 * loss_ref
    assert err < GRAD_TOL[prec]


def test_gradient_sign_by_finite_difference_on_device(rt):
    """Directional derivative of the device loss vs the device gradient (h = 0.5 nm,
    large enough for FP32 forward differences)."""
    print("[device-fd]", flush=True)
    n_in, n_out = 64, 32
    mask, basis, c = _pupil_case(n_in, 6, 9, std=40e-9)
    wls = np.array([0.95e-6, 1.05e-6])
    w = np.array([0.5, 0.5])
    p_out = 2.5e-7
    opd0 = np.tensordot(basis, c, axes=(0, 0))
    norm = 1.0 / np.sqrt(np.sum((mask / n_in ** 2) ** 2))
    target = orc.poly_psf(mask, basis, c * 0.3, wls, w, 1.0, n_out, p_out)
    T = torch.as_tensor(target, dtype=torch.float32, device="cuda")

    def loss_of(o):
        psf = rt.psf_poly(n_in, n_out, wls, w, 1.0 / n_in, p_out, amp=mask, opd=o,
                          scale=norm / n_in ** 2, precision="fp32")
        return torch.sum((psf - T).double() ** 2)

    opd = torch.tensor(opd0, dtype=torch.float32, device="cuda", requires_grad=True)
    L = loss_of(opd)
    L.backward()
    direction = np.tensordot(basis, np.random.default_rng(3).normal(0, 1, c.shape), axes=(0, 0))
    D = torch.as_tensor(direction, dtype=torch.float32, device="cuda")
    analytic = float(torch.sum(opd.grad.double() * D.double()))
    h = 0.5e-9
    with torch.no_grad():
        fd = (float(loss_of(opd + h * D)) - float(loss_of(opd - h * D))) / (2 * h)
    print(f"  analytic {analytic:.6e} fd {fd:.6e}")
    assert abs(analytic - fd) < 2e-3 * abs(fd)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_psf_poly_vjp_amplitude(rt, prec):
    """dL/d amp at fixed scale against a central difference of the oracle PSF
    along a random direction V (pixels where amp == 0 included: the
    cotangent comes from the unit-amplitude phasor, not P / amp).  Bound: the
    directional derivative to 1e-4 (fp32) / 3e-4 (bf16x3) relative."""
    print(f"[vjp amp] prec={prec}", flush=True)
    n_in, n_out, nwl = 64, 32, 3
    mask, basis, c = _pupil_case(n_in, 6, 17)
    assert np.sum(mask == 0) > 100
    rng = np.random.default_rng(23)
    wls = np.linspace(0.9e-6, 1.1e-6, nwl)
    w = rng.uniform(0.5, 1.5, nwl)
    p_out = 2.5e-7
    G = rng.standard_normal((n_out, n_out))
    opd = np.tensordot(basis, c, axes=(0, 0))
    _, _, g_amp = rt.psf_poly_vjp(n_in, n_out, wls, w, 1.0 / n_in, p_out, G, amp=mask, opd=opd,
                                  scale=1.0 / n_in ** 2, precision=prec, need_opd=False,
                                  need_amp=True)
    g_amp = g_amp.double().cpu().numpy()
    V = rng.standard_normal((n_in, n_in))
    h = 1e-6

    def L(A):
        return np.sum(G * orc.poly_psf(A, basis, c, wls, w, 1.0, n_out, p_out,
                                       normalise_power=False))
    fd = (L(mask + h * V) - L(mask - h * V)) / (2 * h)
    got = float(np.sum(g_amp * V))
    err = abs(got - fd) / abs(fd)
    # and along a direction supported only where amp == 0
    V0 = V * (mask == 0)
    fd0 = (L(mask + h * V0) - L(mask - h * V0)) / (2 * h)
    got0 = float(np.sum(g_amp * V0))
    err0 = abs(got0 - fd0) / abs(fd0)
    print(f"  dir err {err:.2e}  zero-support dir err {err0:.2e}")
    assert err < GRAD_TOL[prec] and err0 < 10 * GRAD_TOL[prec]


def test_transmission_gradient_through_api(rt):
    """A TransmissiveLayer whose transmission requires grad (with zeros, and a
    power normalisation): the autograd gradient against a central difference
    of the oracle chain along a random direction."""
    print("[vjp amp] TransmissiveLayer", flush=True)
    import dlux_amd as dl
    from dlux_amd.utils import units
    n, n_out, p_out = 64, 32, 2.5e-7
    mask, basis, c = _pupil_case(n, 5, 29)
    wls = np.array([0.95e-6, 1.05e-6])
    t = torch.tensor(mask, dtype=torch.float32, device="cuda", requires_grad=True)
    optics = dl.AngularOpticalSystem(
        wf_npixels=n, diameter=1.0,
        layers=[("t", dl.layers.TransmissiveLayer(t, normalise=True)),
                ("opd", dl.layers.AberratedLayer(opd=np.tensordot(basis, c, axes=(0, 0))))],
        psf_npixels=n_out, psf_pixel_scale=float(units.rad2arcsec(p_out)))
    psf = optics.propagate(wls)
    G = np.random.default_rng(3).standard_normal((n_out, n_out))
    loss = torch.sum(psf * torch.as_tensor(G, dtype=torch.float32, device="cuda"))
    loss.backward()
    g = t.grad.double().cpu().numpy()
    V = np.random.default_rng(4).standard_normal((n, n))
    w = np.full(2, 0.5)

    def L(A):
        return np.sum(G * orc.poly_psf(A, basis, c, wls, w, 1.0, n_out, p_out))
    h = 1e-6
    fd = (L(mask + h * V) - L(mask - h * V)) / (2 * h)
    got = float(np.sum(g * V))
    err = abs(got - fd) / abs(fd)
    print(f"  got {got:.6e} fd {fd:.6e} err {err:.2e}")
    assert err < 1e-4
