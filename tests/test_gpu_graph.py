"""HIP graph capture of the fused PSF and of its VJP: a captured call replays
with new inputs copied into its static buffers, with eager work between the
replays, and every replay equals the eager result (same kernels, same inputs;
tolerance 1e-6 peak-normalised)."""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def rt():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dlux_amd import runtime
    return runtime


def _peak_err(a, b):
    a = a.double().cpu().numpy()
    b = b.double().cpu().numpy()
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_psf_poly_and_vjp_graph_replays(rt, prec):
    print(f"[graph] prec={prec}", flush=True)
    n, n_out, nwl = 96, 48, 5
    mask, basis = orc.circular_aperture(n, 1.0, oversample=1, zernike_nolls=range(2, 9))
    wls = np.linspace(0.9e-6, 1.1e-6, nwl)
    w = np.ones(nwl) / nwl
    scale = 1.0 / np.sqrt(np.sum(mask ** 2))
    amp = torch.as_tensor(mask, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(5)
    opds = [torch.as_tensor(np.tensordot(basis, rng.normal(0, 40e-9, basis.shape[0]),
                                         axes=(0, 0)), dtype=torch.float32, device="cuda")
            for _ in range(4)]
    G = torch.as_tensor(rng.standard_normal((n_out, n_out)), dtype=torch.float32,
                        device="cuda")
    kw = dict(amp=amp, scale=scale, precision=prec)

    def eager(o):
        with torch.no_grad():
            psf = rt.psf_poly(n, n_out, wls, w, 1.0 / n, 2.5e-7, opd=o, **kw)
            g_opd, _ = rt.psf_poly_vjp(n, n_out, wls, w, 1.0 / n, 2.5e-7, G, opd=o, **kw)
        return psf.clone(), g_opd.clone()

    refs = [eager(o) for o in opds]
    static_opd = opds[0].clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        eager(static_opd)                       # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.no_grad(), torch.cuda.graph(graph):
        psf_s = rt.psf_poly(n, n_out, wls, w, 1.0 / n, 2.5e-7, opd=static_opd, **kw)
        g_s, _ = rt.psf_poly_vjp(n, n_out, wls, w, 1.0 / n, 2.5e-7, G, opd=static_opd, **kw)
    for rep in range(2):
        for i, o in enumerate(opds):
            static_opd.copy_(o)
            graph.replay()
            torch.cuda.synchronize()
            e_psf = _peak_err(psf_s, refs[i][0])
            e_g = _peak_err(g_s, refs[i][1])
            assert e_psf <= 1e-6 and e_g <= 1e-6, (rep, i, e_psf, e_g)
            eager(opds[(i + 1) % len(opds)])    # eager work between replays
            torch.cuda.synchronize()


def test_graphed_gradient_step_matches_eager(rt):
    """runtime.graph_step over forward PSF + Gaussian loss + backward to the
    Zernike coefficients: every replay, with new coefficients copied in and
    eager work between replays, gives the eager loss and gradient."""
    print("[graph] gradient step", flush=True)
    import dlux_amd as dl
    n, n_out, nwl = 128, 64, 6
    mask, basis = orc.circular_aperture(n, 1.0, oversample=1, zernike_nolls=range(2, 12))
    wls = np.linspace(0.9e-6, 1.1e-6, nwl)
    rng = np.random.default_rng(8)
    sets = [rng.normal(0, 30e-9, basis.shape[0]) for _ in range(4)]
    coeffs = torch.tensor(sets[0], dtype=torch.float32, device="cuda", requires_grad=True)
    optics = dl.AngularOpticalSystem(
        wf_npixels=n, diameter=1.0,
        layers=[("aperture", dl.TransmissiveLayer(mask, normalise=True)),
                ("aberrations", dl.BasisLayer(basis, coeffs))],
        psf_npixels=n_out, psf_pixel_scale=dl.utils.rad2arcsec(2.5e-7))
    data = torch.as_tensor(rng.poisson(1e4 * orc.poly_psf(mask, basis, sets[1], wls, None, 1.0,
                                                          n_out, 2.5e-7)),
                           dtype=torch.float32, device="cuda")
    var = torch.clamp(data, min=1.0)

    def step():
        psf = optics.propagate(wls)
        loss = 0.5 * torch.sum((1e4 * psf - data) ** 2 / var)
        if coeffs.grad is not None:
            coeffs.grad.zero_()
        loss.backward()
        return loss.detach()

    def eager(c):
        with torch.no_grad():
            coeffs.copy_(torch.as_tensor(c, dtype=torch.float32))
        loss = step()
        torch.cuda.synchronize()
        return float(loss), coeffs.grad.clone()

    with rt.precision("bf16x3"):
        refs = [eager(c) for c in sets]
        g = rt.graph_step(step)
        for rep in range(2):
            for i, c in enumerate(sets):
                with torch.no_grad():
                    coeffs.copy_(torch.as_tensor(c, dtype=torch.float32))
                loss = g()
                torch.cuda.synchronize()
                assert abs(float(loss) - refs[i][0]) <= 1e-6 * abs(refs[i][0]), (rep, i)
                err = _peak_err(coeffs.grad, refs[i][1])
                assert err <= 1e-6, (rep, i, err)
                eager(sets[(i + 1) % len(sets)])    # eager work between replays


def test_graphed_telescope_model_matches_eager(rt):
    """runtime.graph_step over Telescope.model (point source, 8 wavelengths,
    the detector fused into the wavelength sum): every replay, with a new
    response map copied in and eager calls between replays, equals the eager
    image to the bit."""
    print("[graph] telescope model", flush=True)
    import dlux_amd as dl
    from dlux_amd import layers
    n, n_out, nwl = 128, 64, 8
    mask, basis = orc.circular_aperture(n, 1.0, oversample=1, zernike_nolls=range(2, 8))
    rng = np.random.default_rng(21)
    coeffs = torch.tensor(rng.normal(0, 30e-9, basis.shape[0]), dtype=torch.float32,
                          device="cuda")
    optics = dl.AngularOpticalSystem(
        wf_npixels=n, diameter=1.0,
        layers=[("aperture", dl.TransmissiveLayer(mask, normalise=True)),
                ("aberrations", dl.BasisLayer(basis, coeffs))],
        psf_npixels=n_out, psf_pixel_scale=dl.utils.rad2arcsec(2.5e-7))
    resps = [torch.as_tensor(rng.uniform(0.9, 1.1, (n_out, n_out)), dtype=torch.float32,
                             device="cuda") for _ in range(3)]
    resp = resps[0].clone()
    det = dl.LayeredDetector([layers.ApplyJitter(0.5, 5, 3), layers.ApplyPixelResponse(resp),
                              layers.Downsample(2), layers.AddConstant(1e-6)])
    wls = np.linspace(0.9e-6, 1.1e-6, nwl)
    tel = dl.Telescope(optics, dl.PointSource(wls), det)
    with rt.precision("bf16x3"):
        refs = []
        for r in resps:
            resp.copy_(r)
            refs.append(tel.model().clone())
        torch.cuda.synchronize()
        g = rt.graph_step(tel.model)
        for rep in range(2):
            for k, r in enumerate(resps):
                resp.copy_(r)
                out = g()
                torch.cuda.synchronize()
                assert torch.equal(out, refs[k]), (rep, k)
                tel.model()   # eager work between replays
                torch.cuda.synchronize()
