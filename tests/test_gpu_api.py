"""GPU parity of the dLux-compatible API (systems, layers, FFT, ASM, builders)
against the FP64 oracle.  Error metric: peak-normalised max |x - ref| / max |ref|."""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _err(x, ref):
    x = x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def dl():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dlux_amd
    return dlux_amd


@pytest.mark.parametrize("n,pad,inverse", [(16, 2, False), (16, 2, True), (32, 4, False),
                                           (64, 2, True), (15, 2, False), (12, 3, True),
                                           (256, 8, False)])
def test_fft_vs_oracle(dl, n, pad, inverse):
    print(f"[fft] n={n} pad={pad} inverse={inverse}", flush=True)
    rng = np.random.default_rng(n + pad)
    E = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    out, ps = dl.utils.FFT(E, 1e-6, 1e-3, pad=pad, inverse=inverse)
    ref, ps_ref = orc.fft(E, 1e-6, 1e-3, pad=pad, inverse=inverse)
    assert tuple(out.shape) == ref.shape
    assert np.isclose(ps, ps_ref)
    assert _err(out, ref) < 3e-6


def test_fft_large_sides(dl):
    """The largest row blocks: M = 4096 (four rows, full 160 KiB of LDS) against the
    FP64 oracle, and M = 8192 (one row per workgroup) by the forward-inverse round trip
    and Parseval (sum |X|^2 = sum |x|^2 / 1 for the unitary centred transform)."""
    print("[fft-large]", flush=True)
    rng = np.random.default_rng(41)
    E = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
    out, _ = dl.utils.FFT(E, 1e-6, 1e-3, pad=128)
    ref, _ = orc.fft(E, 1e-6, 1e-3, pad=128)
    assert _err(out, ref) < 3e-6
    x = torch.as_tensor((rng.standard_normal((8192, 8192)) +
                         1j * rng.standard_normal((8192, 8192))).astype(np.complex64)).cuda()
    fwd = dl.runtime.fft2(x, 8192)
    back = dl.runtime.fft2(fwd, 8192, inverse=True)
    assert float(torch.max(torch.abs(back - x))) < 2e-4 * float(torch.max(torch.abs(x)))
    def energy(t):
        return float(torch.sum(t.real.double() ** 2 + t.imag.double() ** 2))
    e_in, e_out = energy(x), energy(fwd)
    assert abs(e_out - e_in) < 1e-5 * e_in


def test_fft_round_trip_reference_fixture(dl):
    """tests/utils/test_propagation.py:74-92: forward then inverse (pad=1) is identity."""
    print("[fft-roundtrip]", flush=True)
    ph = np.ones((32, 32), complex)
    fwd, _ = dl.utils.FFT(ph, 1.0, 0.1, None, pad=1, inverse=False)
    back, _ = dl.utils.FFT(fwd, 1.0, 0.1, None, pad=1, inverse=True)
    assert np.allclose(back.cpu().numpy(), ph, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("n,m,z", [(64, 128, 0.05), (128, 256, 0.001), (32, 64, 0.0)])
def test_asm_vs_oracle(dl, n, m, z):
    print(f"[asm] n={n} m={m} z={z}", flush=True)
    rng = np.random.default_rng(n)
    dx = 4e-6
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y = np.meshgrid(x, x)
    u = np.exp(-(X ** 2 + Y ** 2) / (2 * (n * dx / 8) ** 2)) * np.exp(
        1j * rng.standard_normal((n, n)) * 0.1)
    out = dl.runtime.asm_prop(u, [1e-6], dx, z, m)
    ref = orc.asm_prop(u, dx, 1e-6, z, m)
    assert _err(out, ref) < 1e-5


def test_zernike_basis_vs_oracle(dl):
    print("[zernike]", flush=True)
    js = list(range(1, 38))
    coords_ref = orc.pixel_coords(64, 1.0)
    ref = orc.zernike_basis(js, coords_ref, 1.0)
    out = dl.utils.zernike_basis(js, dl.utils.pixel_coords(64, diameter=1.0), 1.0)
    assert out.shape == ref.shape
    assert float(torch.max(torch.abs(out.cpu().double() - torch.as_tensor(ref)))) < 2e-6


@pytest.mark.parametrize("os_,sec,spw", [(1, None, None), (3, 0.2, 0.05), (5, None, 0.02)])
def test_circular_aperture_vs_oracle(dl, os_, sec, spw):
    print(f"[aperture] os={os_} sec={sec} spider={spw}", flush=True)
    angles = [0, 120, 240] if spw else None
    ref_t, ref_b = orc.circular_aperture(64, 1.0, oversample=os_, secondary_diameter=sec,
                                         spider_width=spw, spider_angles=angles,
                                         zernike_nolls=range(1, 8))
    t, b = dl.utils.circular_aperture(64, 1.0, oversample=os_, secondary_diameter=sec,
                                      spider_width=spw, spider_angles=angles,
                                      zernike_nolls=range(1, 8))
    assert np.max(np.abs(t.cpu().numpy() - ref_t)) < 1e-6
    assert np.max(np.abs(b.cpu().numpy() - ref_b)) < 2e-6


def _system(dl, n, n_out, mask, basis, coeffs, p_out):
    return dl.AngularOpticalSystem(
        wf_npixels=n, diameter=1.0,
        layers=[("aperture", dl.TransmissiveLayer(mask, normalise=True)),
                ("aberrations", dl.BasisLayer(basis, coeffs))],
        psf_npixels=n_out, psf_pixel_scale=dl.utils.rad2arcsec(p_out))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mono_config_fused_vs_oracle(dl, prec):
    """The monochromatic configuration: 256^2 -> 128^2, Noll 2..11 ~ N(0, 20 nm), seed 0."""
    print(f"[mono] prec={prec}", flush=True)
    from dlux_amd import workloads
    cfg = workloads.MONO
    mask_r = orc.circular_aperture(cfg.n_in, 1.0, oversample=1)
    basis_r = orc.zernike_basis(cfg.js, orc.pixel_coords(cfg.n_in, 1.0), 1.0)
    c = cfg.coefficients()
    ref = orc.poly_psf(mask_r, basis_r, c, cfg.wavelengths, cfg.weights, 1.0, cfg.n_out,
                       cfg.p_out)
    optics = workloads.build_system(cfg, c)
    with dl.precision(prec):
        psf = optics.propagate(cfg.wavelengths)
    err = _err(psf, ref)
    print(f"  err {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_fused_equals_layerwise_and_oracle(dl, prec):
    print(f"[fused-vs-layerwise] prec={prec}", flush=True)
    n, n_out = 128, 64
    mask, basis = orc.circular_aperture(n, 1.0, oversample=3, secondary_diameter=0.2,
                                        spider_width=0.03, spider_angles=[45, 135, 225, 315],
                                        zernike_nolls=range(2, 16))
    rng = np.random.default_rng(9)
    c = rng.normal(0, 40e-9, basis.shape[0])
    wls = np.linspace(0.7e-6, 1.3e-6, 5)
    w = np.array([0.1, 0.3, 0.2, 0.25, 0.15])
    p_out = 3e-7
    offset = np.array([2e-6, -1e-6])
    optics = _system(dl, n, n_out, mask, basis, c, p_out)
    with dl.precision(prec):
        fused = optics.propagate(wls, offset=offset, weights=w)
        wf = optics.propagate(wls, offset=offset, weights=w, return_wf=True)
    layerwise = wf.psf.sum(0)
    ref = orc.poly_psf(mask, basis, c, wls, w, 1.0, n_out, p_out, offset=offset)
    assert _err(fused, ref) < 1e-5
    assert _err(layerwise, ref) < 1e-5


def test_layered_system_with_mft_layer(dl):
    print("[layered+MFT]", flush=True)
    n = 64
    mask = orc.circular_aperture(n, 1.0, oversample=1)
    rng = np.random.default_rng(2)
    opd = rng.normal(0, 20e-9, (n, n))
    phase = rng.normal(0, 0.1, (n, n))
    optics = dl.LayeredOpticalSystem(
        n, 1.0, [dl.Optic(mask, opd=opd, phase=phase, normalise=True),
                 dl.Tilt([1e-6, 0.5e-6]), dl.MFT(48, 2e-7)])
    psf = optics.propagate([1e-6, 1.1e-6])
    E = []
    for wl in [1e-6, 1.1e-6]:
        P = orc.wavefront_init(n) * mask
        P = orc.add_opd(P, wl, opd) * np.exp(1j * phase)
        P = orc.normalise(P)
        P = orc.add_opd(P, wl, orc.tilt_opd(n, 1.0 / n, [1e-6, 0.5e-6]))
        E.append(orc.mft(P, wl, 1.0 / n, 48, 2e-7) * np.sqrt(0.5))
    ref = sum(np.abs(e) ** 2 for e in E)
    assert _err(psf, ref) < 1e-5


def test_fft_layer_centered(dl):
    """FFT layer with center=True: ramps + padded FFT + crop (propagators.py:118-142)."""
    print("[fft-layer]", flush=True)
    n = 32
    mask = orc.circular_aperture(n, 1.0, oversample=1)
    wf = dl.Wavefront(1e-6, n, 1.0) * mask
    out = dl.FFT(pad=2, crop=1)(wf)
    ref, ps, c = orc.propagate_fft(orc.wavefront_init(n) * mask, 1e-6, 1.0 / n, pad=2,
                                   spec_c=0.0)
    assert out.npixels == 64
    assert _err(out.phasor, ref) < 3e-6
    assert np.isclose(float(out.pixel_scale), ps)


@pytest.mark.parametrize("centre", [False, True])
def test_propagate_fft_vectorised_equals_per_wavelength(dl, centre):
    """A vectorised wavefront (3 wavelengths) runs one batched padded FFT, with the
    re-centring ramps per wavelength: each slice equals the single-wavelength
    propagation to the bit, with its own pixel scale (the wavefront keeps one
    centre, the first wavelength's, as before)."""
    print(f"[fft-vectorised] centre={centre}", flush=True)
    from dlux_amd.coordinates import CoordSpec
    n, wls = 32, np.array([0.9e-6, 1.0e-6, 1.2e-6])
    rng = np.random.default_rng(5)
    ph = (rng.normal(size=(3, n, n)) + 1j * rng.normal(size=(3, n, n))).astype(np.complex64)
    spec = CoordSpec(c=0.0) if centre else None
    wf = dl.Wavefront.from_phasor(torch.as_tensor(ph).cuda(), wls, diameter=1.0)
    out = wf.propagate_FFT(pad=2, spec_out=spec)
    assert tuple(out.phasor.shape) == (3, 2 * n, 2 * n)
    for i, wl in enumerate(wls):
        one = dl.Wavefront.from_phasor(torch.as_tensor(ph[i]).cuda(), wl, diameter=1.0)
        ref = one.propagate_FFT(pad=2, spec_out=spec)
        assert torch.equal(out.phasor[i], ref.phasor), i
        assert np.isclose(np.ravel(out.pixel_scale)[i], float(ref.pixel_scale), rtol=0, atol=0)
        if centre or i == 0:   # one centre per wavefront: the first wavelength's
            assert np.allclose(out.center, ref.center)


def test_point_sources_sum(dl):
    print("[point-sources]", flush=True)
    from dlux_amd import workloads
    cfg = workloads.MONO
    optics = workloads.build_system(cfg)
    pos = np.array([[0.0, 0.0], [1e-6, -2e-6]])
    flux = np.array([1.0, 3.0])
    src = dl.PointSources(wavelengths=[0.9e-6, 1.0e-6], position=pos, flux=flux)
    psf = src.model(optics)
    mask_r = orc.circular_aperture(cfg.n_in, 1.0, oversample=1)
    basis_r = orc.zernike_basis(cfg.js, orc.pixel_coords(cfg.n_in, 1.0), 1.0)
    ref = orc.point_sources_psf(mask_r, basis_r, cfg.coefficients(), [0.9e-6, 1.0e-6],
                                [1, 1], pos, flux, 1.0, cfg.n_out, cfg.p_out)
    assert _err(psf, ref) < 1e-5


def test_point_sources_batched_bf16x3(dl):
    """PointSources.model under the split-BF16x3 form: all (source, wavelength)
    instances in one batched call, positions as MFT shifts."""
    print("[point-sources-batched]", flush=True)
    from dlux_amd import runtime, workloads
    cfg = workloads.MONO
    optics = workloads.build_system(cfg)
    rng = np.random.default_rng(21)
    pos = rng.uniform(-4e-6, 4e-6, (7, 2))
    flux = np.exp(rng.uniform(0, 4, 7))
    wls = [0.9e-6, 1.0e-6, 1.1e-6]
    src = dl.PointSources(wavelengths=wls, position=pos, flux=flux)
    with runtime.precision("bf16x3"):
        assert optics._sources_psf([(wls, pos, np.ones((7, 3)), 0)]) is not None
        psf = src.model(optics)
    mask_r = orc.circular_aperture(cfg.n_in, 1.0, oversample=1)
    basis_r = orc.zernike_basis(cfg.js, orc.pixel_coords(cfg.n_in, 1.0), 1.0)
    ref = orc.point_sources_psf(mask_r, basis_r, cfg.coefficients(), wls, [1, 1, 1], pos,
                                flux, 1.0, cfg.n_out, cfg.p_out)
    assert _err(psf, ref) < 1e-5
