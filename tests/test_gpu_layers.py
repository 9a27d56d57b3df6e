"""GPU parity of layers and systems that the other suites do not reach:
BasisOptic (layers/optics.py:99-175), Normalise (layers/optical_layers.py:
362-371), CartesianOpticalSystem.to_focus (optical_systems.py:754-775) and the
propagate_MFT wrapper (wavefronts.py:774-805), each through the fused path
(``optics.propagate``) and the layer-by-layer path (``return_wf=True``),
against the FP64 oracle.  Error metric: peak-normalised max |x - ref| / max
|ref| <= 1e-5 for PSFs and 2e-6 for FP32 fields.
"""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _err(x, ref):
    """Peak-normalised max difference; complex tensors keep their imaginary part."""
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu()
        x = x.to(torch.complex128 if x.is_complex() else torch.float64).numpy()
    x = np.asarray(x)
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def dl():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dlux_amd
    return dlux_amd


def _case(n, nolls, seed, std):
    mask, basis = orc.circular_aperture(n, 1.0, oversample=3, secondary_diameter=0.15,
                                        spider_width=0.02, spider_angles=[30, 150, 270],
                                        zernike_nolls=nolls)
    c = np.random.default_rng(seed).normal(0, std, basis.shape[0])
    return mask, basis, c


def _oracle_psf(fields_fn, wls, w, n, n_out, p_out):
    """sum_l w_l |MFT(P_l)|^2 for pupil fields P_l = fields_fn(wl)."""
    out = 0.0
    for wl, wt in zip(wls, w):
        E = orc.mft(fields_fn(wl), wl, 1.0 / n, n_out, p_out)
        out = out + wt * np.abs(E) ** 2
    return out


@pytest.mark.parametrize("effect", ["opd", "phase", "amplitude"])
@pytest.mark.parametrize("normalise", [True, False])
def test_basis_optic_vs_oracle(dl, effect, normalise):
    """BasisOptic: transmission, then the basis applied as OPD / phase /
    amplitude (E (1 + Z.c)), then the optional power normalisation."""
    print(f"[basis-optic] effect={effect} normalise={normalise}", flush=True)
    n, n_out, p_out = 96, 48, 2.8e-7
    std = {"opd": 40e-9, "phase": 0.3, "amplitude": 0.05}[effect]
    mask, basis, c = _case(n, range(2, 12), 31, std)
    basis = np.nan_to_num(basis)
    wls = np.linspace(0.85e-6, 1.15e-6, 4)
    w = np.array([0.1, 0.4, 0.3, 0.2])
    optic = dl.BasisOptic(basis, transmission=mask, coefficients=c, normalise=normalise,
                          effect=effect)
    optics = dl.AngularOpticalSystem(n, 1.0, [("optic", optic)], n_out,
                                     float(dl.utils.rad2arcsec(p_out)))
    z = orc.eval_basis(basis, c)

    def field(wl):
        P = orc.wavefront_init(n) * mask
        if effect == "opd":
            P = orc.add_opd(P, wl, z)
        elif effect == "phase":
            P = P * np.exp(1j * z)
        else:
            P = P * (1 + z)
        return orc.normalise(P) if normalise else P

    ref = _oracle_psf(field, wls, w, n, n_out, p_out)
    for prec in ("fp32", "bf16x3"):
        with dl.precision(prec):
            fused = optics.propagate(wls, weights=w)
            wf = optics.propagate(wls, weights=w, return_wf=True)
        e1, e2 = _err(fused, ref), _err(wf.psf.sum(0), ref)
        print(f"  {prec}: fused {e1:.2e} layer-by-layer {e2:.2e}")
        assert e1 < 1e-5 and e2 < 1e-5


def test_normalise_layer_vs_oracle(dl):
    """Normalise between a scaled transmission and the aberrations: the PSF is
    the one of a unit-power pupil whatever the transmission's scale."""
    print("[normalise-layer]", flush=True)
    n, n_out, p_out = 80, 40, 3e-7
    mask, basis, c = _case(n, range(2, 9), 5, 50e-9)
    basis = np.nan_to_num(basis)
    wls = [0.9e-6, 1.1e-6]
    w = np.array([0.7, 0.3])
    layers = [("t", dl.TransmissiveLayer(3.7 * mask)), ("n", dl.Normalise()),
              ("z", dl.BasisLayer(basis, c))]
    optics = dl.AngularOpticalSystem(n, 1.0, layers, n_out, float(dl.utils.rad2arcsec(p_out)))

    def field(wl):
        P = orc.normalise(orc.wavefront_init(n) * 3.7 * mask)
        return orc.add_opd(P, wl, orc.eval_basis(basis, c))

    ref = _oracle_psf(field, wls, w, n, n_out, p_out)
    fused = optics.propagate(wls, weights=w)
    wf = optics.propagate(wls, weights=w, return_wf=True)
    assert _err(fused, ref) < 1e-5
    assert _err(wf.psf.sum(0), ref) < 1e-5
    # the same system without the Normalise layer is 3.7^2 x (1/sum mask^2 N^-4) brighter
    plain = dl.AngularOpticalSystem(n, 1.0, [layers[0], layers[2]], n_out,
                                    float(dl.utils.rad2arcsec(p_out)))
    ratio = float(plain.propagate(wls, weights=w).double().sum() / fused.double().sum())
    expect = 3.7 ** 2 * np.sum(mask ** 2) / n ** 4
    assert np.isclose(ratio, expect, rtol=1e-5)
    # and a Wavefront run through the layer has unit power
    wf0 = dl.Wavefront(1e-6, n, 1.0) * (3.7 * mask)
    assert abs(float(dl.Normalise()(wf0).power) - 1.0) < 1e-6


@pytest.mark.parametrize("oversample", [1, 2])
def test_cartesian_to_focus_passes_microns_as_radians(dl, oversample):
    """CartesianOpticalSystem.to_focus hands the MFT p = 1e-6 * psf_pixel_scale /
    oversample and no focal length (optical_systems.py:772-775), so the MFT
    reads p as radians: the image equals an angular system whose pixel is
    1e-6 * psf_pixel_scale rad, not the physical focal-plane sampling
    (p / focal_length rad)."""
    print(f"[cartesian] oversample={oversample}", flush=True)
    n, n_out = 64, 32
    mask, basis, c = _case(n, range(2, 8), 8, 30e-9)
    basis = np.nan_to_num(basis)
    focal_length, pps = 4.0, 0.3          # metres, microns per pixel
    wls = [0.95e-6, 1.05e-6]
    layers = [("t", dl.TransmissiveLayer(mask, normalise=True)), ("z", dl.BasisLayer(basis, c))]
    cart = dl.CartesianOpticalSystem(n, 1.0, layers, focal_length, n_out, pps,
                                     oversample=oversample)
    p = 1e-6 * pps / oversample

    def field(wl):
        P = orc.normalise(orc.wavefront_init(n) * mask)
        return orc.add_opd(P, wl, orc.eval_basis(basis, c))

    w = [0.5, 0.5]
    ref = _oracle_psf(field, wls, w, n, n_out * oversample, p)
    fused = cart.propagate(wls)
    wf = cart.propagate(wls, return_wf=True)
    assert tuple(fused.shape) == (n_out * oversample,) * 2
    assert _err(fused, ref) < 1e-5
    assert _err(wf.psf.sum(0), ref) < 1e-5
    # the focal-length form of the same sampling is a different image
    physical = 0.0
    for wl, wt in zip(wls, w):
        E = orc.mft(field(wl), wl, 1.0 / n, n_out * oversample, p, focal_length=focal_length)
        physical = physical + wt * np.abs(E) ** 2
    assert _err(fused, physical) > 1e-2


def test_propagate_mft_with_spec(dl):
    """Wavefront.propagate_MFT(spec_out, focal_length, inverse) (wavefronts.py:774-805)
    takes n and d from the spec and, like the reference, does not read its
    centre c: the output equals the unshifted MFT for any c."""
    print("[propagate_MFT]", flush=True)
    n, n_out, p = 48, 24, 2.5e-7
    mask = orc.circular_aperture(n, 1.0, oversample=1)
    rng = np.random.default_rng(3)
    opd = rng.normal(0, 20e-9, (n, n))
    wf = (dl.Wavefront(1e-6, n, 1.0) * mask).add_opd(opd)
    P = orc.add_opd(orc.wavefront_init(n) * mask, 1e-6, opd)
    for c, focal, inverse in ((0.0, None, None), (3.0 * p, None, False), (0.0, 2.0, True)):
        out = wf.propagate_MFT(dl.CoordSpec(n_out, p, c), focal_length=focal, inverse=inverse)
        ref = orc.mft(P, 1e-6, 1.0 / n, n_out, p, focal_length=focal, inverse=bool(inverse))
        assert out.npixels == n_out
        assert np.isclose(float(np.ravel(out.pixel_scale)[0]), p)
        assert _err(out.phasor, ref) < 2e-6
