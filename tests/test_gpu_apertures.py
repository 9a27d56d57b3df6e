"""GPU parity of the dynamic apertures (dlx_soft_apertures, dlx_aperture_basis)
and of optical systems built from them, against the FP64 oracle.

Bounds: transmissions 1e-6 absolute (values in [0, 1], FP64 in the kernel,
float32 out); bases 1e-5 of the peak (float32 out); PSFs 1e-5 peak-normalised
(the PSF bound of the hot path); coefficient gradients 1e-4 relative to the
largest (the FP32 form's gradient bound).
"""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dl():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dlux_amd
    return dlux_amd


def _np(t):
    return t.detach().double().cpu().numpy()


TF = dict(translation=[0.031, -0.017], rotation=0.3, compression=[1.1, 0.9], shear=[0.05, -0.02])


def _cases(dl, L):
    tf = dl.CoordTransform(**TF)
    return [
        ("circle", L.CircularAperture(0.35, tf, softening=2.0), dict(radius=0.35), False),
        ("circle", L.CircularAperture(0.12, occulting=True, softening=1.0), dict(radius=0.12), True),
        ("square", L.SquareAperture(0.5, tf), dict(width=0.5), False),
        ("rectangle", L.RectangularAperture(0.3, 0.6, tf, softening=3.0),
         dict(width=0.6, height=0.3), False),
        ("reg_polygon", L.RegPolyAperture(6, 0.4, tf), dict(nsides=6, rmax=0.4), False),
        ("reg_polygon", L.RegPolyAperture(5, 0.4), dict(nsides=5, rmax=0.4), False),
        ("spider", L.Spider(0.02, [0.0, 120.0, 240.0], tf), dict(width=0.02, angles=[0, 120, 240]),
         True),
    ]


@pytest.mark.parametrize("case", range(7))
def test_soft_apertures_vs_oracle(dl, case):
    from dlux_amd import layers as L
    kind, ap, shape, occ = _cases(dl, L)[case]
    print(f"[aperture] {kind} occulting={occ}", flush=True)
    n, dx = 128, 1.0 / 128
    coords = orc.pixel_coords(n, 1.0)
    got = _np(ap._grid_transmission(n, dx))
    tfs = [TF] if ap.transformation is not None and ap.transformation.translation is not None \
        else []
    ref = orc.soft_aperture(kind, coords, dx, softening=float(ap.softness), occulting=occ,
                            transforms=tfs, **shape)
    err = np.max(np.abs(got - ref))
    print(f"  max abs err {err:.2e}")
    assert err < 1e-6


def test_composites_vs_oracle(dl):
    """CompoundAperture = product, MultiAperture = sum (of compounds), with the
    enclosing transformations applied before the members' own."""
    print("[aperture] compound / multi", flush=True)
    from dlux_amd import layers as L
    n, dx = 96, 1.0 / 96
    c = orc.pixel_coords(n, 1.0)
    outer = dict(translation=[0.02, 0.01], rotation=-0.2)
    inner = dict(translation=[-0.1, 0.05])
    comp = L.CompoundAperture([L.CircularAperture(0.4),
                               L.CircularAperture(0.1, occulting=True),
                               L.Spider(0.02, [45.0, 225.0])],
                              transformation=dl.CoordTransform(**outer))
    got = _np(comp._grid_transmission(n, dx))
    ref = (orc.soft_aperture("circle", c, dx, transforms=[outer], radius=0.4)
           * orc.soft_aperture("circle", c, dx, occulting=True, transforms=[outer], radius=0.1)
           * orc.soft_aperture("spider", c, dx, occulting=True, transforms=[outer], width=0.02,
                               angles=[45.0, 225.0]))
    assert np.max(np.abs(got - ref)) < 1e-6
    holes = L.MultiAperture([
        L.CompoundAperture([L.CircularAperture(0.08), L.Spider(0.01, [90.0])],
                           transformation=dl.CoordTransform(**inner)),
        L.CircularAperture(0.05, dl.CoordTransform(translation=[0.2, 0.2])),
    ], transformation=dl.CoordTransform(**outer))
    got = _np(holes._grid_transmission(n, dx))
    ref = (orc.soft_aperture("circle", c, dx, transforms=[outer, inner], radius=0.08)
           * orc.soft_aperture("spider", c, dx, occulting=True, transforms=[outer, inner],
                               width=0.01, angles=[90.0])
           + orc.soft_aperture("circle", c, dx,
                               transforms=[outer, dict(translation=[0.2, 0.2])], radius=0.05))
    assert np.max(np.abs(got - ref)) < 1e-6
    # an aperture larger than the grid transmits everywhere (test_apertures.py:137-140)
    wf = dl.Wavefront(1e-6, 16, 1.0)
    for big in (L.CircularAperture(2.4), L.SquareAperture(4.8)):
        assert torch.all(big.transmission(wf.coordinates(), wf.pixel_scale) == 1)


@pytest.mark.parametrize("which", ["circle", "hexagon", "square"])
def test_aperture_basis_vs_oracle(dl, which):
    """AberratedAperture basis: Zernike (circle) or polike (n-gon) on the
    aperture's transformed coordinates divided by its extent."""
    print(f"[aperture basis] {which}", flush=True)
    from dlux_amd import layers as L
    n, dx = 128, 1.0 / 128
    tf = dl.CoordTransform(**TF)
    ap = {"circle": L.CircularAperture(0.4, tf), "hexagon": L.RegPolyAperture(6, 0.45, tf),
          "square": L.SquareAperture(0.5, tf)}[which]
    js = [1, 2, 3, 4, 5, 6, 7, 8, 11]
    ab = L.AberratedAperture(ap, js)
    got = _np(ab._grid_basis(n, dx))
    coords = orc.coord_transform(orc.pixel_coords(n, 1.0), **TF) / float(ap.extent)
    if ap.nsides == 0:
        ref = np.array([orc.zernike(j, coords) for j in js])
    else:
        ref = np.array([orc.polike(ap.nsides, j, coords) for j in js])
    ref = np.nan_to_num(ref)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"  err {err:.2e}")
    assert err < 1e-5


def test_dynamic_aperture_system_psf_and_gradient(dl):
    """An AngularOpticalSystem built from an aberrated, transformed, normalising
    circular aperture and a spider lowers into the fused PSF path: the PSF and
    the coefficient gradient match the oracle chain on the oracle's own
    transmission and basis."""
    print("[aperture system]", flush=True)
    from dlux_amd import layers as L
    from dlux_amd.utils import units
    n, n_out, p = 128, 64, 2.5e-7
    tf = dict(translation=[0.01, -0.02], rotation=0.1)
    js = list(range(2, 12))
    coeffs = np.random.default_rng(31).normal(0, 30e-9, len(js))
    ab = L.AberratedAperture(L.CircularAperture(0.5, dl.CoordTransform(**tf), normalise=True),
                             js, torch.tensor(coeffs, dtype=torch.float32, device="cuda",
                                              requires_grad=True))
    optics = dl.AngularOpticalSystem(
        wf_npixels=n, diameter=1.0,
        layers=[("pupil", ab), ("spider", L.Spider(0.01, [0.0, 90.0, 180.0, 270.0]))],
        psf_npixels=n_out, psf_pixel_scale=float(units.rad2arcsec(p)))
    wls = np.array([0.95e-6, 1.05e-6])
    psf = optics.propagate(wls)
    c = orc.pixel_coords(n, 1.0)
    dx = 1.0 / n
    basis = np.array([orc.zernike(j, orc.coord_transform(c, **tf) / 0.5) for j in js])
    # the aperture normalises the power of the circle alone; the spider comes after
    circ = orc.soft_aperture("circle", c, dx, transforms=[tf], radius=0.5)
    spider = orc.soft_aperture("spider", c, dx, occulting=True, width=0.01,
                               angles=[0.0, 90.0, 180.0, 270.0])
    norm = 1.0 / np.sqrt(np.sum((circ / n ** 2) ** 2))
    amp = circ * norm * spider
    w = np.full(len(wls), 1.0 / len(wls))
    ref = orc.poly_psf(amp, basis, coeffs, wls, w, 1.0, n_out, p, normalise_power=False)
    err = float(np.max(np.abs(_np(psf) - ref)) / np.max(ref))
    print(f"  psf err {err:.2e}")
    assert err < 1e-5
    # a target from other coefficients, so the loss is away from its minimum
    data = 1e5 * orc.poly_psf(amp, basis, coeffs * 0.5 + 5e-9, wls, w, 1.0, n_out, p,
                              normalise_power=False)
    loss = 0.5 * torch.sum((psf * 1e5 - torch.as_tensor(data, device="cuda")) ** 2
                           / torch.clamp(torch.as_tensor(data, device="cuda"), min=1.0))
    loss.backward()
    G = 1e5 * (1e5 * ref - data) / np.maximum(data, 1.0)
    g_opd, _ = orc.poly_psf_vjp(amp, basis, coeffs, wls, w, 1.0, n_out, p, G,
                                normalise_power=False)
    gref = np.tensordot(basis, g_opd, axes=([1, 2], [0, 1]))
    g = _np(ab.coefficients.grad)
    gerr = float(np.max(np.abs(g - gref)) / np.max(np.abs(gref)))
    print(f"  grad err {gerr:.2e}")
    assert gerr < 1e-4


def test_compound_normalises_after_aberrations(dl):
    """tests/layers/test_apertures.py:230-236 on the device path."""
    from dlux_amd import layers as L
    wf = dl.Wavefront(1e-6, 16, 1.0)
    ab = L.AberratedAperture(L.CircularAperture(2.4), [2], coefficients=np.array([1e-7]))
    out = L.CompoundAperture([ab], normalise=True)(wf)
    assert abs(float(out.power) - 1.0) < 1e-5
    assert not torch.allclose(out.phase, torch.zeros_like(out.phase))
    plain = L.CircularAperture(0.1)(wf)
    amp = L.AberratedAperture(L.CircularAperture(0.1), [1, 2, 3, 4], effect="amplitude")(wf)
    assert torch.allclose(amp.amplitude, plain.amplitude)
