"""GPU parity of DistortedCoords (coordinates.py:318-357) wherever a coordinate
transformation is accepted: dynamic apertures (alone, after a CoordTransform
and inside a compound), the aberrated aperture's basis, and PSF / Wavefront
resampling, each against the FP64 oracle on the distorted coordinates
(oracle.distort_coords, pinned by closed forms in test_oracle).

Bounds as for the undistorted forms: transmissions 1e-6 absolute, bases 1e-5
of the peak, resampling 2e-6 of the image peak.
"""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

D2 = np.array([[0.02, -0.01, 0.3, 0.1, -0.2], [0.01, 0.03, -0.1, 0.25, 0.15]])
TF = dict(translation=[0.031, -0.017], rotation=0.3, compression=[1.1, 0.9], shear=[0.05, -0.02])


@pytest.fixture(scope="module")
def dl():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dlux_amd
    return dlux_amd


def _np(t):
    return t.detach().double().cpu().numpy()


def _distort(c, D=D2, order=2):
    return orc.coord_transform(c, distortion=D, order=order)


def test_apertures_with_distortion(dl):
    from dlux_amd import layers as L
    print("[distortion] apertures", flush=True)
    n, dx = 128, 1.0 / 128
    grid = orc.pixel_coords(n, 1.0)
    # alone
    ap = L.CircularAperture(0.35, dl.DistortedCoords(2, D2), softening=2.0)
    ref = orc.soft_aperture("circle", _distort(grid), dx, softening=2.0, radius=0.35)
    assert np.max(np.abs(_np(ap._grid_transmission(n, dx)) - ref)) < 1e-6
    # a first-order distortion on a hexagon
    D1 = np.array([[0.1, -0.05], [0.02, 0.08]])
    ap = L.RegPolyAperture(6, 0.4, dl.DistortedCoords(1, D1))
    ref = orc.soft_aperture("reg_polygon", _distort(grid, D1, 1), dx, nsides=6, rmax=0.4)
    assert np.max(np.abs(_np(ap._grid_transmission(n, dx)) - ref)) < 1e-6
    # after an enclosing CoordTransform, inside a compound with a plain member
    comp = L.CompoundAperture(
        [L.CircularAperture(0.45), L.SquareAperture(0.2, dl.DistortedCoords(2, D2),
                                                    occulting=True)],
        transformation=dl.CoordTransform(**TF))
    inner = orc.coord_transform(grid, **TF)
    ref = orc.soft_aperture("circle", inner, dx, radius=0.45) * \
        orc.soft_aperture("square", _distort(inner), dx, occulting=True, width=0.2)
    assert np.max(np.abs(_np(comp._grid_transmission(n, dx)) - ref)) < 1e-6
    # the cached transmission follows a new distortion
    ap = L.CircularAperture(0.35, dl.DistortedCoords(2, D2))
    ap._grid_transmission(n, dx)
    ap2 = ap.set(transformation=dl.DistortedCoords(2, -D2))
    ref = orc.soft_aperture("circle", _distort(grid, -D2), dx, radius=0.35)
    assert np.max(np.abs(_np(ap2._grid_transmission(n, dx)) - ref)) < 1e-6


def test_aberrated_aperture_basis_with_distortion(dl):
    from dlux_amd import layers as L
    print("[distortion] aperture basis", flush=True)
    n, dx = 128, 1.0 / 128
    js = [1, 2, 3, 4, 5, 6, 7, 8, 11]
    for ap in (L.CircularAperture(0.4, dl.DistortedCoords(2, D2)),
               L.RegPolyAperture(6, 0.45, dl.DistortedCoords(2, D2))):
        got = _np(L.AberratedAperture(ap, js)._grid_basis(n, dx))
        coords = _distort(orc.pixel_coords(n, 1.0)) / float(ap.extent)
        if ap.nsides == 0:
            ref = np.array([orc.zernike(j, coords) for j in js])
        else:
            ref = np.array([orc.polike(ap.nsides, j, coords) for j in js])
        ref = np.nan_to_num(ref)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"  nsides {ap.nsides} err {err:.2e}")
        assert err < 1e-5


def test_resampling_with_distortion(dl):
    from dlux_amd import layers
    print("[distortion] resampling", flush=True)
    rng = np.random.default_rng(11)
    n = 32
    data = rng.uniform(0, 1, (n, n)).astype(np.float32)
    tf = dl.DistortedCoords(2, D2)
    psf = dl.PSF(torch.as_tensor(data, device="cuda"), 0.05)
    xs = orc.spec_xs(n, 0.05)
    q = _distort(np.array(np.meshgrid(xs, xs, indexing="xy")))
    ref = orc.interp_grid(data, n, 0.05, 0.0, q[0], q[1], "linear", 0.0)
    err = float(np.max(np.abs(_np(psf.interpolate(tf).data) - ref)) / np.max(ref))
    assert err < 2e-6
    err = float(np.max(np.abs(_np(layers.ApplyInterpolation(tf)(psf).data) - ref)) / np.max(ref))
    assert err < 2e-6
    wf = dl.Wavefront(1e-6, n, 1.0)
    ph = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))).astype(np.complex64)
    wf = wf.set(phasor=torch.as_tensor(ph, device="cuda"))
    xs = orc.spec_xs(n, 1.0 / n)
    q = _distort(np.array(np.meshgrid(xs, xs, indexing="xy")))
    ref = orc.interp_grid(ph.real, n, 1 / n, 0.0, q[0], q[1], "linear", 0.0) + \
        1j * orc.interp_grid(ph.imag, n, 1 / n, 0.0, q[0], q[1], "linear", 0.0)
    got = wf.interpolate(tf).phasor.cpu().numpy()
    assert float(np.max(np.abs(got - ref)) / np.max(np.abs(ref))) < 2e-6


def test_size_gradient_under_distortion(dl):
    """A tracked radius under a DistortedCoords: the transmission's gradient
    (dlx_soft_apertures_vjp on the coordinate planes, run in the backward
    after the forward's planes went out of scope) against central differences
    of the oracle; bound 1e-5 relative.  A tracked distortion raises."""
    from dlux_amd import layers as L
    print("[distortion] size gradient", flush=True)
    n, dx = 128, 1.0 / 128
    coords = _distort(orc.pixel_coords(n, 1.0))
    r = torch.tensor(0.35, dtype=torch.float64, requires_grad=True)
    ap = L.CircularAperture(r, dl.DistortedCoords(2, D2), softening=2.0)
    dout = np.random.default_rng(4).normal(size=(n, n))
    t = ap._grid_transmission(n, dx)
    torch.cuda.empty_cache()
    scratch = torch.full((4, n, n), 7.0, dtype=torch.float64, device="cuda")
    (t.double() * torch.as_tensor(dout, device="cuda")).sum().backward()
    del scratch
    h = 1e-7
    ref = np.sum(dout * (orc.soft_aperture("circle", coords, dx, softening=2.0, radius=0.35 + h)
                         - orc.soft_aperture("circle", coords, dx, softening=2.0,
                                             radius=0.35 - h))) / (2 * h)
    err = abs(float(r.grad) - ref) / abs(ref)
    print(f"  grad {float(r.grad):.6e} ref {ref:.6e} err {err:.2e}")
    assert err < 1e-5
    dd = torch.tensor(D2, requires_grad=True)
    with pytest.raises(NotImplementedError):
        L.CircularAperture(0.35, dl.DistortedCoords(2, dd))._grid_transmission(n, dx)
