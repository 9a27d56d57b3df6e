"""GPU parity of the Zernike parameterisations (polynomials.py) against the
FP64 oracle: DynamicZernike, ZernikeBasis and DynamicZernikeBasis, on
explicit coordinates (dlx_zernike_basis; the polike wedge radius on the
device) and on a wavefront's pixel grid (dlx_aperture_basis).

Bound: 1e-5 of the basis peak (float32 output; the coordinates themselves
are float32 on explicit grids); the n-sided bases 1e-4, whose 1/r_alpha
scaling is evaluated from float32 coordinates and so moves the disk edge by
an ulp, compared away from the edge.
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


def _peak(got, ref):
    return float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))


def test_dynamic_zernike_vs_oracle(dl):
    """tests/test_polynomials.py:17-44: shape, the diameter contract and the
    polygonal form, on the oracle's values."""
    print("[polynomials] DynamicZernike", flush=True)
    c8 = orc.pixel_coords(8, 2.0)
    z = dl.DynamicZernike(4)
    assert tuple(z.calculate(c8).shape) == (8, 8)
    c = orc.pixel_coords(64, 1.1)
    got = _np(dl.DynamicZernike(7).calculate(c, diameter=1.1))
    assert _peak(got, orc.zernike(7, c, diameter=1.1)) < 1e-5
    c = orc.pixel_coords(96, 2.0)
    ref = orc.polike(6, 4, c)
    got = _np(dl.DynamicZernike(4).calculate(c, nsides=6))
    inner = np.hypot(*c) < 0.8
    assert np.max(np.abs(got - ref)[inner]) / np.max(np.abs(ref)) < 1e-4


def test_zernike_basis_radial_orders_and_diameter(dl):
    """tests/test_polynomials.py:47-95."""
    print("[polynomials] ZernikeBasis", flush=True)
    c = orc.pixel_coords(32, 2.0)
    coefficients = np.arange(3.0)
    by_index = dl.ZernikeBasis(c, js=[4, 5, 6], coefficients=coefficients)
    by_order = dl.ZernikeBasis(c, radial_orders=[2], coefficients=coefficients)
    assert torch.equal(by_index.basis, by_order.basis)
    assert torch.allclose(by_index.evaluate(), by_order.evaluate())
    ref = np.tensordot(coefficients, orc.zernike_basis([4, 5, 6], c), 1)
    assert _peak(_np(by_index.evaluate()), ref) < 1e-5
    with pytest.raises(ValueError, match="Coefficient shape"):
        dl.ZernikeBasis(c, js=[4, 5], coefficients=np.ones(3))
    c = orc.pixel_coords(32, 1.1)
    b = dl.ZernikeBasis(c, js=[4, 5], diameter=1.1)
    assert _peak(_np(b.basis), orc.zernike_basis([4, 5], c, diameter=1.1)) < 1e-5


def test_dynamic_zernike_basis_from_wavefront_and_coordinates(dl):
    """tests/test_polynomials.py:98-150: the wavefront infers the diameter (its
    pixel grid generated in the kernel) and equals the explicit basis; an
    explicit diameter is honoured; the n-sided form on transformed (not
    grid) coordinates; gradients reach the coefficients."""
    print("[polynomials] DynamicZernikeBasis", flush=True)
    diameter = 1.1
    wf = dl.Wavefront(700e-9, 64, diameter=diameter)
    coefficients = np.arange(3.0)
    dyn = dl.DynamicZernikeBasis(radial_orders=[2], coefficients=coefficients)
    explicit = dl.ZernikeBasis(wf.coordinates(), radial_orders=[2], coefficients=coefficients,
                               diameter=diameter)
    ref = np.tensordot(coefficients, orc.zernike_basis([4, 5, 6], orc.pixel_coords(64, 1.1),
                                                      diameter=1.1), 1)
    assert _peak(_np(dyn.evaluate(wavefront=wf)), ref) < 1e-5
    assert _peak(_np(explicit.evaluate()), ref) < 1e-5
    c = orc.pixel_coords(48, 0.8)
    d = dl.DynamicZernikeBasis(js=[4, 5], coefficients=np.ones(2))
    e = dl.ZernikeBasis(c, js=[4, 5], coefficients=np.ones(2), diameter=0.8)
    assert torch.allclose(d.evaluate(coordinates=c, diameter=0.8), e.evaluate(), atol=1e-6)
    # polike on rotated coordinates
    cr = orc.rotate_coords(orc.pixel_coords(80, 2.0), 0.3)
    poly = dl.DynamicZernikeBasis(js=[1, 4, 7], nsides=5)
    got = _np(poly.calculate_basis(coordinates=cr))
    ref = np.array([orc.polike(5, j, cr) for j in (1, 4, 7)])
    inner = np.hypot(*cr) < 0.75
    assert np.max(np.abs(got - ref)[:, inner]) / np.max(np.abs(ref)) < 1e-4
    # the n-sided basis from a wavefront grid (one kernel launch)
    wf2 = dl.Wavefront(700e-9, 64, diameter=2.0)
    poly_wf = _np(poly.calculate_basis(wavefront=wf2))
    ref_wf = np.array([orc.polike(5, j, orc.pixel_coords(64, 2.0)) for j in (1, 4, 7)])
    assert _peak(np.nan_to_num(poly_wf), np.nan_to_num(ref_wf)) < 1e-5
    # solve_basis recovers the coefficients
    sol = dyn.solve_basis(dyn.evaluate(wavefront=wf), wavefront=wf)
    assert np.allclose(sol, coefficients, atol=1e-5)
    # gradients: d/dc sum(v^2) = 2 B v
    coeffs = torch.ones(3, dtype=torch.float32, device="cuda", requires_grad=True)
    b = dl.DynamicZernikeBasis(radial_orders=[2], coefficients=coeffs)
    v = b.evaluate(wavefront=wf)
    torch.sum(v ** 2).backward()
    B = b.calculate_basis(wavefront=wf)
    ref_g = _np(2 * torch.tensordot(B, v.detach(), dims=([1, 2], [0, 1])))
    assert np.all(np.isfinite(_np(coeffs.grad)))
    assert _peak(_np(coeffs.grad), ref_g) < 1e-5


def test_odd_polygon_vertical_edge(dl):
    """A pentagon (an edge whose end points share x exactly: nandiv's
    infinite slope, the vertical rule) on the device matches the oracle and
    the closed-form n-gon."""
    print("[polynomials] odd polygon", flush=True)
    from dlux_amd import layers as L
    n, dx = 128, 1.0 / 128
    c = orc.pixel_coords(n, 1.0)
    for nsides in (5, 7):
        got = _np(L.RegPolyAperture(nsides, 0.4)._grid_transmission(n, dx))
        ref = orc.soft_aperture("reg_polygon", c, dx, nsides=nsides, rmax=0.4)
        assert np.max(np.abs(got - ref)) < 1e-6
        area = nsides / 2 * 0.4 ** 2 * np.sin(2 * np.pi / nsides)
        assert abs(got.sum() * dx * dx - area) / area < 0.01
