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


# --- geometry gradients -------------------------------------------------------

_SLOT_FIELDS = [("size", 0), ("size", 1), ("translation", 0), ("translation", 1), ("shear", 0),
                ("shear", 1), ("compression", 0), ("compression", 1), ("rotation", None)]


def _oracle_of(kind, shape, tf):
    """Oracle transmission as a function of the 9 parameter slots."""
    def f(theta, coords, dx, softening, occulting):
        sh = dict(shape)
        if kind == "circle":
            sh["radius"] = theta[0]
        elif kind in ("square", "spider"):
            sh["width"] = theta[0]
        elif kind == "rectangle":
            sh["width"], sh["height"] = theta[0], theta[1]
        elif kind == "reg_polygon":
            sh["rmax"] = theta[0]
        t = {}
        if "translation" in tf:
            t["translation"] = [theta[2], theta[3]]
        if "shear" in tf:
            t["shear"] = [theta[4], theta[5]]
        if "compression" in tf:
            t["compression"] = [theta[6], theta[7]]
        if "rotation" in tf:
            t["rotation"] = theta[8]
        return orc.soft_aperture(kind, coords, dx, softening=softening, occulting=occulting,
                                 transforms=[t] if t else [], **sh)
    return f


def _theta0(kind, shape, tf):
    th = np.zeros(9)
    th[0] = {"circle": shape.get("radius"), "square": shape.get("width"),
             "spider": shape.get("width"), "rectangle": shape.get("width"),
             "reg_polygon": shape.get("rmax")}[kind]
    if kind == "rectangle":
        th[1] = shape["height"]
    for k, (name, e) in enumerate(_SLOT_FIELDS[2:], start=2):
        if name in tf:
            th[k] = tf[name] if e is None else tf[name][e]
    return th


@pytest.mark.parametrize("case", range(7))
def test_soft_aperture_geometry_vjp_vs_fd(dl, case):
    """dlx_soft_apertures_vjp against central differences of the oracle
    transmission, L = sum(dout * T) for a random dout; every slot the
    aperture has.  Bound 1e-5 of the largest gradient."""
    from dlux_amd import layers as L, runtime
    kind, ap, shape, occ = _cases(dl, L)[case]
    print(f"[aperture vjp] {kind} occulting={occ}", flush=True)
    n, dx = 128, 1.0 / 128
    coords = orc.pixel_coords(n, 1.0)
    tf = TF if ap.transformation is not None and ap.transformation.translation is not None \
        else {}
    descs = [d for _, d, _ in ap._descs(())]
    dout = np.random.default_rng(case).normal(size=(n, n))
    g = runtime.soft_apertures_vjp(n, dx, descs, torch.as_tensor(dout, device="cuda")) \
        .double().cpu().numpy()[0]
    f = _oracle_of(kind, shape, tf)
    th0 = _theta0(kind, shape, tf)
    slots = [0] + ([1] if kind == "rectangle" else []) + (list(range(2, 9)) if tf else [])
    ref = np.zeros(9)
    for k in slots:
        h = 1e-7 * max(1.0, abs(th0[k]))
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        ref[k] = np.sum(dout * (f(tp, coords, dx, float(ap.softness), occ)
                                - f(tm, coords, dx, float(ap.softness), occ))) / (2 * h)
    err = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
    print(f"  grad {np.array2string(g[slots], precision=4)}\n  err {err:.2e}")
    assert err < 1e-5
    others = [k for k in range(9) if k not in slots]
    assert np.all(g[others] == 0)


def test_compound_geometry_vjp_vs_fd(dl):
    """Geometry gradients through a CompoundAperture (product) inside a
    MultiAperture (sum): the other members of a group enter the cotangent."""
    print("[aperture vjp] compound / multi", flush=True)
    from dlux_amd import layers as L, runtime
    n, dx = 96, 1.0 / 96
    c = orc.pixel_coords(n, 1.0)
    inner = dict(translation=[-0.1, 0.05], rotation=0.2)
    ap = L.MultiAperture([
        L.CompoundAperture([L.CircularAperture(0.3, dl.CoordTransform(**inner)),
                            L.Spider(0.02, [30.0, 150.0], dl.CoordTransform(**inner))]),
        L.SquareAperture(0.2, dl.CoordTransform(translation=[0.25, 0.2])),
    ])
    from dlux_amd.layers.apertures import _grouped
    descs = _grouped(ap._descs(()))
    dout = np.random.default_rng(5).normal(size=(n, n))
    g = runtime.soft_apertures_vjp(n, dx, descs, torch.as_tensor(dout, device="cuda")) \
        .double().cpu().numpy()
    assert len(descs) == 3

    def T(r, w, s, t0):
        tf_in = dict(translation=[t0, 0.05], rotation=0.2)
        return (orc.soft_aperture("circle", c, dx, transforms=[tf_in], radius=r)
                * orc.soft_aperture("spider", c, dx, occulting=True, transforms=[tf_in],
                                    width=w, angles=[30.0, 150.0])
                + orc.soft_aperture("square", c, dx, transforms=[dict(translation=[0.25, 0.2])],
                                    width=s))
    base = [0.3, 0.02, 0.2, -0.1]
    ref = []
    for k in range(4):
        h = 1e-7
        p, m = list(base), list(base)
        p[k] += h
        m[k] -= h
        ref.append(np.sum(dout * (T(*p) - T(*m))) / (2 * h))
    # translation x of the circle and spider's transforms both feed t0
    got = [g[0, 0], g[1, 0], g[2, 0], g[0, 2] + g[1, 2]]
    err = np.max(np.abs(np.array(got) - ref)) / np.max(np.abs(ref))
    print(f"  err {err:.2e}")
    assert err < 1e-5


def test_geometry_gradients_through_psf(dl):
    """A radius and a translation given as tensors that require grad: the
    gradient of a PSF loss through the fused pupil path (amplitude cotangent
    of dlx_psf_poly_vjp, then dlx_soft_apertures_vjp) against central
    differences of the oracle PSF on the oracle's transmission.  Bound 1e-3
    of the largest (FP32 PSF path; FD in FP64)."""
    print("[aperture vjp] through the PSF", flush=True)
    from dlux_amd import layers as L
    from dlux_amd.utils import units
    n, n_out, p = 128, 48, 2.5e-7
    radius = torch.tensor(0.45, dtype=torch.float64, requires_grad=True)
    shift = torch.tensor([0.02, -0.01], dtype=torch.float64, requires_grad=True)
    ap = L.CircularAperture(radius, dl.CoordTransform(translation=shift, rotation=0.0),
                            softening=2.0)
    sp = L.Spider(0.015, [0.0, 120.0, 240.0])
    optics = dl.AngularOpticalSystem(
        wf_npixels=n, diameter=1.0, layers=[("pupil", ap), ("spider", sp)],
        psf_npixels=n_out, psf_pixel_scale=float(units.rad2arcsec(p)))
    wls = np.array([0.9e-6, 1.1e-6])
    psf = optics.propagate(wls)
    c = orc.pixel_coords(n, 1.0)
    dx = 1.0 / n
    spider = orc.soft_aperture("spider", c, dx, occulting=True, width=0.015,
                               angles=[0.0, 120.0, 240.0])
    w = np.full(2, 0.5)

    def model(r, tx, ty):
        amp = orc.soft_aperture("circle", c, dx, softening=2.0,
                                transforms=[dict(translation=[tx, ty], rotation=0.0)],
                                radius=r) * spider
        return orc.poly_psf(amp, np.zeros((1, n, n)), np.zeros(1), wls, w, 1.0, n_out, p,
                            normalise_power=False)
    ref = model(0.45, 0.02, -0.01)
    assert float(np.max(np.abs(psf.detach().double().cpu().numpy() - ref)) / ref.max()) < 1e-5
    wgt = np.random.default_rng(9).normal(size=(n_out, n_out))
    loss = torch.sum(psf * torch.as_tensor(wgt, device=psf.device, dtype=psf.dtype)) * 1e3
    loss.backward()
    got = np.array([float(radius.grad), float(shift.grad[0]), float(shift.grad[1])])
    fd = []
    for k in range(3):
        h = 1e-7   # below the 8e-7 by which one pixel clears the clip bound (a kink)
        a, b = [0.45, 0.02, -0.01], [0.45, 0.02, -0.01]
        a[k] += h
        b[k] -= h
        fd.append(1e3 * np.sum(wgt * (model(*a) - model(*b))) / (2 * h))
    fd = np.array(fd)
    err = np.max(np.abs(got - fd)) / np.max(np.abs(fd))
    print(f"  got {got}  fd {fd}  err {err:.2e}")
    assert err < 1e-3
    # gradients that are not implemented say so
    ab = L.AberratedAperture(L.CircularAperture(radius), [2, 3])
    with pytest.raises(NotImplementedError):
        ab._grid_basis(32, 1 / 32)
    with pytest.raises(NotImplementedError):
        L.CircularAperture(0.4, softening=torch.tensor(1.0, requires_grad=True)) \
            ._grid_transmission(32, 1 / 32)
