"""GPU parity of the resampling kernels (dlx_interp_transform,
dlx_interp_points) and of the API built on them (utils.interp / scale /
rotate, Wavefront.scale_to / interpolate / rotate, PSF.rotate / interpolate,
Interpolate, ApplyInterpolation) against the FP64 oracle's bilinear / nearest
rules, which test_oracle pins against scipy.  interpax, where the reference
keeps these, is absent: parity with it is unpinned.

Bound: 2e-6 of the image peak (float32 output of FP64 weights).
"""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 2e-6
TF = dict(translation=[0.3, -0.2], rotation=0.4, compression=[1.1, 0.9], shear=[0.05, -0.03])


@pytest.fixture(scope="module")
def dl():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dlux_amd
    return dlux_amd


def _err(x, ref):
    x = x.detach().double().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


def _oracle_tf(img, n_in, d_in, c_in, n_out, d_out, c_out, tf, method, fill):
    xo = orc.spec_xs(n_out, d_out, c_out)
    grid = np.array(np.meshgrid(xo, xo, indexing="xy"))
    q = orc.coord_transform(grid, **tf) if tf else grid
    return orc.interp_grid(img, n_in, d_in, c_in, q[0], q[1], method, fill)


@pytest.mark.parametrize("method", ["linear", "nearest"])
def test_interp_transform_and_points_vs_oracle(dl, method):
    print(f"[interp] {method}", flush=True)
    from dlux_amd import runtime
    rng = np.random.default_rng(3)
    n_in, d_in, c_in = 48, 0.5, 0.2
    imgs = rng.normal(size=(3, n_in, n_in))
    tf = dl.CoordTransform(**TF)._struct()
    got = runtime.interp_transform(torch.as_tensor(imgs, dtype=torch.float32), d_in, c_in, 40,
                                   0.6, -0.1, [tf], method, fill=1.5)
    for b in range(3):
        ref = _oracle_tf(imgs[b].astype(np.float32), n_in, d_in, c_in, 40, 0.6, -0.1, TF,
                         method, 1.5)
        assert _err(got[b], ref) < TOL
    xs, ys = rng.uniform(-14, 14, (7, 9)), rng.uniform(-14, 14, (7, 9))
    got = runtime.interp_points(torch.as_tensor(imgs, dtype=torch.float32), d_in, c_in, xs, ys,
                                method, fill=-2.0)
    assert tuple(got.shape) == (3, 7, 9)
    for b in range(3):
        ref = orc.interp_grid(imgs[b].astype(np.float32), n_in, d_in, c_in, xs, ys, method, -2.0)
        assert _err(got[b], ref) < TOL
    with pytest.raises(NotImplementedError):
        runtime.interp_points(torch.zeros((4, 4)), 1.0, 0.0, [0.0], [0.0], "cubic")


def test_nearest_takes_the_upper_knot_at_half_pixel_ties(dl):
    """Samples exactly half-way between knots (a 2x upsampling by an integer
    grid shift) pick the upper knot on both axes, in the points kernel and the
    grid kernel alike."""
    print("[interp] nearest ties", flush=True)
    from dlux_amd import runtime
    n, d = 8, 0.5
    img = np.arange(n * n, dtype=np.float32).reshape(n, n)
    half = orc.spec_xs(n, d, 0.0)[:-1] + d / 2
    xs, ys = np.meshgrid(half, half[::-1])
    got = runtime.interp_points(torch.as_tensor(img), d, 0.0, xs, ys, "nearest", fill=-1.0)
    ref = orc.interp_grid(img, n, d, 0.0, xs, ys, "nearest", -1.0)
    assert np.array_equal(got.cpu().numpy(), ref.astype(np.float32))
    assert np.array_equal(ref[0], img[n - 1, 1:])
    # grid form: n - 1 output knots centred between the input ones
    got = runtime.interp_transform(torch.as_tensor(img[None]), d, 0.0, n - 1, d, d / 2, [],
                                   "nearest", fill=-1.0)
    xo = orc.spec_xs(n - 1, d, d / 2)
    gx, gy = np.meshgrid(xo, xo)
    ref = orc.interp_grid(img, n, d, 0.0, gx, gy, "nearest", -1.0)
    assert np.array_equal(got[0].cpu().numpy(), ref.astype(np.float32))


def test_resampling_api_vs_oracle(dl):
    """PSF.rotate / interpolate, Wavefront.rotate / scale_to / interpolate and
    the two layers, each against the oracle rule on the same grids."""
    print("[interp] API", flush=True)
    from dlux_amd import layers
    rng = np.random.default_rng(5)
    n = 32
    data = rng.uniform(0, 1, (n, n)).astype(np.float32)
    psf = dl.PSF(torch.as_tensor(data, device="cuda"), 0.25)
    ref = _oracle_tf(data, n, 1.0, 0.0, n, 1.0, 0.0, dict(rotation=0.3), "linear", 0.0)
    assert _err(psf.rotate(0.3).data, ref) < TOL
    ref = _oracle_tf(data, n, 0.25, 0.0, n, 0.25, 0.0, TF, "linear", 0.0)
    assert _err(psf.interpolate(dl.CoordTransform(**TF)).data, ref) < TOL
    assert _err(layers.ApplyInterpolation(dl.CoordTransform(**TF))(psf).data, ref) < TOL
    det = dl.LayeredDetector([layers.ApplyInterpolation(dl.CoordTransform(**TF)),
                              layers.Downsample(2)])
    ref_det = ref.reshape(n // 2, 2, n // 2, 2).sum((1, 3))
    assert _err(det(psf), ref_det) < TOL
    # wavefronts: real and imaginary parts resampled separately
    wf = dl.Wavefront(1e-6, n, 1.0)
    ph = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))).astype(np.complex64)
    wf = wf.set(phasor=torch.as_tensor(ph, device="cuda"))

    def both(f):
        return f(ph.real) + 1j * f(ph.imag)
    got = wf.rotate(0.2).phasor.cpu().numpy()
    assert _err(got, both(lambda a: _oracle_tf(a, n, 1.0, 0.0, n, 1.0, 0.0,
                                                dict(rotation=0.2), "linear", 0.0))) < TOL
    got = wf.interpolate(dl.CoordTransform(**TF)).phasor.cpu().numpy()
    ref = both(lambda a: _oracle_tf(a, n, 1 / n, 0.0, n, 1 / n, 0.0, TF, "linear", 0.0))
    assert _err(got, ref) < TOL
    assert _err(layers.Interpolate(dl.CoordTransform(**TF))(wf).phasor.cpu().numpy(), ref) < TOL
    wf2 = wf.scale_to(24, 1.5 / n)
    r = (1.5 / n) / (1.0 / n) * 24 / n
    ref = both(lambda a: _oracle_tf(a, n, 1 / n, 0.0, 24, 1 / 24, 0.0,
                                    dict(compression=[r, r]), "linear", 0.0))
    assert wf2.npixels == 24 and np.isclose(float(np.ravel(wf2.pixel_scale)[0]), 1.5 / n)
    assert _err(wf2.phasor.cpu().numpy(), ref) < TOL


def _sampling_matrix(n_in, d_in, c_in, xs, ys, method):
    """The oracle's sampling rule as a matrix: column k is the resampling of
    the k-th unit image (fill 0)."""
    cols = []
    for k in range(n_in * n_in):
        e = np.zeros(n_in * n_in)
        e[k] = 1.0
        cols.append(orc.interp_grid(e.reshape(n_in, n_in), n_in, d_in, c_in, xs, ys, method,
                                    0.0).ravel())
    return np.array(cols).T


@pytest.mark.parametrize("method", ["linear", "nearest"])
def test_resampling_gradient_is_the_transposed_sampling(dl, method):
    """Gradients through the resampling (dlx_interp_transform_vjp,
    dlx_interp_points_vjp): equal to S^T dout with S the oracle's sampling
    matrix (1e-6 of the largest entry), for a transform chain and for
    explicit points, over a batch; the dot-product identity <S x, y> =
    <x, S^T y> at a larger size (1e-5 relative)."""
    print(f"[interp vjp] {method}", flush=True)
    from dlux_amd import runtime
    rng = np.random.default_rng(8)
    n_in, d_in, c_in, n_out, d_out = 9, 0.5, 0.1, 7, 0.6
    tf = dl.CoordTransform(**TF)
    xo = orc.spec_xs(n_out, d_out, 0.05)
    q = orc.coord_transform(np.array(np.meshgrid(xo, xo, indexing="xy")), **TF)
    S = _sampling_matrix(n_in, d_in, c_in, q[0], q[1], method)
    x = torch.tensor(rng.normal(size=(2, n_in, n_in)), dtype=torch.float32, device="cuda",
                     requires_grad=True)
    y = rng.normal(size=(2, n_out, n_out))
    out = runtime.interp_transform(x, d_in, c_in, n_out, d_out, 0.05, [tf._struct()], method, 0.0)
    (out * torch.as_tensor(y, dtype=torch.float32, device="cuda")).sum().backward()
    ref = np.stack([S.T @ y[b].ravel() for b in range(2)]).reshape(2, n_in, n_in)
    err = np.max(np.abs(x.grad.double().cpu().numpy() - ref)) / np.max(np.abs(ref))
    print(f"  transform: err {err:.2e}")
    assert err < 1e-6
    # explicit points (some outside the span), through the same rule
    px = rng.uniform(-3, 3, size=(5, 6))
    py = rng.uniform(-3, 3, size=(5, 6))
    S = _sampling_matrix(n_in, d_in, c_in, px, py, method)
    x.grad = None
    out = runtime.interp_points(x, d_in, c_in, px, py, method, -1.0)
    yp = rng.normal(size=(2, 5, 6))
    (out * torch.as_tensor(yp, dtype=torch.float32, device="cuda")).sum().backward()
    ref = np.stack([S.T @ yp[b].ravel() for b in range(2)]).reshape(2, n_in, n_in)
    err = np.max(np.abs(x.grad.double().cpu().numpy() - ref)) / np.max(np.abs(ref))
    print(f"  points: err {err:.2e}")
    assert err < 1e-6
    # dot-product identity at 256^2 -> 200^2
    big = torch.tensor(rng.normal(size=(256, 256)), dtype=torch.float32, device="cuda",
                       requires_grad=True)
    yb = torch.tensor(rng.normal(size=(200, 200)), dtype=torch.float32, device="cuda")
    ob = runtime.interp_transform(big, 1.0, 0.0, 200, 1.1, 0.0, [tf._struct()], method, 0.0)
    (ob * yb).sum().backward()
    lhs = float((ob.detach().double() * yb.double()).sum())
    rhs = float((big.detach().double() * big.grad.double()).sum())
    print(f"  <Sx,y> {lhs:.6e} <x,S^T y> {rhs:.6e}")
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), 1.0)


def test_resampling_gradients_through_the_api(dl):
    """PSF.interpolate, ApplyInterpolation in a LayeredDetector (then a
    downsample) and Wavefront.rotate carry gradients to their images; a
    tracked transformation raises."""
    print("[interp vjp] API", flush=True)
    from dlux_amd import layers
    rng = np.random.default_rng(9)
    n = 16
    data = torch.tensor(rng.uniform(0, 1, (n, n)), dtype=torch.float32, device="cuda",
                        requires_grad=True)
    tfd = dict(rotation=0.3, translation=[0.1, -0.05])
    xs = orc.spec_xs(n, 0.25)
    q = orc.coord_transform(np.array(np.meshgrid(xs, xs, indexing="xy")), **tfd)
    S = _sampling_matrix(n, 0.25, 0.0, q[0], q[1], "linear")
    psf = dl.PSF(data, 0.25)
    w = rng.normal(size=(n, n))
    (psf.interpolate(dl.CoordTransform(**tfd)).data
     * torch.as_tensor(w, dtype=torch.float32, device="cuda")).sum().backward()
    ref = (S.T @ w.ravel()).reshape(n, n)
    assert np.max(np.abs(data.grad.double().cpu().numpy() - ref)) / np.max(np.abs(ref)) < 1e-6
    data.grad = None
    det = dl.LayeredDetector([layers.ApplyInterpolation(dl.CoordTransform(**tfd)),
                              layers.Downsample(2)])
    w2 = rng.normal(size=(n // 2, n // 2))
    (det(dl.PSF(data, 0.25)) * torch.as_tensor(w2, dtype=torch.float32, device="cuda")).sum() \
        .backward()
    ref = (S.T @ np.kron(w2, np.ones((2, 2))).ravel()).reshape(n, n)
    assert np.max(np.abs(data.grad.double().cpu().numpy() - ref)) / np.max(np.abs(ref)) < 1e-6
    ph = torch.tensor(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)),
                      dtype=torch.complex64, device="cuda", requires_grad=True)
    wf = dl.Wavefront(1e-6, n, 1.0).set(phasor=ph)
    wf.rotate(0.2).psf.sum().backward()
    assert ph.grad is not None and torch.isfinite(ph.grad).all()
    rot = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    with pytest.raises(NotImplementedError):
        psf.interpolate(dl.CoordTransform(rotation=rot))


def test_intensity_gradient(dl):
    """dlx_intensity carries gradients to the fields: dL/dE = 2 E dL/dI, summed
    over the batch or per plane (1e-6 relative)."""
    print("[intensity vjp]", flush=True)
    from dlux_amd import runtime
    rng = np.random.default_rng(10)
    e = rng.normal(size=(3, 12, 12)) + 1j * rng.normal(size=(3, 12, 12))
    for sum_batch in (True, False):
        E = torch.tensor(e, dtype=torch.complex64, device="cuda", requires_grad=True)
        g = rng.normal(size=(12, 12) if sum_batch else (3, 12, 12))
        I = runtime.intensity(E, sum_batch=sum_batch)
        ref_i = (np.abs(e) ** 2).sum(0) if sum_batch else np.abs(e) ** 2
        assert np.max(np.abs(I.detach().double().cpu().numpy() - ref_i)) < 1e-5 * ref_i.max()
        (I * torch.as_tensor(g, dtype=torch.float32, device="cuda")).sum().backward()
        ref = 2 * e * (g[None] if sum_batch else g)
        err = np.max(np.abs(E.grad.cpu().numpy() - ref)) / np.max(np.abs(ref))
        assert err < 1e-6, (sum_batch, err)


def test_non_differentiable_entry_points_raise_on_tracked_inputs(dl):
    """The unfused MFT, FFT, ASM, LCT, power, pupil-field and batched-MFT
    entry points raise on an input that requires grad (rather than return a
    result with no gradient path); under no_grad they run."""
    print("[untracked inputs]", flush=True)
    from dlux_amd import runtime
    e = torch.ones((1, 16, 16), dtype=torch.complex64, device="cuda", requires_grad=True)
    calls = [lambda: runtime.mft(e, 1e-6, 1 / 16, 8, 1e-7),
             lambda: runtime.fft2(e, 32),
             lambda: runtime.asm_prop(e, 1e-6, 1 / 16, 1e-3, 32),
             lambda: runtime.power(e),
             lambda: runtime.pupil_field(16, 1e-6, 1 / 16, opd=e.real[0]),
             lambda: runtime.mft_batched_psf([dict(field=e[0], wavelength=1e-6,
                                                   pixel_scale_in=1 / 16, n_out=8,
                                                   pixel_scale_out=1e-7)])]
    for f in calls:
        with pytest.raises(NotImplementedError):
            f()
        with torch.no_grad():
            f()
