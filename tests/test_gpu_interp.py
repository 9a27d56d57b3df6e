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
