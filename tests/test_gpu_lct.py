"""GPU parity of the linear canonical transforms (dlx_lct, dlx_lct_fft) and
the ABCD propagator layers against the FP64 oracle's Collins sum, which
test_oracle pins by closed forms (a tilted Gaussian beam through free space,
lenses, mirrors and the conjugate plane; the conjugate plane as the MFT; the
FFT form as the Collins sum on the FFT grid).  abcdLux, where the reference
keeps these transforms, is absent: parity with it is unpinned.

Bounds (peak-normalised max error of the complex field): FP32 GEMM form 2e-5,
split-BF16x3 form 5e-5; FFT form 2e-5.
"""

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


def _peak(got, ref):
    return float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))


_MATS = {
    "free": orc.abcd_free_space(0.5),
    "free_back": orc.abcd_free_space(-0.4),
    "lens_free": orc.compose_abcd([orc.abcd_free_space(0.3), orc.abcd_lens(0.8),
                                   orc.abcd_free_space(0.2)]),
    "conjugate": orc.abcd_conjugate_plane(0.6),
}


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("case", list(_MATS))
def test_lct_vs_oracle(rt, case, prec):
    print(f"[lct] {case} {prec}", flush=True)
    rng = np.random.default_rng(len(case))
    n_in, n_out, nb = 64, 48, 3
    dx, dxo, c_in, c_out = 40e-6, 35e-6, 2e-5, -7e-5
    wls = np.array([0.9e-6, 1.0e-6, 1.2e-6])
    u = rng.normal(size=(nb, n_in, n_in)) + 1j * rng.normal(size=(nb, n_in, n_in))
    ut = torch.as_tensor(u.astype(np.complex64), device="cuda")
    # three calls: generated, then cached on the second sighting, then a hit
    runs = [rt.lct(ut, wls, _MATS[case], dx, c_in, n_out, dxo, c_out, precision=prec)
            for _ in range(3)]
    assert all(torch.equal(runs[0], r) for r in runs[1:])
    got = runs[0].cpu().numpy()
    tol = {"fp32": 2e-5, "bf16x3": 5e-5}[prec]
    for b in range(nb):
        ref = orc.lct(u[b], wls[b], orc.spec_xs(n_in, dx, c_in), orc.spec_xs(n_out, dxo, c_out),
                      _MATS[case])
        err = _peak(got[b], ref)
        print(f"  wl {wls[b]:.2e} err {err:.2e}")
        assert err < tol


@pytest.mark.parametrize("n,m", [(48, 64), (64, 128), (32, 32)])
def test_lct_fft_vs_oracle(rt, n, m):
    print(f"[lct fft] n={n} m={m}", flush=True)
    rng = np.random.default_rng(n * m)
    wls = np.array([0.8e-6, 1.1e-6])
    for case in ("lens_free", "free_back", "conjugate"):
        u = rng.normal(size=(2, n, n)) + 1j * rng.normal(size=(2, n, n))
        got, ps = rt.lct_fft(torch.as_tensor(u.astype(np.complex64), device="cuda"), wls,
                             _MATS[case], 30e-6, 1e-5, m, -4e-5)
        got = got.cpu().numpy()
        for b in range(2):
            ref, d_out = orc.lct_fft(u[b], wls[b], 30e-6, 1e-5, -4e-5, m, _MATS[case])
            assert abs(ps[b] - d_out) <= 1e-15 * d_out
            err = _peak(got[b], ref)
            print(f"  {case} wl {wls[b]:.2e} err {err:.2e}")
            assert err < 2e-5


def test_conjugate_plane_lct_is_the_device_mft(rt):
    """The conjugate-plane LCT on the device is the library's MFT with focal
    length f, times 1/i."""
    print("[lct] conjugate plane vs dlx_mft", flush=True)
    rng = np.random.default_rng(4)
    n, n_out, dx, dxo, wl, f = 64, 40, 1e-3 / 64, 3e-6, 1e-6, 2.0
    u = torch.as_tensor((rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
                        .astype(np.complex64), device="cuda")
    a = rt.lct(u, [wl], orc.abcd_conjugate_plane(f), dx, 0.0, n_out, dxo, 0.0)
    b = rt.mft(u, [wl], [dx], n_out, [dxo], focal_length=f) / 1j
    assert _peak(a.cpu().numpy(), b.cpu().numpy()) < 2e-5


def test_abcd_propagator_layers(rt):
    """tests/layers/test_abcd_propagators.py calls, and the layers against the
    oracle: MFTPropagator to a CoordSpec grid; FFTPropagator on a power-of-two
    FFT grid (dlx_lct_fft) and on a 20-point one (dlx_lct's GEMMs), then the
    PadSpec crop."""
    print("[lct] layers", flush=True)
    import dlux_amd as dl
    from dlux_amd import layers as L
    from dlux_amd.coordinates import CoordSpec, PadSpec
    wf = dl.Wavefront(1e-6, 16, 1.0)
    out = L.MFTPropagator([L.ABCDFreeSpace(1.0)], CoordSpec(n=8, d=1 / 32, c=0.0))(wf)
    assert out.npixels == 8 and np.allclose(out.pixel_scale, 1 / 32)
    # (a metre of free space after a metre-wide 16-pixel pupil aliases the
    # chirp beyond use: the numbers are compared on a sampled configuration)
    wf_s = dl.Wavefront(1e-6, 64, 64 * 40e-6)
    ap = orc.circle(orc.pixel_coords(64, 64 * 40e-6), 0.8e-3)
    wf_s = wf_s * torch.as_tensor(ap.astype(np.float32), device="cuda")
    out = L.MFTPropagator([L.ABCDFreeSpace(0.5)], CoordSpec(n=48, d=35e-6, c=1e-5))(wf_s)
    assert out.npixels == 48 and np.allclose(out.center, 1e-5)
    us = wf_s.phasor.cpu().numpy().astype(np.complex128)
    ref = orc.lct(us, 1e-6, orc.spec_xs(64, 40e-6), orc.spec_xs(48, 35e-6, 1e-5),
                  orc.abcd_free_space(0.5))
    assert _peak(out.phasor.cpu().numpy(), ref) < 2e-5
    u = wf.phasor.cpu().numpy().astype(np.complex128)
    out = L.FFTPropagator([L.ABCDConjugatePlane(1.0)], CoordSpec(n=20, c=0.0))(wf)
    assert out.npixels == 20
    d_out = 1e-6 * 1.0 / (20 * (1 / 16))
    assert np.allclose(out.pixel_scale, d_out)
    ref = orc.lct(u, 1e-6, orc.spec_xs(16, 1 / 16), orc.spec_xs(20, d_out),
                  orc.abcd_conjugate_plane(1.0))
    assert _peak(out.phasor.cpu().numpy(), ref) < 2e-5
    out = L.FFTPropagator([L.ABCDConjugatePlane(1.0)], PadSpec(pad=2, crop=2, c=0.0))(wf)
    assert out.npixels == 16
    full, d32 = orc.lct_fft(u, 1e-6, 1 / 16, 0.0, 0.0, 32, orc.abcd_conjugate_plane(1.0))
    assert np.allclose(out.pixel_scale, d32)
    assert _peak(out.phasor.cpu().numpy(), full[8:24, 8:24]) < 2e-5
