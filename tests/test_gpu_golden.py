"""GPU path against the committed FP64 golden fixtures (tests/golden/).

Tolerances (peak-normalised max error): MFT fields 2e-6 (FP32) / 3e-5
(split-BF16x3); PSFs 1e-5 (both forms); FFT fields 2e-6; Zernike basis 1e-5
(FP64 evaluation, float32 output, NaN where the reference yields NaN);
coefficient gradients 1e-4 (FP32) / 3e-4 (split-BF16x3), loss 1e-4 relative.
"""

from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLDEN = Path(__file__).resolve().parent / "golden"
FIELD_TOL = {"fp32": 2e-6, "bf16x3": 3e-5}
GRAD_TOL = {"fp32": 1e-4, "bf16x3": 3e-4}


@pytest.fixture(scope="module")
def rt():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dlux_amd import runtime
    return runtime


def _peak_err(x, ref):
    return float(np.nanmax(np.abs(np.asarray(x) - ref)) / np.nanmax(np.abs(ref)))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mft_golden(rt, prec):
    print(f"[golden-mft] {prec}", flush=True)
    d = np.load(GOLDEN / "mft_32to16.npz")
    for i in range(len(d["out"])):
        f = None if np.isnan(d["focal"][i]) else float(d["focal"][i])
        out = rt.mft(d["field"], 1.0, 0.1, 16, 0.05, focal_length=f,
                     shift_x=float(d["shift"][i][0]), shift_y=float(d["shift"][i][1]),
                     inverse=bool(d["inverse"][i]), precision=prec).cpu().numpy()
        assert _peak_err(out, d["out"][i]) < FIELD_TOL[prec], i


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("name,n_out", [("psf_mono_64to32.npz", 32),
                                        ("psf_poly3_48to24.npz", 24)])
def test_psf_golden(rt, prec, name, n_out):
    print(f"[golden-psf] {name} {prec}", flush=True)
    d = np.load(GOLDEN / name)
    mask = d["mask"]
    n = mask.shape[0]
    wls = d["wavelengths"] if "wavelengths" in d else np.array([float(d["wavelength"])])
    w = d["weights"] if "weights" in d else np.array([1.0])
    opd = np.tensordot(d["basis"], d["coeffs"], axes=(0, 0))
    norm = 1.0 / np.sqrt(np.sum((mask / n ** 2) ** 2))
    psf = rt.psf_poly(n, n_out, wls, w, 1.0 / n, float(d["p_out"]), amp=mask, opd=opd,
                      scale=norm / n ** 2, precision=prec).cpu().numpy()
    assert _peak_err(psf, d["psf"]) < 1e-5


def test_fft_golden(rt):
    print("[golden-fft]", flush=True)
    d = np.load(GOLDEN / "fft_16to32.npz")
    fwd = rt.fft2(d["field"], 32).cpu().numpy()
    inv = rt.fft2(d["field"], 32, inverse=True).cpu().numpy()
    assert _peak_err(fwd, d["forward"]) < 2e-6
    assert _peak_err(inv, d["inverse"]) < 2e-6


def test_zernike_golden(rt):
    print("[golden-zernike]", flush=True)
    from dlux_amd.utils import coordinates, zernikes
    d = np.load(GOLDEN / "zernike_16.npz")
    zb = zernikes.zernike_basis(d["js"], coordinates.pixel_coords(16, diameter=1.0), 1.0)
    zb = zb.cpu().numpy() if hasattr(zb, "cpu") else np.asarray(zb)
    assert np.array_equal(np.isnan(zb), np.isnan(d["basis"]))
    assert _peak_err(zb, d["basis"]) < 1e-5


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_gradient_golden(rt, prec):
    print(f"[golden-grad] {prec}", flush=True)
    d = np.load(GOLDEN / "grad_32to16.npz")
    mask, basis = d["mask"], d["basis"]
    n = mask.shape[0]
    norm = 1.0 / np.sqrt(np.sum((mask / n ** 2) ** 2))
    c = torch.tensor(d["coeffs"], dtype=torch.float32, device="cuda", requires_grad=True)
    B = torch.as_tensor(basis, dtype=torch.float32, device="cuda")
    opd = rt.eval_basis(B, c)
    psf = rt.psf_poly(n, 16, d["wavelengths"], d["weights"], 1.0 / n, float(d["p_out"]),
                      amp=mask, opd=opd, scale=norm / n ** 2, precision=prec)
    data = torch.as_tensor(d["data"], dtype=torch.float32, device="cuda")
    flux = float(d["flux"])
    loss = 0.5 * torch.sum((flux * psf - data).double() ** 2 / torch.clamp(data, min=1.0))
    loss.backward()
    assert abs(loss.item() - float(d["loss"])) < 1e-4 * float(d["loss"])
    g = c.grad.double().cpu().numpy()
    assert _peak_err(g, d["grad"]) < GRAD_TOL[prec]
