"""dlx_psf_poly_noalloc, the PSF entry point that allocates nothing (its buffers
live in the library's fixed per-GPU scratch array), called through the C ABI
and compared with dlx_psf_poly (caller workspace, device matrix cache) on the
same inputs: two calls in a row, the second with a smaller input.

Bar: the polychromatic PSF bound, 1e-5 of the peak against the workspace entry
point; both forms run the same kernels per wavelength and sum the wavelengths
in the same order, so the results are expected to be identical and the test
reports whether they are.
"""

import ctypes

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


def _noalloc(lib):
    """The entry point (argument types from the binding's table, which the host
    tests match against include/dlux_amd.h) and the scratch array's size."""
    return lib.dlx_psf_poly_noalloc, int(lib.dlx_psf_poly_noalloc_scratch_bytes())


def _case(n, n_out, nwl, seed):
    mask, basis = orc.circular_aperture(n, 1.0, oversample=1, spider_width=8.0 / n,
                                        spider_angles=[0, 120, 240],
                                        zernike_nolls=range(2, 12))
    c = np.random.default_rng(seed).normal(0, 30e-9, basis.shape[0])
    opd = np.tensordot(np.nan_to_num(basis), c, axes=(0, 0))
    wls = np.linspace(0.8e-6, 1.2e-6, nwl)
    w = np.random.default_rng(seed + 1).uniform(0.5, 1.5, nwl)
    return (torch.as_tensor(mask, dtype=torch.float32, device="cuda").contiguous(),
            torch.as_tensor(opd, dtype=torch.float32, device="cuda").contiguous(), wls, w)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_noalloc_matches_workspace_entry_point(rt, prec):
    from dlux_amd import _lib
    lib = _lib.load()
    noalloc, cap = _noalloc(lib)
    assert cap == 64 << 20
    code = _lib.DLX_PREC_BF16X3 if prec == "bf16x3" else _lib.DLX_PREC_FP32
    p_out = 2.5e-7
    # a large call, then a smaller one: the array keeps the first call's contents
    for n, n_out, nwl, seed in ((512, 256, 64, 3), (96, 40, 7, 4)):
        print(f"[noalloc] prec={prec} {n}->{n_out} x {nwl}", flush=True)
        amp, opd, wls, w = _case(n, n_out, nwl, seed)
        scale = 1.0 / float(torch.sqrt(torch.sum(amp.double() ** 2)))
        ref = rt.psf_poly(n, n_out, wls, w, 1.0 / n, p_out, amp=amp, opd=opd, scale=scale,
                          precision=prec)
        out = torch.full((n_out, n_out), float("nan"), device="cuda")
        g = rt.geom(n, n_out)
        st = noalloc(ctypes.byref(g), nwl, _lib.dptr(wls), _lib.dptr(w), 1.0 / n, p_out, 0.0,
                     0.0, 0.0, amp.data_ptr(), opd.data_ptr(), None, scale, out.data_ptr(),
                     code, torch.cuda.current_stream().cuda_stream)
        _lib.check(st, "dlx_psf_poly_noalloc")
        torch.cuda.synchronize()
        r, o = ref.double().cpu().numpy(), out.double().cpu().numpy()
        err = float(np.max(np.abs(o - r)) / np.max(r))
        print(f"  peak-normalised difference {err:.3e}, identical {np.array_equal(o, r)}")
        assert np.all(np.isfinite(o))
        assert err <= 1e-5


def test_noalloc_rejects_sizes_beyond_the_array(rt):
    from dlux_amd import _lib
    lib = _lib.load()
    noalloc, cap = _noalloc(lib)
    n, n_out = 4096, 16     # one 4096^2 split4 pupil alone is 128 MiB
    g = rt.geom(n, n_out)
    out = torch.zeros((n_out, n_out), device="cuda")
    st = noalloc(ctypes.byref(g), 1, _lib.dptr([1e-6]), _lib.dptr([1.0]), 1.0 / n, 2.5e-7, 0.0,
                 0.0, 0.0, None, None, None, 1.0, out.data_ptr(), _lib.DLX_PREC_BF16X3,
                 torch.cuda.current_stream().cuda_stream)
    assert st == 1
    assert b"scratch" in lib.dlx_last_error()
    torch.cuda.synchronize()
    assert float(out.abs().sum()) == 0.0
