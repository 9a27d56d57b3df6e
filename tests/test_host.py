"""Host-side tests (CPU only): the C ABI library loads and exports every
declared symbol, its argument validation rejects bad calls without touching a
GPU, and the Python layer reproduces dLux's eager validation errors."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "dlux_amd.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dlx_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from dlux_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("native library not built (run __graft_entry__.build())")
    return _lib.load()


def test_library_exports_every_declared_symbol(lib):
    names = _declared()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name


def test_ctypes_signature_table_matches_header():
    from dlux_amd import _lib
    assert sorted(_lib.SIGNATURES) == _declared()


def test_version_and_noll(lib):
    assert b"gfx950" in lib.dlx_version()
    n, m = ctypes.c_int32(), ctypes.c_int32()
    for j, nm in zip(range(1, 11), [(0, 0), (1, 1), (1, -1), (2, 0), (2, -2), (2, 2), (3, -1),
                                    (3, 1), (3, -3), (3, 3)]):
        assert lib.dlx_noll_indices(j, ctypes.byref(n), ctypes.byref(m)) == 0
        assert (n.value, m.value) == nm
    assert lib.dlx_noll_indices(0, ctypes.byref(n), ctypes.byref(m)) == 1
    assert b"greater than 0" in lib.dlx_last_error()


def test_abi_rejects_bad_arguments_without_gpu(lib):
    from dlux_amd import _lib
    g = _lib.MftGeom(n_in=0, n_out=8, inverse=0, has_focal=0, focal_length=0.0,
                     focal_shift=0.0)
    one = _lib.dptr([1.0])
    st = lib.dlx_mft(g, 1, 1, 1, one, one, one, None, None, None, 0, None, 0, None)
    assert st == 1 and b"n_in" in lib.dlx_last_error()
    g.n_in = 16
    st = lib.dlx_mft(g, 1, 1, 1, one, one, one, None, None, None, 7, None, 0, None)
    assert st == 1 and b"precision" in lib.dlx_last_error()
    st = lib.dlx_mft(g, 1, 1, 1, one, one, one, None, None, None, 0, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()
    g.has_focal, g.focal_length = 1, 0.0
    st = lib.dlx_transfer_matrix(g, 1, one, one, one, None, 1, None)
    assert st == 1 and b"focal" in lib.dlx_last_error()
    # empty batches are valid no-ops
    g.has_focal = 0
    assert lib.dlx_transfer_matrix(g, 0, None, None, None, None, None, None) == 0
    assert lib.dlx_power(16, 0, None, None, None) == 0
    st = lib.dlx_fft2(16, 15, 1, 0, 1, 1, None, 0, None)
    assert st == 1
    st = lib.dlx_fft2(16, 24, 1, 0, 1, 1, None, 0, None)
    assert st == 4 and b"power-of-two" in lib.dlx_last_error()


def test_workspace_queries(lib):
    from dlux_amd import _lib
    g = _lib.MftGeom(n_in=512, n_out=256, inverse=0, has_focal=0, focal_length=0.0,
                     focal_shift=0.0)
    ws = lib.dlx_psf_poly_workspace_size(g, 64, 1)
    mat = 64 * 256 * 512 * 8
    assert ws >= 64 * 512 * 512 * 8 + 2 * mat + 64 * 256 * 256 * 4
    assert lib.dlx_mft_workspace_size(g, 4, 0) >= 3 * 4 * 256 * 512 * 8
    assert lib.dlx_fft2_workspace_size(1024, 2048, 1) >= 2048 * 2048 * 8


def test_batched_and_vjp_abi_validation_without_gpu(lib):
    from dlux_amd import _lib
    arr = (_lib.MftInstance * 2)()
    for e in arr:
        e.field, e.n_in, e.n_out, e.out_index = 4096, 64, 32, 0
        e.wavelength, e.pixel_scale_in, e.pixel_scale_out, e.weight = 1e-6, 1 / 64, 2.5e-7, 1.0
    outs = (ctypes.c_void_p * 1)(8192)
    arr[1].out_index = 1
    st = lib.dlx_mft_batched_psf(2, arr, 1, outs, None, 0, None)
    assert st == 1 and b"out_index" in lib.dlx_last_error()
    arr[1].out_index, arr[1].n_out = 0, 16
    st = lib.dlx_mft_batched_psf(2, arr, 1, outs, None, 0, None)
    assert st == 1 and b"different n_out" in lib.dlx_last_error()
    arr[1].n_out, arr[1].wavelength = 32, -1.0
    st = lib.dlx_mft_batched_psf(2, arr, 1, outs, None, 0, None)
    assert st == 1 and b"wavelength" in lib.dlx_last_error()
    arr[1].wavelength = 1e-6
    # two instances sharing one 64 x 64 field: one conversion, two matrices each
    need = lib.dlx_mft_batched_workspace_size(2, arr, 1)
    assert need >= 64 * 64 * 8 + 2 * 32 * 64 * 8 + 2 * 32 * 64 * 8 + 2 * 32 * 32 * 4
    st = lib.dlx_mft_batched_psf(2, arr, 1, outs, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()
    assert lib.dlx_mft_batched_psf(0, None, 0, None, None, 0, None) == 0

    g = _lib.MftGeom(n_in=64, n_out=32, inverse=0, has_focal=0, focal_length=0.0,
                     focal_shift=0.0)
    one = _lib.dptr([1e-6])
    w = _lib.dptr([1.0])
    st = lib.dlx_psf_poly_vjp(g, 1, one, w, 1 / 64, 2.5e-7, 0.0, 0.0, 0.0, None, None, None,
                              1.0, 4096, None, None, 0, None, 0, None)
    assert st == 1 and b"grad_opd or grad_phase" in lib.dlx_last_error()
    st = lib.dlx_psf_poly_vjp(g, 1, one, w, 1 / 64, 2.5e-7, 0.0, 0.0, 0.0, None, None, None,
                              1.0, 4096, 8192, None, 0, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()
    assert lib.dlx_psf_poly_vjp_workspace_size(g, 1, 1) > lib.dlx_psf_poly_workspace_size(g, 1, 1)
    assert lib.dlx_project_basis_workspace_size(36, 512 * 512, 1) >= 36 * 4
    st = lib.dlx_project_basis(3, 100, 1, 4096, 8192, 12288, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()


def test_wavefront_constructor_errors():
    import dlux_amd as dl
    with pytest.raises(ValueError, match="Provide one"):
        dl.Wavefront(1e-6, 16)
    with pytest.raises(ValueError, match="Cannot specify both"):
        dl.Wavefront(1e-6, 16, diameter=1.0, pixel_scale=0.1)


def test_optical_system_validation_errors():
    import dlux_amd as dl
    optics = dl.AngularOpticalSystem(16, 1.0, [], 8, 0.1)
    with pytest.raises(ValueError, match="Wavelength and weight shape mismatch"):
        optics.propagate([1e-6, 2e-6], weights=[1.0, 2.0, 3.0])
    with pytest.raises(ValueError, match=r"offset must be \[x, y\]"):
        optics.propagate([1e-6], offset=[0.0, 0.0, 0.0])
    with pytest.raises(ValueError, match="Cannot return both"):
        optics.propagate([1e-6], return_wf=True, return_psf=True)


def test_layer_validation_errors():
    import dlux_amd as dl
    with pytest.raises(ValueError, match="effect must be"):
        dl.BasisLayer(None, None, effect="bad")
    with pytest.raises(ValueError, match="coefficient shape"):
        dl.BasisLayer(np.zeros((3, 4, 4)), np.zeros(2))
    with pytest.raises(ValueError, match="angles must be"):
        dl.Tilt([1.0, 2.0, 3.0])
    with pytest.raises(ValueError, match="opd and phase"):
        dl.AberratedLayer(np.zeros((4, 4)), np.zeros((5, 5)))
    with pytest.raises(ValueError, match="transmission and opd"):
        dl.Optic(np.zeros((4, 4)), opd=np.zeros((5, 5)))
    with pytest.raises(ValueError, match="can not specify d or c"):
        dl.ASMPropagator(1.0, dl.CoordSpec(n=32, d=1.0))


def test_spectrum_and_sources():
    import dlux_amd as dl
    s = dl.Spectrum([1e-6, 2e-6], [1.0, 3.0])
    assert np.allclose(s.weights, [0.25, 0.75])
    s2 = dl.Spectrum([1e-6, 2e-6], np.ones((3, 2)))
    assert np.allclose(s2.weights.sum(-1), 1)
    with pytest.raises(ValueError):
        dl.Spectrum([1e-6, 2e-6], [1.0, 2.0, 3.0])
    src = dl.PointSources([1e-6], position=np.zeros((3, 2)), flux=[1.0, 2.0, 3.0])
    assert src.flux.shape == (3,)
    with pytest.raises(ValueError, match="Length of flux"):
        dl.PointSources([1e-6], position=np.zeros((3, 2)), flux=[1.0, 2.0])


def test_units_and_fft_spec():
    import dlux_amd.utils as dlu
    assert np.isclose(dlu.arcsec2rad(1.0), np.pi / 180 / 3600)
    assert np.isclose(dlu.unit_factor_to_rad("marcsec"), 1e-3 * np.pi / 180 / 3600)
    assert np.isclose(dlu.unit_factor_to_rad("mrad"), 1e-3)
    # like utils/units.py:71-90, a prefix applies to a canonical base only
    with pytest.raises(ValueError):
        dlu.unit_factor_to_rad("mas")
    with pytest.raises(ValueError):
        dlu.unit_factor_to_rad("furlong")
    d, c = dlu.fft_spec(16, 0.1, 1.0)
    assert c == -0.5 * d and dlu.fft_spec(15, 0.1, 1.0)[1] == 0.0
    assert np.isclose(dlu.calc_nfringes(1.0, 32, 0.1, 16, 0.05), 16 * 0.05 / (1.0 / 3.2))


def test_immutable_set():
    import dlux_amd as dl
    optics = dl.AngularOpticalSystem(16, 1.0, [], 8, 0.1)
    o2 = optics.set("psf_npixels", 12)
    assert optics.psf_npixels == 8 and o2.psf_npixels == 12
    o3 = optics.set(["psf_npixels", "oversample"], [10, 2])
    assert (o3.psf_npixels, o3.oversample) == (10, 2)


def test_workload_definitions():
    from dlux_amd import workloads as w
    assert np.isclose(w.POLY.mft_flops / 1e9, 51.54, atol=0.01)
    assert np.isclose(w.MONO.mft_flops / 1e9, 0.1007, atol=1e-3)
    assert len(w.POLY.js) == 36 and w.POLY.n_wl == 64
    c1, c2 = w.POLY.coefficients(), w.POLY.coefficients()
    assert np.array_equal(c1, c2)
