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


def test_psf_poly_detector_abi_validation_without_gpu(lib):
    """The fused PSF + detector entry point validates the chain against the PSF
    side and its workspace before any device work."""
    from dlux_amd import _lib
    g = _lib.MftGeom(n_in=64, n_out=30, inverse=0, has_focal=0, focal_length=0.0,
                     focal_shift=0.0)
    ops = (_lib.DetectorOp * 1)()
    ops[0].kind, ops[0].size = _lib.DLX_DET_DOWNSAMPLE, 4
    one = _lib.dptr([1e-6])
    st = lib.dlx_psf_poly_detector(g, 1, one, _lib.dptr([1.0]), 1 / 64, 2.5e-7, 0.0, 0.0, 0.0,
                                   None, None, None, 1.0, ops, 1, 4096, 1, None, 0, None)
    assert st == 1 and b"divisible" in lib.dlx_last_error()
    ops[0].size = 3
    need = lib.dlx_psf_poly_detector_workspace_size(g, 1, 1, ops, 1)
    assert need >= lib.dlx_psf_poly_workspace_size(g, 1, 1)
    st = lib.dlx_psf_poly_detector(g, 1, one, _lib.dptr([1.0]), 1 / 64, 2.5e-7, 0.0, 0.0, 0.0,
                                   None, None, None, 1.0, ops, 1, 4096, 1, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()


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
                              1.0, 4096, None, None, None, 0, None, 0, None)
    assert st == 1 and b"grad_opd, grad_phase or grad_amp" in lib.dlx_last_error()
    st = lib.dlx_psf_poly_vjp(g, 1, one, w, 1 / 64, 2.5e-7, 0.0, 0.0, 0.0, None, None, None,
                              1.0, 4096, 8192, None, None, 0, None, 0, None)
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


def test_detector_abi_plans_without_gpu(lib):
    """dlx_detector_out_side / workspace_size plan the chain on the host; the
    entry points reject what the reference's layers reject."""
    from dlux_amd import _lib

    def ops(*specs):
        arr = (_lib.DetectorOp * max(len(specs), 1))()
        for e, (kind, size, value, data) in zip(arr, specs):
            e.kind, e.size, e.value, e.data = kind, size, value, data
        return arr, len(specs)

    side = ctypes.c_int32()
    a, n = ops((_lib.DLX_DET_CONVOLVE, 9, 0.0, 4096), (_lib.DLX_DET_PIXEL_RESPONSE, 64, 0.0, 8192),
               (_lib.DLX_DET_DOWNSAMPLE, 4, 0.0, None), (_lib.DLX_DET_SATURATE, 0, 1e4, None),
               (_lib.DLX_DET_ADD_CONSTANT, 0, 3.0, None), (_lib.DLX_DET_DOWNSAMPLE, 2, 0.0, None))
    assert lib.dlx_detector_out_side(64, a, n, ctypes.byref(side)) == 0 and side.value == 8
    # two structural passes: one intermediate image pair of 64x64 floats per batch
    assert lib.dlx_detector_workspace_size(64, 3, a, n, 0) >= 2 * 3 * 64 * 64 * 4
    assert (lib.dlx_detector_workspace_size(64, 3, a, n, 1)
            > lib.dlx_detector_workspace_size(64, 3, a, n, 0))
    a, n = ops((_lib.DLX_DET_DOWNSAMPLE, 5, 0.0, None),)
    assert lib.dlx_detector_out_side(64, a, n, ctypes.byref(side)) == 1
    assert b"not divisible by 5" in lib.dlx_last_error()
    a, n = ops((_lib.DLX_DET_DOWNSAMPLE, 2, 0.0, None), (_lib.DLX_DET_PIXEL_RESPONSE, 64, 0.0, 8))
    assert lib.dlx_detector_out_side(64, a, n, ctypes.byref(side)) == 1
    assert b"32x32" in lib.dlx_last_error()
    a, n = ops((_lib.DLX_DET_CONVOLVE, 65, 0.0, 8),)
    assert lib.dlx_detector_out_side(64, a, n, ctypes.byref(side)) == 1
    a, n = ops((9, 1, 0.0, None),)
    assert lib.dlx_detector_out_side(64, a, n, ctypes.byref(side)) == 1
    assert b"unknown" in lib.dlx_last_error()
    # an empty chain is the identity; empty batches are no-ops
    assert lib.dlx_detector_out_side(7, None, 0, ctypes.byref(side)) == 0 and side.value == 7
    assert lib.dlx_detector_apply(7, 0, None, None, 0, None, None, 0, None) == 0
    a, n = ops((_lib.DLX_DET_CONVOLVE, 3, 0.0, 8), (_lib.DLX_DET_DOWNSAMPLE, 2, 0.0, None))
    st = lib.dlx_detector_apply(16, 1, 4096, a, n, 8192, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()
    st = lib.dlx_detector_vjp(16, 1, 4096, a, n, 8192, 12288, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()


def test_detector_layer_construction():
    """tests/layers/test_detector_layers.py:21-51 and tests/test_detectors.py:10-31."""
    import dlux_amd as dl
    from dlux_amd import layers
    with pytest.raises(ValueError):
        layers.ApplyPixelResponse(np.ones(16))
    jit = layers.ApplyJitter(0.1, 5)
    assert jit.sigma.shape == () and jit.kernel.shape == (5, 5)
    with pytest.raises(ValueError):
        layers.ApplyJitter(0.1, 0)
    assert layers.ApplySaturation(1e4).threshold.shape == ()
    assert layers.AddConstant(1).value.shape == ()
    with pytest.raises(ValueError):
        layers.Downsample(0)
    det = dl.LayeredDetector([layers.Downsample(4)])
    assert isinstance(det.Downsample, layers.Downsample) and det.kernel_size == 4
    with pytest.raises(AttributeError):
        det.not_an_attr
    inserted = det.insert_layer(layers.Downsample(4), 1)
    assert list(inserted.layers.keys()) == ["Downsample_0", "Downsample_1"]
    removed = inserted.remove_layer("Downsample_0")
    assert list(removed.layers.keys()) == ["Downsample_1"]
    assert list(det.layers.keys()) == ["Downsample"]


def test_jitter_kernel_matches_oracle():
    from dlux_amd import layers
    from dlux_amd.utils import gaussian
    from oracle import dlux_oracle as orc
    assert np.allclose(gaussian(0.0, [0.7, 1.3], 11), orc.gaussian(0.0, [0.7, 1.3], 11),
                       rtol=1e-6)
    for sigma, k, os_ in [(0.5, 9, 3), (1.7, 5, 1), (0.1, 7, 4)]:
        ref = orc.jitter_kernel(sigma, k, os_)
        layer = layers.ApplyJitter(sigma, k, os_)
        got = layer.kernel
        assert got.dtype == np.float32 and np.allclose(got, ref, rtol=1e-6, atol=1e-9)
        # the separable form the HIP chain runs is the same kernel
        f = layer._factor()
        assert np.max(np.abs(np.outer(f, f) - ref)) < 2e-7 * ref.max()


def test_binary_source_and_scene_host_logic():
    """tests/test_sources.py:76-129 (construction, weights, keys) and
    utils/source.py:10-52."""
    import dlux_amd as dl
    from dlux_amd import sources as S
    from dlux_amd.utils import fluxes_from_contrast, positions_from_sep
    wl = np.ones(2)
    src = dl.BinarySource(wl, separation=1.0)
    assert src.mean_flux.shape == () and src.separation.shape == ()
    assert src.position_angle.shape == () and src.contrast.shape == ()
    with pytest.raises(ValueError):
        dl.BinarySource(wl, separation=1.0, position=np.ones(1))
    src = dl.BinarySource(wl, weights=np.ones((2, 2)))
    assert src.weights.shape == (2, 2) and np.allclose(src.weights.sum(-1), 1)
    # fluxes average to the mean and keep the contrast; positions straddle the centre
    f = fluxes_from_contrast(10.0, 4.0)
    assert np.allclose(f, [16.0, 4.0]) and np.isclose(f.mean(), 10.0)
    p = positions_from_sep(np.array([1.0, 2.0]), 2.0, np.pi / 2)
    assert np.allclose(p, [[2.0, 2.0], [0.0, 2.0]])
    p = positions_from_sep(np.zeros(2), 2.0, 0.0)
    assert np.allclose(p, [[0.0, 1.0], [0.0, -1.0]])
    (wls, pos, w), = dl.BinarySource([1e-6, 2e-6], mean_flux=3.0, contrast=2.0,
                                     separation=1e-6)._point_terms()
    assert pos.shape == (2, 2) and np.allclose(w.sum(-1), [4.0, 2.0])

    keyed = dl.Scene([("source", dl.PointSource([1e-6]))])
    assert isinstance(keyed.source, dl.PointSource)
    assert isinstance(dl.Scene(dl.PointSource([1e-6])).PointSource, dl.PointSource)
    tup = dl.Scene((dl.PointSource([1e-6]), dl.PointSource([1e-6])))
    assert isinstance(tup.PointSource_0, dl.PointSource)
    with pytest.raises(AttributeError):
        tup.not_an_attr
    assert len(tup._point_terms()) == 2

    assert S._as_wavelengths_1d(None) is None
    with pytest.raises(ValueError):
        S._as_wavelengths_1d(np.ones((2, 2)))
    assert S._infer_n_wavelengths(np.ones(3), None) == 3
    assert S._infer_n_wavelengths(None, dl.Spectrum([1e-6])) == 1
    with pytest.raises(ValueError):
        S._infer_n_wavelengths(None, None)
    with pytest.raises(TypeError):
        S._infer_n_wavelengths(None, 1)
    with pytest.raises(TypeError):
        dl.PointSource([1e-6], spectrum=1)
    with pytest.raises(ValueError, match="Cannot return both"):
        dl.PointSources([1e-6]).model(None, return_wf=True, return_psf=True)


def test_instrument_construction():
    """instruments.py:65-114 and :211-229 validation."""
    import dlux_amd as dl
    from dlux_amd import layers
    optics = dl.LayeredOpticalSystem(16, 1.0, [dl.Optic()])
    with pytest.raises(TypeError, match="OpticalSystem"):
        dl.Telescope(1, dl.PointSource([1e-6]))
    with pytest.raises(TypeError, match="source tuple"):
        dl.Telescope(optics, ("a", 1))
    with pytest.raises(TypeError, match="Detector"):
        dl.Telescope(optics, dl.PointSource([1e-6]), detector=1)
    tel = dl.Telescope(optics, [dl.PointSource([1e-6]), ("b", dl.PointSource([1e-6]))],
                       dl.LayeredDetector([layers.Downsample(2)]))
    assert isinstance(tel.source, dl.Scene) and list(tel.source.sources) == ["PointSource", "b"]
    assert tel.wf_npixels == 16 and tel.kernel_size == 2
    with pytest.raises(ValueError, match="ndithers"):
        dl.Dither(np.zeros(2), optics, dl.PointSource([1e-6]))
    d = dl.Dither(np.zeros((3, 2)), optics, dl.PointSource([1e-6]))
    moved = d.add("source.position", np.ones(2))
    assert np.allclose(moved.source.position, 1.0) and np.allclose(d.source.position, 0.0)


def test_resolved_source_construction():
    """tests/test_sources.py:69-101."""
    import dlux_amd as dl
    wl = np.ones(2)
    src = dl.ResolvedSource(wl)
    assert np.isclose(src.distribution.sum(), 1.0)
    with pytest.raises(ValueError):
        dl.ResolvedSource(wl, distribution=np.ones(1))
    # normalise floors negative intensities before renormalising (sources.py:475-488)
    d = np.array([[1.0, -1.0], [2.0, 1.0]])
    n = dl.ResolvedSource(wl, distribution=d).normalise()
    assert np.allclose(n.distribution, np.array([[1.0, 0.0], [2.0, 1.0]]) / 4.0)
    prs = dl.PointResolvedSource(wl)
    assert prs.contrast.shape == ()
    prs = dl.PointResolvedSource(wl, weights=np.ones((2, 2)))
    assert prs.weights.shape == (2, 2) and np.allclose(prs.weights.sum(-1), 1)
    with pytest.raises(NotImplementedError, match="Wavefront information"):
        prs.model(None, return_wf=True)
    with pytest.raises(NotImplementedError):
        src._point_terms()


def test_dynamic_aperture_construction():
    """tests/layers/test_apertures.py:56-245 (construction, attributes, errors)."""
    import dlux_amd as dl
    from dlux_amd import layers as L
    ap = L.CircularAperture(0.1, dl.CoordTransform(translation=[1.0, 1.0]), False, 1.0, True)
    assert ap.softness.shape == () and ap.radius.shape == ()
    assert np.allclose(ap.translation, [1.0, 1.0])
    with pytest.raises(AttributeError):
        ap.not_an_attr
    with pytest.raises(TypeError):
        L.CircularAperture(0.1, transformation=2)
    with pytest.raises(ValueError):
        L.CircularAperture(0.1, softening=0.0)
    assert L.CircularAperture(0.1).nsides == 0 and np.isclose(L.CircularAperture(0.1).extent, 0.1)
    sq = L.SquareAperture(0.2)
    assert sq.nsides == 4 and np.isclose(sq.extent, np.sqrt(2) * 0.2)
    rect = L.RectangularAperture(0.2, 0.4)
    assert rect.height.shape == () and np.isclose(rect.extent, np.hypot(0.1, 0.2))
    poly = L.RegPolyAperture(6, 0.1)
    assert poly.nsides == 6 and poly.rmax.shape == ()
    sp = L.Spider(0.1, np.linspace(0, 360, 6, endpoint=False))
    assert sp.occulting and sp.width.shape == ()
    with pytest.raises(TypeError):
        sp.extent
    with pytest.raises(TypeError):
        sp.nsides
    with pytest.raises(TypeError):
        L.AberratedAperture(L.CircularAperture(0.1, occulting=True), [1, 2])
    with pytest.raises(TypeError):
        L.AberratedAperture(L.Spider(0.1, [0.0], occulting=False), [1, 2])
    with pytest.raises(TypeError):
        L.AberratedAperture(None, [1, 2])
    with pytest.raises(NotImplementedError):
        L.AberratedAperture(L.CircularAperture(0.1), [1, 2]).calculate()
    with pytest.raises(ValueError, match="effect"):
        L.AberratedAperture(L.CircularAperture(0.1), [1, 2], effect="invalid")
    ab = L.AberratedAperture(L.CircularAperture(1.0), [1, 2, 3, 4])
    comp = L.CompoundAperture([L.CircularAperture(0.1), ab])
    assert isinstance(comp.CircularAperture, L.CircularAperture)
    with pytest.raises(AttributeError):
        comp.not_an_attr
    with pytest.raises(TypeError):
        L.CompoundAperture([ab, ab])
    multi = L.MultiAperture([comp, L.CircularAperture(0.05)])
    assert len(multi._aberrated_apertures()) == 1
    assert dl.CoordTransform()._struct().flags == 0
    with pytest.raises(ValueError, match="translation"):
        dl.CoordTransform(translation=[1.0])


def test_aperture_abi_validation_without_gpu(lib):
    from dlux_amd import _lib
    aps = (_lib.SoftAperture * 2)()
    aps[0].kind, aps[0].softening, aps[0].size[0] = _lib.DLX_AP_CIRCLE, 1.0, 0.3
    aps[1].kind, aps[1].softening, aps[1].nsides = _lib.DLX_AP_REG_POLYGON, 1.0, 2
    st = lib.dlx_soft_apertures(64, 1 / 64, 0.0, 2, aps, 4096, None, 0, None)
    assert st == 1 and b"3 to 64 sides" in lib.dlx_last_error()
    aps[1].nsides, aps[1].softening = 6, 0.0
    st = lib.dlx_soft_apertures(64, 1 / 64, 0.0, 2, aps, 4096, None, 0, None)
    assert st == 1 and b"softening" in lib.dlx_last_error()
    aps[1].softening, aps[1].kind = 1.0, 9
    st = lib.dlx_soft_apertures(64, 1 / 64, 0.0, 2, aps, 4096, None, 0, None)
    assert st == 1 and b"unknown aperture kind" in lib.dlx_last_error()
    aps[1].kind = _lib.DLX_AP_SPIDER
    assert lib.dlx_soft_apertures(64, 1 / 64, 0.0, 2, aps, 4096, None, 0, None) == 1
    aps[1].n_angles = 3
    need = lib.dlx_soft_apertures_workspace_size(64, 2)
    assert need >= 2 * 64 * 64 * 8
    st = lib.dlx_soft_apertures(64, 1 / 64, 0.0, 2, aps, 4096, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()
    js = (ctypes.c_int32 * 2)(1, 2)
    st = lib.dlx_aperture_basis(64, 1 / 64, 0.0, 0, None, None, 1.0, 2, 2, js, 4096, None)
    assert st == 1 and b"nsides" in lib.dlx_last_error()
    st = lib.dlx_aperture_basis(64, 1 / 64, 0.0, 4, None, None, 1.0, 0, 2, js, 4096, None)
    assert st == 1 and b"at most 3" in lib.dlx_last_error()


def test_soft_apertures_vjp_abi_validation_without_gpu(lib):
    from dlux_amd import _lib
    aps = (_lib.SoftAperture * 1)()
    aps[0].kind, aps[0].softening, aps[0].size[0] = _lib.DLX_AP_CIRCLE, 1.0, 0.3
    need = lib.dlx_soft_apertures_vjp_workspace_size(64, 1)
    assert need > lib.dlx_soft_apertures_workspace_size(64, 1)
    st = lib.dlx_soft_apertures_vjp(64, 1 / 64, 0.0, 1, aps, None, 4096, None, 0, None)
    assert st == 1 and b"null argument" in lib.dlx_last_error()
    st = lib.dlx_soft_apertures_vjp(64, 1 / 64, 0.0, 1, aps, 4096, 4096, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()
    aps[0].softening = -1.0
    st = lib.dlx_soft_apertures_vjp(64, 1 / 64, 0.0, 1, aps, 4096, 4096, None, need, None)
    assert st == 1 and b"softening" in lib.dlx_last_error()


def test_geometry_gradient_sources():
    """Which tensors the geometry VJP feeds: sizes and the innermost
    transformation of each aperture; the rest raise NotImplementedError."""
    torch = pytest.importorskip("torch")
    import dlux_amd as dl
    from dlux_amd import layers as L
    from dlux_amd.layers.apertures import _grad_sources
    r = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    t = torch.tensor([0.01, 0.02], dtype=torch.float64, requires_grad=True)
    rot = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    ap = L.CircularAperture(r, dl.CoordTransform(translation=t, rotation=rot))
    d = ap._descs(())
    assert d[0][1].size[0] == pytest.approx(0.3) and d[0][1].transforms[0].rotation == \
        pytest.approx(0.1)
    params, smap = _grad_sources(d)
    assert len(params) == 3
    assert sorted(s[1] for s in smap) == [0, 2, 3, 8]
    # a member without its own transform differentiates the enclosing one
    comp = L.CompoundAperture([L.CircularAperture(0.4), L.SquareAperture(0.2)],
                              transformation=dl.CoordTransform(translation=t))
    params, smap = _grad_sources(comp._descs(()))
    assert len(params) == 1 and [(a, s) for a, s, _, _ in smap] == [(0, 2), (0, 3), (1, 2),
                                                                    (1, 3)]
    # ... but not under a member's own transformation
    comp = L.CompoundAperture([L.CircularAperture(0.4, dl.CoordTransform(rotation=0.2))],
                              transformation=dl.CoordTransform(translation=t))
    with pytest.raises(NotImplementedError):
        comp._descs(())
    with torch.no_grad():
        comp._descs(())
    with pytest.raises(NotImplementedError):
        L.Spider(0.02, torch.tensor([0.0, 90.0], requires_grad=True))._descs(())
    # keys are plain floats whatever the parameter type
    assert ap._param_key() == L.CircularAperture(
        0.3, dl.CoordTransform(translation=[0.01, 0.02], rotation=0.1))._param_key()


def test_polynomial_construction_and_errors():
    """tests/test_polynomials.py construction and error cases that need no
    evaluation."""
    import dlux_amd as dl
    import dlux_amd.utils as dlu
    z = dl.DynamicZernike(4)
    assert z.j == 4 and z.name == dlu.zernike_name(4) == "Defocus" and (z.n, z.m) == (2, 0)
    with pytest.raises(ValueError, match="greater than 0"):
        dl.DynamicZernike(0)
    assert dlu.radial_orders_to_indices([2]) == [4, 5, 6]
    assert dlu.radial_orders_to_indices([0, 1]) == [1, 2, 3]
    assert dlu.zernike_name(36) == "Heptafoil Y" and dlu.zernike_name(37) == "Zernike 37"
    coords = np.zeros((2, 8, 8))
    with pytest.raises(ValueError, match="exactly one"):
        dl.ZernikeBasis(coords)
    with pytest.raises(ValueError, match="exactly one"):
        dl.ZernikeBasis(coords, js=[1], radial_orders=[0])
    with pytest.raises(ValueError, match="At least one"):
        dl.ZernikeBasis(coords, js=[])
    with pytest.raises(ValueError, match="greater than zero"):
        dl.ZernikeBasis(coords, js=[0])
    with pytest.raises(ValueError, match="non-negative"):
        dl.ZernikeBasis(coords, radial_orders=[-1])
    b = dl.DynamicZernikeBasis(radial_orders=[2])
    assert b.coefficient_shape == (3,) and b.js == [4, 5, 6]
    with pytest.raises(ValueError, match="wavefront or coordinates"):
        dl.DynamicZernikeBasis(js=[1]).evaluate()
    for nsides in (-1, 1, 2):
        with pytest.raises(ValueError, match="zero or greater"):
            dl.DynamicZernikeBasis(js=[1], nsides=nsides)
    with pytest.raises(ValueError, match="Coefficient shape"):
        dl.DynamicZernikeBasis(js=[4, 5], coefficients=np.ones(3))
    # solve_basis: least squares over the trailing dimensions
    rng = np.random.default_rng(0)
    basis = rng.normal(size=(3, 5, 5))
    c = np.array([1.0, -2.0, 0.5])
    assert np.allclose(dlu.solve_basis(np.tensordot(c, basis, 1), basis), c)
    with pytest.raises(ValueError, match="trailing dimensions"):
        dlu.solve_basis(np.ones((4, 4)), basis)
    # AberratedAperture keeps its Zernikes as a DynamicZernikeBasis
    from dlux_amd import layers as L
    ab = L.AberratedAperture(L.RegPolyAperture(6, 0.4), [2, 3], np.array([1.0, 2.0]))
    assert isinstance(ab.aberrations, dl.DynamicZernikeBasis) and ab.aberrations.nsides == 6
    assert np.allclose(ab.set("coefficients", np.array([3.0, 4.0])).coefficients, [3, 4])
    assert np.allclose(ab.coefficients, [1, 2])


def test_abcd_elements_and_propagators_host():
    """tests/layers/test_abcd_propagators.py construction cases."""
    import dlux_amd as dl
    from dlux_amd import layers as L
    from dlux_amd.coordinates import CoordSpec
    els = [L.ABCDFreeSpace(1.0), L.ABCDLens(2.0), L.ABCDMirror(3.0), L.ABCDConjugatePlane(4.0)]
    for el in els:
        assert el.abcd.shape == (2, 2) and abs(np.linalg.det(el.abcd) - 1) < 1e-12
    assert np.allclose(L.ABCDFreeSpace(1.5).abcd, [[1, 1.5], [0, 1]])
    assert np.allclose(L.ABCDLens(2.0).abcd, [[1, 0], [-0.5, 1]])
    assert np.allclose(L.ABCDMirror(3.0).abcd, [[1, 0], [-2 / 3, 1]])
    assert np.allclose(L.ABCDConjugatePlane(4.0).abcd, [[0, 4], [-0.25, 0]])
    # f - lens - f is the conjugate plane
    prop = L.MFTPropagator([L.ABCDFreeSpace(2.0), L.ABCDLens(2.0), ("back", L.ABCDFreeSpace(2.0))],
                           CoordSpec(n=8, d=1 / 32, c=0.0))
    assert np.allclose(prop.abcd, L.ABCDConjugatePlane(2.0).abcd)
    assert prop.n == 8 and np.allclose(prop.focal_length, 2.0)
    assert isinstance(prop.ABCDLens, L.ABCDLens) and isinstance(prop.back, L.ABCDFreeSpace)
    with pytest.raises(AttributeError):
        prop.not_an_attr
    with pytest.raises(TypeError):
        L.MFTPropagator([L.ABCDFreeSpace(1.0), 3.0], CoordSpec(n=8, d=1.0, c=0.0))
    with pytest.raises(ValueError, match="can not specify d"):
        L.FFTPropagator([L.ABCDFreeSpace(1.0)], CoordSpec(n=8, d=1.0, c=0.0))
    with pytest.raises(NotImplementedError):
        L.Fraunhofer()
    with pytest.raises(NotImplementedError):
        L.Fresnel()
    assert dl.layers.ABCDPropagator is L.ABCDPropagator


def test_lct_abi_validation_without_gpu(lib):
    abcd = (ctypes.c_double * 4)(1.0, 0.0, 0.0, 1.0)
    wl = (ctypes.c_double * 1)(1e-6)
    st = lib.dlx_lct(16, 8, 1, abcd, wl, 1e-3, 0.0, 1e-3, 0.0, 4096, 4096, 0, None, 0, None)
    assert st == 1 and b"non-zero B" in lib.dlx_last_error()
    abcd[1] = 0.5
    st = lib.dlx_lct(16, 8, 1, abcd, wl, 1e-3, 0.0, 1e-3, 0.0, 4096, 4096, 0, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()
    assert lib.dlx_lct_workspace_size(16, 8, 1, 0) >= 2 * 8 * 16 * 8
    assert lib.dlx_lct_workspace_size(16, 8, 1, 1) > lib.dlx_lct_workspace_size(16, 8, 1, 0)
    st = lib.dlx_lct(16, 8, 1, abcd, wl, -1e-3, 0.0, 1e-3, 0.0, 4096, 4096, 0, None, 0, None)
    assert st == 1 and b"pixel scales" in lib.dlx_last_error()
    st = lib.dlx_lct_fft(16, 31, 1, abcd, wl, 1e-3, 0.0, 0.0, 4096, 4096, None, None, 0, None)
    assert st == 1 and b"m - n even" in lib.dlx_last_error()
    need = lib.dlx_lct_fft_workspace_size(16, 32, 1)
    assert need >= 16 * 16 * 8 + 32 * 32 * 8
    st = lib.dlx_lct_fft(16, 32, 1, abcd, wl, 1e-3, 0.0, 0.0, 4096, 4096, None, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()


def test_detector_param_vjp_abi_validation_without_gpu(lib):
    from dlux_amd import _lib
    ops = (_lib.DetectorOp * 3)()
    ops[0].kind, ops[0].size, ops[0].data = _lib.DLX_DET_CONVOLVE, 3, 4096
    ops[1].kind, ops[1].size = _lib.DLX_DET_DOWNSAMPLE, 2
    ops[2].kind, ops[2].value = _lib.DLX_DET_ADD_CONSTANT, 1.0
    ptrs = (ctypes.c_void_p * 3)(None, 4096, None)
    st = lib.dlx_detector_vjp_params(16, 1, 4096, ops, 3, 4096, 4096, ptrs, None, 0, None)
    assert st == 1 and b"downsample factor has no gradient" in lib.dlx_last_error()
    ptrs[0], ptrs[1], ptrs[2] = 4096, None, 8192
    need = lib.dlx_detector_vjp_params_workspace_size(16, 1, ops, 3)
    assert need > lib.dlx_detector_workspace_size(16, 1, ops, 3, 1)
    st = lib.dlx_detector_vjp_params(16, 1, 4096, ops, 3, 4096, 4096, ptrs, None, 0, None)
    assert st == 3 and b"workspace" in lib.dlx_last_error()
    st = lib.dlx_detector_vjp_params(16, 1, 4096, ops, 3, 4096, 4096, None, None, 0, None)
    assert st == 1 and b"null parameter-gradient list" in lib.dlx_last_error()


def test_spectra_classes():
    """spectra.py: BaseSpectrum / SimpleSpectrum / Spectrum / PolySpectrum."""
    import dlux_amd as dl
    wl = np.linspace(1e-6, 2e-6, 5)
    ps = dl.PolySpectrum(wl, [1.0, 2e5])
    ref = 1.0 + 2e5 * wl
    assert np.allclose(ps.weights, ref / ref.sum()) and ps.normalise() is ps
    with pytest.raises(ValueError, match="1d"):
        dl.PolySpectrum(wl, np.ones((2, 2)))
    with pytest.raises(ValueError, match="1d array"):
        dl.Spectrum(np.ones((2, 2)))
    assert isinstance(ps, dl.BaseSpectrum) and isinstance(dl.Spectrum(wl), dl.SimpleSpectrum)
    src = dl.PointSource(spectrum=ps)
    assert np.allclose(src.spectrum.weights, ps.weights)


def test_distorted_coords_host():
    """DistortedCoords: shapes and errors; the FP64 coordinate planes that the
    aperture and resampling kernels read (a CoordTransform then a distortion)
    against the oracle's chain on the CPU."""
    torch = pytest.importorskip("torch")
    import dlux_amd as dl
    from oracle import dlux_oracle as orc
    from dlux_amd.coordinates import _chain_coords
    assert dl.DistortedCoords(2).powers.tolist() == [[1, 0, 2, 1, 0], [0, 1, 0, 1, 2]]
    with pytest.raises(ValueError):
        dl.DistortedCoords(1, np.zeros((2, 5)))
    D = np.array([[0.02, -0.01, 0.3, 0.1, -0.2], [0.01, 0.03, -0.1, 0.25, 0.15]])
    tf = dict(translation=[0.03, -0.02], rotation=0.3, compression=[1.1, 0.9],
              shear=[0.05, -0.02])
    chain = (dl.CoordTransform(**tf), dl.DistortedCoords(2, D))
    got = _chain_coords(chain, 24, 0.05, 0.01, "cpu").numpy()
    xs = orc.spec_xs(24, 0.05, 0.01)
    ref = np.array(np.meshgrid(xs, xs, indexing="xy"))
    ref = orc.coord_transform(orc.coord_transform(ref, **tf), distortion=D, order=2)
    assert np.max(np.abs(got - ref)) < 1e-14
    # the host call form (float32 for the CoordTransform) matches the planes
    c = torch.as_tensor(np.array(np.meshgrid(xs, xs, indexing="xy")))
    out = chain[1](chain[0](c))
    assert np.max(np.abs(np.asarray(out) - ref)) < 1e-6
    with pytest.raises(TypeError):
        _chain_coords((object(),), 4, 1.0, 0.0, "cpu")


def test_interp_vjp_abi_validation_without_gpu(lib):
    """dlx_interp_transform_vjp / dlx_interp_points_vjp reject bad arguments
    before any launch; an empty batch is a no-op."""
    from dlux_amd import _lib
    tf = (_lib.CoordTransformC * 1)()
    assert lib.dlx_interp_transform_vjp(0, 1.0, 0.0, 4, 1.0, 0.0, 0, tf, 1, 1, None, None, None,
                                        None) != 0
    assert lib.dlx_interp_transform_vjp(4, 1.0, 0.0, 4, 1.0, 0.0, 0, tf, 2, 1, None, None, None,
                                        None) != 0
    assert lib.dlx_interp_transform_vjp(4, 1.0, 0.0, 4, 1.0, 0.0, 0, tf, 1, 1, None, None, None,
                                        None) != 0
    assert lib.dlx_interp_transform_vjp(4, 1.0, 0.0, 4, 1.0, 0.0, 0, tf, 1, 0, None, None, None,
                                        None) == 0
    assert lib.dlx_interp_points_vjp(4, -1.0, 0.0, 3, None, None, 1, 1, None, None, None,
                                     None) != 0
    assert lib.dlx_interp_points_vjp(4, 1.0, 0.0, 3, None, None, 1, 1, None, None, None,
                                     None) != 0
    assert lib.dlx_interp_points_vjp(4, 1.0, 0.0, 3, None, None, 1, 0, None, None, None,
                                     None) == 0
