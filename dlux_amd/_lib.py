"""ctypes binding of ``libdlux_amd.so`` (the C ABI declared in ``include/dlux_amd.h``).

The library is built in-tree by ``dlux_amd/build.py``.  There is no fallback:
if the library is missing every compute entry point raises.
"""

from __future__ import annotations

import ctypes
from ctypes import c_float, POINTER, c_char_p, c_double, c_int32, c_int64, c_size_t, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_native" / "libdlux_amd.so"

DLX_OK = 0
DLX_PREC_FP32 = 0
DLX_PREC_BF16X3 = 1

_dp = POINTER(c_double)
_fp = c_void_p  # float* / void* device pointers are passed as integers


class MftGeom(ctypes.Structure):
    """``dlx_mft_geom``."""

    _fields_ = [
        ("n_in", c_int32),
        ("n_out", c_int32),
        ("inverse", c_int32),
        ("has_focal", c_int32),
        ("focal_length", c_double),
        ("focal_shift", c_double),
    ]


# name -> (restype, argtypes); must match include/dlux_amd.h
class MftInstance(ctypes.Structure):
    """``dlx_mft_instance``."""

    _fields_ = [
        ("field", c_void_p),
        ("n_in", c_int32),
        ("n_out", c_int32),
        ("out_index", c_int32),
        ("inverse", c_int32),
        ("has_focal", c_int32),
        ("reserved_", c_int32),
        ("focal_length", c_double),
        ("focal_shift", c_double),
        ("wavelength", c_double),
        ("pixel_scale_in", c_double),
        ("pixel_scale_out", c_double),
        ("shift_x", c_double),
        ("shift_y", c_double),
        ("weight", c_double),
    ]


# numpy view of dlx_mft_instance with the same field offsets (vectorised filling)
INSTANCE_DTYPE = np.dtype({
    "names": [f[0] for f in MftInstance._fields_],
    "formats": [np.uint64 if f[1] is c_void_p else (np.int32 if f[1] is c_int32 else np.float64)
                for f in MftInstance._fields_],
    "offsets": [getattr(MftInstance, f[0]).offset for f in MftInstance._fields_],
    "itemsize": ctypes.sizeof(MftInstance),
})


DLX_DET_PIXEL_RESPONSE = 1
DLX_DET_CONVOLVE = 2
DLX_DET_DOWNSAMPLE = 3
DLX_DET_SATURATE = 4
DLX_DET_ADD_CONSTANT = 5
DLX_DET_CONVOLVE_SEPARABLE = 6


class DetectorOp(ctypes.Structure):
    """``dlx_detector_op``."""

    _fields_ = [
        ("kind", c_int32),
        ("size", c_int32),
        ("value", c_double),
        ("data", c_void_p),
    ]


DLX_TF_TRANSLATION = 1
DLX_TF_SHEAR = 2
DLX_TF_COMPRESSION = 4
DLX_TF_ROTATION = 8

DLX_AP_CIRCLE = 1
DLX_AP_SQUARE = 2
DLX_AP_RECTANGLE = 3
DLX_AP_REG_POLYGON = 4
DLX_AP_SPIDER = 5


class CoordTransformC(ctypes.Structure):
    """``dlx_coord_transform``."""

    _fields_ = [
        ("flags", c_int32),
        ("reserved_", c_int32),
        ("translation", c_double * 2),
        ("shear", c_double * 2),
        ("compression", c_double * 2),
        ("rotation", c_double),
    ]


class SoftAperture(ctypes.Structure):
    """``dlx_soft_aperture``."""

    _fields_ = [
        ("kind", c_int32),
        ("occulting", c_int32),
        ("nsides", c_int32),
        ("n_angles", c_int32),
        ("n_transforms", c_int32),
        ("group", c_int32),
        ("size", c_double * 2),
        ("softening", c_double),
        ("angles_deg", c_double * 16),
        ("transforms", CoordTransformC * 3),
        ("coords", c_void_p),
    ]


SIGNATURES = {
    "dlx_version": (c_char_p, []),
    "dlx_last_error": (c_char_p, []),
    "dlx_transfer_matrix": (c_int32, [POINTER(MftGeom), c_int32, _dp, _dp, _dp, _dp, c_void_p,
                                      c_void_p]),
    "dlx_mft_workspace_size": (c_size_t, [POINTER(MftGeom), c_int32, c_int32]),
    "dlx_mft": (c_int32, [POINTER(MftGeom), c_int32, c_void_p, c_void_p, _dp, _dp, _dp, _dp,
                          _dp, _dp, c_int32, c_void_p, c_size_t, c_void_p]),
    "dlx_pupil_field": (c_int32, [c_int32, c_int32, _dp, c_double, c_double, c_double,
                                  c_double, _fp, _fp, _fp, c_double, c_void_p, c_void_p,
                                  c_void_p]),
    "dlx_eval_basis": (c_int32, [c_int32, c_int64, c_int32, _fp, _fp, _fp, _fp, c_void_p]),
    "dlx_power": (c_int32, [c_int64, c_int32, c_void_p, _fp, c_void_p]),
    "dlx_intensity": (c_int32, [c_int64, c_int32, c_void_p, c_int32, _fp, c_void_p]),
    "dlx_psf_poly_workspace_size": (c_size_t, [POINTER(MftGeom), c_int32, c_int32]),
    "dlx_psf_poly": (c_int32, [POINTER(MftGeom), c_int32, _dp, _dp, c_double, c_double,
                               c_double, c_double, c_double, _fp, _fp, _fp, c_double, _fp,
                               c_int32, c_void_p, c_size_t, c_void_p]),
    "dlx_psf_poly_noalloc_scratch_bytes": (c_size_t, []),
    "dlx_psf_poly_noalloc": (c_int32, [POINTER(MftGeom), c_int32, _dp, _dp, c_double,
                                       c_double, c_double, c_double, c_double, _fp, _fp, _fp,
                                       c_double, _fp, c_int32, c_void_p]),
    "dlx_psf_poly_vjp_workspace_size": (c_size_t, [POINTER(MftGeom), c_int32, c_int32]),
    "dlx_psf_poly_vjp": (c_int32, [POINTER(MftGeom), c_int32, _dp, _dp, c_double, c_double,
                                   c_double, c_double, c_double, _fp, _fp, _fp, c_double, _fp,
                                   _fp, _fp, _fp, c_int32, c_void_p, c_size_t, c_void_p]),
    "dlx_project_basis_workspace_size": (c_size_t, [c_int32, c_int64, c_int32]),
    "dlx_project_basis": (c_int32, [c_int32, c_int64, c_int32, _fp, _fp, _fp, c_void_p,
                                    c_size_t, c_void_p]),
    "dlx_detector_out_side": (c_int32, [c_int32, POINTER(DetectorOp), c_int32,
                                        POINTER(c_int32)]),
    "dlx_detector_workspace_size": (c_size_t, [c_int32, c_int32, POINTER(DetectorOp), c_int32,
                                               c_int32]),
    "dlx_detector_apply": (c_int32, [c_int32, c_int32, _fp, POINTER(DetectorOp), c_int32, _fp,
                                     c_void_p, c_size_t, c_void_p]),
    "dlx_psf_poly_detector_workspace_size": (c_size_t, [POINTER(MftGeom), c_int32, c_int32,
                                                        POINTER(DetectorOp), c_int32]),
    "dlx_psf_poly_detector": (c_int32, [POINTER(MftGeom), c_int32, _dp, _dp, c_double,
                                        c_double, c_double, c_double, c_double, _fp, _fp, _fp,
                                        c_double, POINTER(DetectorOp), c_int32, _fp, c_int32,
                                        c_void_p, c_size_t, c_void_p]),
    "dlx_detector_vjp": (c_int32, [c_int32, c_int32, _fp, POINTER(DetectorOp), c_int32, _fp,
                                   _fp, c_void_p, c_size_t, c_void_p]),
    "dlx_detector_vjp_params_workspace_size": (c_size_t, [c_int32, c_int32, POINTER(DetectorOp),
                                                          c_int32]),
    "dlx_detector_vjp_params": (c_int32, [c_int32, c_int32, _fp, POINTER(DetectorOp), c_int32,
                                          _fp, _fp, POINTER(c_void_p), c_void_p, c_size_t,
                                          c_void_p]),
    "dlx_soft_apertures_workspace_size": (c_size_t, [c_int32, c_int32]),
    "dlx_soft_apertures": (c_int32, [c_int32, c_double, c_double, c_int32,
                                     POINTER(SoftAperture), _fp, c_void_p, c_size_t,
                                     c_void_p]),
    "dlx_lct_workspace_size": (c_size_t, [c_int32, c_int32, c_int32, c_int32]),
    "dlx_lct": (c_int32, [c_int32, c_int32, c_int32, POINTER(c_double), POINTER(c_double),
                          c_double, c_double, c_double, c_double, _fp, _fp, c_int32, c_void_p,
                          c_size_t, c_void_p]),
    "dlx_lct_fft_workspace_size": (c_size_t, [c_int32, c_int32, c_int32]),
    "dlx_lct_fft": (c_int32, [c_int32, c_int32, c_int32, POINTER(c_double), POINTER(c_double),
                              c_double, c_double, c_double, _fp, _fp, POINTER(c_double),
                              c_void_p, c_size_t, c_void_p]),
    "dlx_soft_apertures_vjp_workspace_size": (c_size_t, [c_int32, c_int32]),
    "dlx_soft_apertures_vjp": (c_int32, [c_int32, c_double, c_double, c_int32,
                                         POINTER(SoftAperture), _fp, _fp, c_void_p, c_size_t,
                                         c_void_p]),
    "dlx_interp_transform": (c_int32, [c_int32, c_double, c_double, c_int32, c_double, c_double,
                                       c_int32, POINTER(CoordTransformC), c_int32, c_float,
                                       c_int32, _fp, _fp, c_void_p]),
    "dlx_interp_points": (c_int32, [c_int32, c_double, c_double, c_int64, _fp, _fp, c_int32,
                                    c_float, c_int32, _fp, _fp, c_void_p]),
    "dlx_interp_transform_vjp": (c_int32, [c_int32, c_double, c_double, c_int32, c_double,
                                           c_double, c_int32, POINTER(CoordTransformC), c_int32,
                                           c_int32, _fp, _fp, _fp, c_void_p]),
    "dlx_interp_points_vjp": (c_int32, [c_int32, c_double, c_double, c_int64, _fp, _fp,
                                        c_int32, c_int32, _fp, _fp, _fp, c_void_p]),
    "dlx_aperture_basis": (c_int32, [c_int32, c_double, c_double, c_int32,
                                     POINTER(CoordTransformC), c_void_p, c_double, c_int32,
                                     c_int32, POINTER(c_int32), _fp, c_void_p]),
    "dlx_mft_batched_workspace_size": (c_size_t, [c_int32, POINTER(MftInstance), c_int32]),
    "dlx_batched_graphs": (c_int32, [c_int32]),
    "dlx_batched_graph_stats": (c_int32, [POINTER(ctypes.c_uint64), POINTER(ctypes.c_uint64),
                                          POINTER(c_int32)]),
    "dlx_mft_batched_psf": (c_int32, [c_int32, POINTER(MftInstance), c_int32,
                                      POINTER(c_void_p), c_void_p, c_size_t, c_void_p]),
    "dlx_noll_indices": (c_int32, [c_int32, POINTER(c_int32), POINTER(c_int32)]),
    "dlx_zernike_basis": (c_int32, [c_int32, POINTER(c_int32), c_int64, _fp, _fp, c_double,
                                    _fp, c_void_p]),
    "dlx_circular_aperture": (c_int32, [c_int32, c_double, c_int32, c_double, c_double,
                                        c_double, c_int32, _dp, _fp, _fp, c_void_p]),
    "dlx_fft2_workspace_size": (c_size_t, [c_int32, c_int32, c_int32]),
    "dlx_fft2": (c_int32, [c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p,
                           c_size_t, c_void_p]),
    "dlx_asm_workspace_size": (c_size_t, [c_int32, c_int32, c_int32]),
    "dlx_asm_prop": (c_int32, [c_int32, c_int32, c_int32, _dp, c_double, c_double, c_void_p,
                               c_void_p, c_void_p, c_size_t, c_void_p]),
    "dlx_cache_clear": (c_int32, []),
    "dlx_cache_set_budget": (c_int32, [c_size_t]),
    "dlx_cache_bypass": (c_int32, [c_int32]),
    "dlx_mfma_rate": (c_int32, [c_int32, POINTER(c_double), c_void_p, c_size_t, c_void_p]),
    "dlx_cache_stats": (c_int32, [POINTER(ctypes.c_uint64), POINTER(ctypes.c_uint64),
                                  POINTER(c_size_t), POINTER(c_int32)]),
    "dlx_profile_enable": (c_int32, [c_int32]),
    "dlx_profile_read": (c_int32, [c_int32, POINTER(c_double), POINTER(c_int32)]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the native library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"dlux_amd native library not found at {LIB_PATH}; build it with "
            "`python -m dlux_amd.build` (hipcc, gfx950).  There is no CPU fallback."
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, name: str) -> None:
    """Raise RuntimeError carrying dlx_last_error() for a non-zero status."""
    if status != DLX_OK:
        msg = load().dlx_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed (status {status}): {msg}")


def dptr(values) -> ctypes.Array:
    """A host double array for the C ABI."""
    vals = [float(v) for v in values]
    return (c_double * len(vals))(*vals)
