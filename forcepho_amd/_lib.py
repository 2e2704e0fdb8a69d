"""ctypes binding of libfpx.so (declarations: include/forcepho_amd.h).

There is no fallback: if the library is missing or fails to load, every
device entry point raises.  Host-only helpers (struct sizes, tile planning)
work without a GPU.
"""

import ctypes
import os

import numpy as np

from .build import LIB_PATH

__all__ = ["lib", "check", "FpxError", "TILE_DTYPE", "SLOT_DTYPE", "SLOT_MODEL_ONLY",
           "STATUS_LIMIT", "STATUS_RADIUS_BIN", "STATUS_SPLIT_WAIT", "MAX_TILE_PARTS",
           "STRUCT_PATCH", "STRUCT_SOURCE", "STRUCT_RESPONSE", "STRUCT_PSFGAUSS",
           "STRUCT_TILE", "STRUCT_SLOT", "STRUCT_AMPMODEL", "EXPORTED_SYMBOLS",
           "AmplitudeModel", "AMP_LOGNORMAL", "AMP_SPLINE", "MAX_SPLINE_DEGREE"]

(STRUCT_PATCH, STRUCT_SOURCE, STRUCT_RESPONSE, STRUCT_PSFGAUSS, STRUCT_TILE, STRUCT_SLOT,
 STRUCT_AMPMODEL) = range(7)

# FpxTile / FpxSlot as numpy dtypes (include/forcepho_amd.h)
TILE_DTYPE = np.dtype([("exposure", np.int32), ("band", np.int16), ("part", np.uint8),
                       ("parts", np.uint8), ("pix_start", np.int32), ("pix_count", np.int32)])
SLOT_DTYPE = np.dtype([("patch", np.uintp), ("proposal", np.uintp), ("tiles", np.uintp),
                       ("pixel_weight", np.uintp), ("pchi2", np.uintp),
                       ("pdchi2_dp", np.uintp), ("accum", np.uintp), ("status", np.uintp),
                       ("n_bands", np.int32), ("n_tiles", np.int32),
                       ("flags", np.int32), ("reserved", np.int32), ("split", np.uintp)],
                      align=True)

SLOT_MODEL_ONLY = 1    # FpxSlot.flags: model pass only (FPX_SLOT_MODEL_ONLY)
STATUS_LIMIT = 1       # FpxSlot.status bit: a tile broke a size limit (FPX_STATUS_LIMIT)
STATUS_RADIUS_BIN = 2  # FpxSlot.status bit: radius bin outside [0, n_radii)
STATUS_SPLIT_WAIT = 4  # FpxSlot.status bit: a split tile's piece timed out waiting
MAX_TILE_PARTS = 16    # FPX_MAX_TILE_PARTS: pieces of one split tile

AMP_LOGNORMAL, AMP_SPLINE = 0, 1   # FpxAmplitudeModel.kind
MAX_SPLINE_DEGREE = 5              # FPX_MAX_SPLINE_DEGREE


class AmplitudeModel(ctypes.Structure):
    """FpxAmplitudeModel (include/forcepho_amd.h); array fields hold device
    addresses."""
    _fields_ = [("kind", ctypes.c_int32), ("n_radii", ctypes.c_int32),
                ("lnr", ctypes.c_void_p), ("lnr_count", ctypes.c_void_p),
                ("kx", ctypes.c_int32), ("ky", ctypes.c_int32),
                ("tx_stride", ctypes.c_int32), ("ty_stride", ctypes.c_int32),
                ("cx", ctypes.c_int32), ("cy", ctypes.c_int32),
                ("tx", ctypes.c_void_p), ("ty", ctypes.c_void_p),
                ("tx_len", ctypes.c_void_p), ("ty_len", ctypes.c_void_p),
                ("coef", ctypes.c_void_p)]

_c_int, _c_size, _p = ctypes.c_int32, ctypes.c_size_t, ctypes.c_void_p

# name -> (restype, argtypes); every symbol include/forcepho_amd.h declares
_SIGNATURES = {
    "fpx_abi_version": (_c_int, []),
    "fpx_last_error": (ctypes.c_char_p, []),
    "fpx_struct_size": (_c_size, [_c_int]),
    "fpx_accum_doubles": (_c_size, [_c_int]),
    "fpx_split_bytes": (_c_size, [_c_int, _c_int]),
    "fpx_plan_tiles": (_c_int, [_p, _p, _p, _p, _c_int, _c_int, _p, _c_int]),
    "fpx_n_configs": (_c_int, []),
    "fpx_config_info": (_c_int, [_c_int, _p, _p, _p]),
    "fpx_lds_bytes": (_c_int, [_c_int, _c_int, _p]),
    "fpx_init": (_c_int, [_c_int]),
    "fpx_device_info": (_c_int, [_p, _p, _p]),
    "fpx_evaluate_proposal": (_c_int, [_p, _c_int, _c_int, _c_int, _p]),
    "fpx_evaluate_sync": (_c_int, [_p, _c_int, _c_int, _c_int, _c_int, _p, _p, _p, _c_size,
                                   _p, _p, _c_size, _p, _p]),
    "fpx_bind": (_c_int, [_p, _c_int, _c_int, _c_int, _p]),
    "fpx_bind_tiles": (_c_int, [_p, _c_int, _c_int, _c_int, _p, _c_int, _p]),
    "fpx_evaluate": (_c_int, [_p, _p, _p, _p, _p]),
    "fpx_evaluate_vjp": (_c_int, [_p, _p, _p, _p, _p, _p]),
    "fpx_evaluate_model": (_c_int, [_p, _p, _p, _p, _p]),
    "fpx_evaluate_host": (_c_int, [_p, _p, _c_int, _p, _p, _p]),
    "fpx_evaluator_status": (_c_int, [_p, _p]),
    "fpx_evaluator_info": (_c_int, [_p, _p, _p, _p]),
    "fpx_release": (_c_int, [_p]),
    "fpx_evaluate_static": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _p, _p, _p, _c_int,
                                     _c_int, _p]),
    "fpx_evaluate_batch": (_c_int, [_p, _c_int, _p, _c_int, _c_int, _c_int, _c_int, _p, _p]),
    "fpx_evaluate_tiles": (_c_int, [_p, _c_int, _p, _c_int, _c_int, _c_int, _c_int, _p, _p]),
    "fpx_finalize": (_c_int, [_p, _c_int, _p]),
    "fpx_count_pairs": (_c_int, [_p, _c_int, _p, _c_int, _c_int, _p, _p, _p]),
    "fpx_resident_groups": (_c_int, [_c_int, _c_int, _p]),
    "fpx_build_sources": (_c_int, [_p, _p, _p, _c_int, _c_int, _c_int, _p, _p, _p, _p, _p]),
    "fpx_leapfrog_update": (_c_int, [_p, _p, _p, _p, _p, _p, _c_int, _c_int, _c_int,
                                     ctypes.c_double, _c_int, _p, _p, _p, _p, _p]),
    "fpx_stack_gradients": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _p, _p, _p]),
    "fpx_hmc_metropolis": (_c_int, [_p, _p, _p, _p, _p, _p, _c_int, _c_int, _p, _p, _p]),
    "fpx_probe_valu": (_c_int, [_c_int, _c_int, _p, _c_int, _p, _p]),
}
EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_LIB = None


class FpxError(RuntimeError):
    pass


def lib():
    """Load libfpx.so once and return the ctypes handle."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FpxError(f"{LIB_PATH} is not built; run forcepho_amd.build.build_library() "
                           "(or __graft_entry__.build()) first")
        # torch (the host layer's device memory and streams) ships its own HIP
        # runtime under the same soname as the system one libfpx.so links: load
        # torch first so that the library binds to that runtime instead of
        # bringing in a second one (torch then finds no device)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def check(rc, what=""):
    """Raise FpxError for a negative return code of a C entry point."""
    if rc < 0:
        msg = lib().fpx_last_error().decode(errors="replace")
        raise FpxError(f"{what} failed ({rc}): {msg}")
    return rc
