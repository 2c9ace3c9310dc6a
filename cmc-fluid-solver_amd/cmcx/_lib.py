"""ctypes binding of libcmcx.so (the C ABI declared in include/cmcx.h).

The library is built in-tree by cmc-fluid-solver_amd/build.py.  There is no
fallback: if the library is missing, loading fails with an error.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# CMCX_LIB overrides the library file (another build of the same sources).
LIB_PATH = os.environ.get("CMCX_LIB") or os.path.join(os.path.dirname(HERE), "lib", "libcmcx.so")

CMCX_OK = 0
CMCX_DIVERGED = 1
CMCX_ERR_ARG = -1
CMCX_ERR_MASK = -2
CMCX_ERR_HIP = -3
CMCX_ERR_STATE = -4
CMCX_ERR_UNSUPPORTED = -5
CMCX_ERR_TOO_LARGE = -6
CMCX_ERR_PICARD = -7

CMCX_F32, CMCX_F64 = 0, 1
CMCX_PATH_PLAIN, CMCX_PATH_FUSED = 0, 1
CMCX_MERGE_CPU, CMCX_MERGE_GPU_ORDER = 0, 1
CMCX_SHAPE3D, CMCX_SHAPE2D = 3, 2

_u8p = ctypes.POINTER(ctypes.c_uint8)
_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


class cmcx_domain(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32 * 3), ("h", ctypes.c_double * 3), ("base_t", ctypes.c_double),
                ("type", _u8p), ("bc_vel", _u8p), ("bc_temp", _u8p), ("v", _dp * 3), ("T", _dp)]


class cmcx2d_domain(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32 * 2), ("h", ctypes.c_double * 2), ("start_t", ctypes.c_double),
                ("type", _u8p), ("bc", _u8p), ("v", _dp * 2), ("T", _dp)]


class cmcx_coards_header(ctypes.Structure):
    _fields_ = [("bbox_min", ctypes.c_double * 3), ("bbox_max", ctypes.c_double * 3),
                ("timestep", ctypes.c_double), ("time", ctypes.c_double),
                ("outdims", ctypes.c_int32 * 3), ("vars", ctypes.c_char_p),
                ("xy_degree_units", ctypes.c_int), ("depth", ctypes.POINTER(ctypes.c_float))]


class cmcx_params(ctypes.Structure):
    _fields_ = [("v_T", ctypes.c_double), ("v_vis", ctypes.c_double), ("t_vis", ctypes.c_double),
                ("t_phi", ctypes.c_double)]


# (name, restype, argtypes)
SIGNATURES = [
    ("cmcx_version", ctypes.c_char_p, []),
    ("cmcx_last_error", ctypes.c_char_p, []),
    ("cmcx_params_normalized", None, [ctypes.c_double] * 3 + [ctypes.POINTER(cmcx_params)]),
    ("cmcx_params_dimensional", None, [ctypes.c_double] * 5 + [ctypes.POINTER(cmcx_params)]),
    ("cmcx_create", ctypes.c_int, [ctypes.POINTER(cmcx_domain), ctypes.POINTER(cmcx_params), ctypes.c_int,
                                   _vp, ctypes.POINTER(_vp)]),
    ("cmcx_destroy", ctypes.c_int, [_vp]),
    ("cmcx_set_options", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    ("cmcx_set_nblockz", ctypes.c_int, [_vp, ctypes.c_int]),
    ("cmcx_build_segments", ctypes.c_int, [_vp]),
    ("cmcx_num_segments", ctypes.c_int, [_vp, ctypes.c_int, _i64p]),
    ("cmcx_get_segments", ctypes.c_int, [_vp, ctypes.c_int, _i32p, ctypes.c_int64]),
    ("cmcx_get_segment_nodes", ctypes.c_int, [_vp, ctypes.c_int, _u8p, _dp, ctypes.c_int64]),
    ("cmcx_update_boundaries", ctypes.c_int, [_vp]),
    ("cmcx_step", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    ("cmcx_run", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    ("cmcx_div_error", ctypes.c_int, [_vp, _dp]),
    ("cmcx_get_fields", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("cmcx_get_fields_device", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("cmcx_set_fields", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("cmcx_state_device_ptrs", ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    ("cmcx_get_layer", ctypes.c_int, [_vp, _vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("cmcx_set_boundary", ctypes.c_int, [_vp, _u8p, _u8p, _dp, _dp, _dp, _dp]),
    ("cmcx_profile", ctypes.c_int, [_vp, ctypes.c_int]),
    ("cmcx_profile_read", ctypes.c_int, [_vp, _dp, _i64p, ctypes.c_int]),
    ("cmcx_step_bytes", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    ("cmcx_div_error_fields", ctypes.c_int, [_vp, _vp, _vp, _vp, _i32p, _dp, ctypes.c_int, _vp, _dp]),
    ("cmcx_autotune", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    ("cmcx_set_tile_variant", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    ("cmcx_synchronize", ctypes.c_int, [_vp]),
    ("cmcx_create_batch", ctypes.c_int, [ctypes.POINTER(cmcx_domain), ctypes.POINTER(cmcx_params),
                                         ctypes.c_int, ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    ("cmcx_batch_size", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    ("cmcx_select", ctypes.c_int, [_vp, ctypes.c_int]),
    ("cmcx_grid_errors", ctypes.c_int, [_vp, _dp]),
    ("cmcx_set_graph", ctypes.c_int, [_vp, ctypes.c_int]),
    ("cmcx_set_params", ctypes.c_int, [_vp, ctypes.POINTER(cmcx_params)]),
    ("cmcx_adjoint_forward", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("cmcx_adjoint_backward", ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                             ctypes.POINTER(_vp), _dp]),
    ("cmcx_adjoint_backward_device", ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                    ctypes.POINTER(_vp), _dp]),
    ("cmcx_adjoint_release", ctypes.c_int, [_vp]),
    ("cmcx_adjoint_generation", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    ("cmcx_geometry_load", ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, _dp, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("cmcx_geometry_info", ctypes.c_int, [_vp, _i32p, _dp, _dp, ctypes.POINTER(ctypes.c_int), _dp]),
    ("cmcx_geometry_build", ctypes.c_int, [_vp, ctypes.c_double, _u8p, _u8p, _u8p, _dp, _dp, _dp, _dp]),
    ("cmcx_geometry_build2d", ctypes.c_int, [_vp, ctypes.c_double, _u8p, _u8p, _dp, _dp, _dp]),
    ("cmcx_geometry_destroy", ctypes.c_int, [_vp]),
    ("cmcx_coards_write_header", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(cmcx_coards_header)]),
    ("cmcx_coards_write_layer", ctypes.c_int, [ctypes.c_char_p, _vp, ctypes.c_int, _dp, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]),
    ("cmcx2d_create", ctypes.c_int, [ctypes.POINTER(cmcx2d_domain), ctypes.POINTER(cmcx_params), ctypes.c_int,
                                     _vp, ctypes.POINTER(_vp)]),
    ("cmcx2d_update_boundaries", ctypes.c_int, [_vp]),
    ("cmcx2d_step", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp,
                                   ctypes.POINTER(ctypes.c_int)]),
    ("cmcx2d_get_fields", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("cmcx2d_destroy", ctypes.c_int, [_vp]),
    ("cmcx2d_create_method", ctypes.c_int, [ctypes.POINTER(cmcx2d_domain), ctypes.POINTER(cmcx_params),
                                            ctypes.c_int, ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    ("cmcx2d_set_poisson", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int64]),
    ("cmcx2d_set_grid", ctypes.c_int, [_vp, _u8p, _u8p, _dp, _dp, _dp]),
    ("cmcx2d_poisson_sweeps", ctypes.c_int, [_vp, _i32p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
]

_lib = None


def lib():
    """Load libcmcx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libcmcx.so not found at {LIB_PATH}: build it with "
                           "`python cmc-fluid-solver_amd/build.py` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class CmcxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cmcx error {code}: {msg}")
        self.code = code


class DivergenceError(CmcxError):
    """The divergence guard fired (the reference throws runtime_error, AdiSolver3D.cpp:371-374)."""


def check(rc: int, what: str = ""):
    if rc == CMCX_OK:
        return rc
    msg = lib().cmcx_last_error().decode(errors="replace")
    if rc == CMCX_DIVERGED:
        raise DivergenceError(rc, msg or "divergence error above threshold")
    raise CmcxError(rc, f"{what}: {msg}")
