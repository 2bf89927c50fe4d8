"""ctypes binding of libsbp.so (the C ABI declared in include/sbp.h).

The library is loaded from the package directory.  There is no fallback: if
libsbp.so is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsbp.so"

SBP_OK = 0
SBP_ERR_INVALID_SIZE = 1
SBP_ERR_INVALID_ORDER = 2
SBP_ERR_INVALID_ARG = 3
SBP_ERR_HIP = 4
SBP_ERR_NONFINITE = 5
SBP_ERR_BREAKDOWN = 6
SBP_ERR_NOT_CONVERGED = 7
SBP_ERR_NO_DEVICE = 8
SBP_ERR_VENDOR = 9
SBP_ERR_SCRATCH = 10

SBP_F64 = 0
SBP_F32 = 1
SBP_FUSED = 2   # or-ed into sbp_set_tiles' dtype: shape of the fused two-sweep kernels
SBP_ALIGNED = 4   # or-ed into sbp_set_tiles' dtype: the line-aligned variant of a stencil tile shape

# Every exported entry point: name -> (restype, argtypes)
_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)


class SbpParams(C.Structure):
    _fields_ = [
        ("nu_pre", C.c_int),
        ("nu_post", C.c_int),
        ("omega", C.c_double),
        ("coarse_n", C.c_int),
        ("restrict_scale", C.c_double),
        ("smoother", C.c_int),
        ("richardson_omega", C.c_double),
        ("poll_every", C.c_int),
        ("inner_rtol", C.c_double),
        ("inner_maxiter", C.c_int),
        ("max_outer", C.c_int),
        ("tile_i", C.c_int),
        ("tile_j", C.c_int),
        ("use_graph", C.c_int),
        ("ir_rtol", C.c_double),
        ("bottom_n", C.c_int),
        ("fuse_n", C.c_int),
        ("drop_coarse_rhs", C.c_int),
    ]


SIGNATURES = {
    "sbp_version": (C.c_char_p, []),
    "sbp_status_string": (C.c_char_p, [_i]),
    "sbp_default_params": (None, [C.POINTER(SbpParams)]),
    "sbp_create": (_i, [_i, _i, C.POINTER(SbpParams), C.POINTER(_vp)]),
    "sbp_create_batch": (_i, [_i, _i, _i, C.POINTER(SbpParams), C.POINTER(_vp)]),
    "sbp_create_ragged": (_i, [_i, _i, _pi, C.POINTER(SbpParams), C.POINTER(_vp)]),
    "sbp_get_batch": (_i, [_vp, _pi]),
    "sbp_destroy": (_i, [_vp]),
    "sbp_set_stream": (_i, [_vp, _vp]),
    "sbp_get_info": (_i, [_vp, _pi, _pi, _pi]),
    "sbp_set_tiles": (_i, [_vp, _i, _i, _i, _i]),
    "sbp_apply": (_i, [_vp, _i, _vp, _vp, _i, _i]),
    "sbp_apply_dot": (_i, [_vp, _vp, _vp, _pd]),
    "sbp_apply_dot_noalloc": (_i, [_vp, _vp, _vp, _vp]),
    "sbp_diag": (_i, [_vp, _vp]),
    "sbp_assemble_rhs": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sbp_rhs_transpose": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sbp_hnorm": (_i, [_vp, _vp, _pd]),
    "sbp_h_weights": (_i, [_vp, _pd]),
    "sbp_dot": (_i, [_vp, _vp, _vp, _pd]),
    "sbp_restrict": (_i, [_vp, _i, _vp, _vp, _i]),
    "sbp_prolong_add": (_i, [_vp, _i, _vp, _vp, _i]),
    "sbp_smooth": (_i, [_vp, _i, _vp, _vp, _i, _i]),
    "sbp_coarse_solve": (_i, [_vp, _vp, _vp]),
    "sbp_vcycle": (_i, [_vp, _vp, _vp, _i]),
    "sbp_cg": (_i, [_vp, _vp, _vp, _d, _i, _pi, _pd, _vp]),
    "sbp_mgpcg": (_i, [_vp, _vp, _vp, _d, _i, _pi, _pd, _vp]),
    "sbp_mgpcg_mixed": (_i, [_vp, _vp, _vp, _d, _i, _pi, _pi, _pd]),
    "sbp_cg_interp": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _d, _i, _pi, _pd, _vp]),
    "sbp_create_var": (_i, [_i, _i, _pd, _pd, C.POINTER(SbpParams), C.POINTER(_vp)]),
    "sbp_mgpcg_batch": (_i, [_vp, _vp, _vp, _pd, _i, _pi, _pd, _pd, _pi]),
    "sbp_mgpcg_mixed_batch": (_i, [_vp, _vp, _vp, _pd, _i, _pi, _pi, _pd, _pd, _pi]),
    "sbp_adjoint_grad": (_i, [_vp, _vp, _vp, _d, _i, _pd, _vp, _vp, _vp, _vp, _vp, _vp, _pi]),
    "sbp_csr_assemble": (_i, [_vp, _vp, _vp, _vp, C.POINTER(C.c_int64)]),
    "sbp_csr_spmv": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "sbp_csr_cg": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _d, _i, _pi, _pd]),
    "sbp_csr_vendor_attach": (_i, [_vp, _vp, _vp, _vp, C.c_int64, _i]),
    "sbp_csr_vendor_spmv": (_i, [_vp, _vp, _vp]),
    "sbp_csr_vendor_cg": (_i, [_vp, _vp, _vp, _d, _i, _pi, _pd]),
    "sbp_copy": (_i, [_vp, _vp, _vp, C.c_size_t]),
}


class SbpError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = _LIB.sbp_status_string(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{where}: sbp status {status} ({msg})")


_LIB = None


def load(path: os.PathLike | None = None):
    """Loads libsbp.so and binds every entry point (raises if absent)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python -m sc_2022_amd.build); "
            "there is no CPU fallback")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(status, where=""):
    if status != SBP_OK:
        raise SbpError(status, where)
    return status


def default_params(**overrides) -> SbpParams:
    lib = load()
    prm = SbpParams()
    lib.sbp_default_params(C.byref(prm))
    for k, v in overrides.items():
        if not hasattr(prm, k):
            raise KeyError(f"unknown solver parameter {k!r}")
        setattr(prm, k, v)
    return prm
