"""sc_2022_amd -- MI355X (gfx950) matrix-free SBP-SAT 2-D Poisson solver.

The compute path is libsbp.so (hand-written HIP kernels behind the C ABI of
include/sbp.h); this package is its host side.  Build with
``python -m sc_2022_amd.build``.
"""
from . import _capi
from ._capi import SbpError, SbpParams, default_params, load

__all__ = ["SbpError", "SbpParams", "default_params", "load", "SbpPoisson", "reference_api"]


def __getattr__(name):
    if name == "SbpPoisson":
        from .solver import SbpPoisson
        return SbpPoisson
    if name == "reference_api":
        import importlib
        return importlib.import_module(__name__ + ".reference_api")
    raise AttributeError(name)
