"""canary.jl_amd -- MI355X-native nodal-DG right-hand side and LSRK54 update.

Drop-in device path for Canary.jl's Euler (set2c, 2D/3D) and shallow-water
examples: `volumerhs`, `facerhs`, `updatesolution` with the reference
layouts, a fused single-pass stage, host preparation with the reference's
`Canary` module API, and a device time stepper.  All numerical work runs in
libcanary_dg.so (C++ host preparation + HIP kernels for gfx950).
"""
from . import _lib  # noqa: F401
from ._lib import EULER, SWE, FP64, FP32, CanaryError, default_phys  # noqa: F401


def built():
    import os
    return os.path.exists(_lib.LIBPATH)


def build(jobs=None, verbose=False):
    from .build import build as _b
    return _b(jobs, verbose)
