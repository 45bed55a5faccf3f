"""vocalsim: MI355X-native 2.5D FDTD vocal-tract solver (Python side).

The solver itself is libvocalsim.so (hand-written HIP kernels for gfx950 and
a C ABI, include/vocalsim_abi.h).  This package is a thin ctypes binding plus
workload generators and spectral analysis helpers.
"""
from . import tubes
from ._abi import VocalSimError, load
from .solver import Solver

__all__ = ["tubes", "Solver", "VocalSimError", "load"]
