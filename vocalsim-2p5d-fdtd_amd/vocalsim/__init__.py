"""vocalsim: MI355X-native 2.5D FDTD vocal-tract solver (Python side).

The solver itself is libvocalsim.so (hand-written HIP kernels for gfx950 and
a C ABI, include/vocalsim_abi.h).  This package is a thin ctypes binding plus
workload generators and spectral analysis helpers.
"""
from . import tubes
from ._abi import VocalSimError, check, load
from .solver import Solver


def release_cached_memory() -> None:
    """Frees the device memory the library keeps after a handle is closed
    for the next handle to take (vs_release_cached_memory)."""
    check(load().vs_release_cached_memory())


__all__ = ["tubes", "Solver", "VocalSimError", "load", "release_cached_memory"]
