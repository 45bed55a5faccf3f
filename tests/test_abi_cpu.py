"""CPU-side checks of the C ABI library: it loads, exports every symbol the
header declares, and rejects bad input on the host (no GPU compute calls)."""
import ctypes
import os
import re

import pytest

import vocalsim
from vocalsim import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vocalsim_abi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(vs_\w+)\s*\(", text, flags=re.M)))


def test_library_loads_and_exports_all_declared_symbols():
    L = vocalsim.load()
    syms = declared_symbols()
    assert len(syms) >= 19
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_abi.EXPORTS)


def test_abi_version():
    assert vocalsim.load().vs_abi_version() == 2


def test_null_arguments_rejected_without_device():
    L = vocalsim.load()
    h = ctypes.c_void_p()
    assert L.vs_create(None, 1, 64, ctypes.byref(h)) == -1
    assert b"specs" in L.vs_last_error()
    assert L.vs_run(None, 10) == -1
    assert L.vs_adjoint_forward(None, 512, None, 0) == -1
    assert L.vs_adjoint_backward(None, None, 0, None, None, None, 0) == -1
    assert L.vs_destroy(None) == 0
    # nothing cached yet: releasing makes no device call
    assert L.vs_release_cached_memory() == 0


def test_device_trace_buffers_are_checked():
    """The adjoint's device-memory trace arguments must be contiguous float64
    GPU tensors of the right size; anything else is refused before any call."""
    import torch
    from vocalsim.solver import _check_device_f64

    for bad in (torch.zeros(10, dtype=torch.float64), torch.zeros(10, dtype=torch.float32),
                torch.zeros(20, dtype=torch.float64)[::2], [0.0] * 10):
        with pytest.raises(ValueError):
            _check_device_f64(bad, 10, "probe output")
