"""ctypes binding of libvocalsim.so (the C ABI declared in include/vocalsim_abi.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C vocalsim-2p5d-fdtd_amd``).  There is no fallback: if the library is
missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libvocalsim.so")

VS_OK = 0
VS_FP32, VS_FP64 = 32, 64

F_CELL_TYPE, F_BOUNDARY_SEG, F_DEPTH_P, F_DEPTH_X, F_DEPTH_Y, F_Z_INV, F_SOURCE = range(7)
F_UPPER_WALL, F_LOWER_WALL, F_SEGMENT, F_PRESSURE, F_VX, F_VY, F_CELL_CLASS = range(7, 14)
F_COEFF0 = 100
COEFFS = ["N_out", "N_in", "minVxBeta", "minVyBeta", "maxVxSigmaPrimedt",
          "maxVySigmaPrimedt", "betaVxSqr_dt_inv_rho_inv_ds", "betaVySqr_dt_inv_rho_inv_ds",
          "pressureSigmaPrimedt", "areWeNotExcitationVx", "areWeNotExcitationVy",
          "excitationWeightVx", "excitationWeightVy", "xor_val1", "xor_val2", "xor_val3", "xor_val4"]

(OPT_STEPS_PER_LAUNCH, OPT_THREADS, OPT_USE_GRAPH, OPT_KERNEL, OPT_TIMING, OPT_ENGINE_VARIANT,
 OPT_ADJOINT_KERNEL, OPT_ADJOINT_BLOCK, OPT_EXACT_GROUPS, OPT_EXACT_STEPS, OPT_BARRIER_SPIN,
 OPT_FAST_WIDE, OPT_FAST_LAYOUT_K, OPT_FAST_PAIR_CLASSES, OPT_FAST_ROW_QUANTUM) = range(15)
F_AS_F32 = 0x1000
K_EXACT, K_FAST = 0, 1

EXPORTS = [
    "vs_last_error", "vs_abi_version", "vs_create", "vs_destroy", "vs_num_tubes", "vs_get_dims",
    "vs_get_field", "vs_run", "vs_sync", "vs_steps_done", "vs_reset", "vs_get_probe",
    "vs_get_source_trace", "vs_set_option", "vs_get_option", "vs_field_ptr", "vs_stream",
    "vs_last_run_timing", "vs_cells_per_step", "vs_dtft", "vs_probe_dtft", "vs_autotune",
    "vs_adjoint_forward", "vs_adjoint_backward", "vs_probe_audio", "vs_butter2_lowpass", "vs_write_wav",
    "vs_copy_to_host", "vs_measure_lds_bandwidth", "vs_dtft_noalloc", "vs_release_cached_memory",
    "vs_adjoint_timing",
]

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)


class TubeSpecC(ctypes.Structure):
    _fields_ = [
        ("fdtd_solver_type", ctypes.c_int), ("cross_sectional_shape", ctypes.c_int),
        ("mouth_radiation_cond", ctypes.c_int), ("srate_mul", ctypes.c_int),
        ("pml_flag", ctypes.c_int), ("source_type", ctypes.c_int),
        ("num_steps", ctypes.c_int), ("audio_dur", ctypes.c_float),
        ("num_sections", ctypes.c_int),
        ("length_cm", _F), ("area_cm2", _F), ("aspect", _F), ("depth_scale", _F), ("mu3d", _F),
        ("mu3d_default", ctypes.c_float),
        ("gauss_amp", ctypes.c_double), ("gauss_t0", ctypes.c_double), ("gauss_width", ctypes.c_double),
        ("tm_ps", ctypes.c_double), ("tm_q", ctypes.c_double),
        ("engine_variant", ctypes.c_int),
        ("open_space", ctypes.c_int),
    ]


class Dims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "frameH", "frameW", "domainH", "domainW", "listenerX", "listenerY",
        "excitationX", "excitationY", "num_steps", "tube_start_x", "tube_end_x",
        "tube_len_cells", "midY", "num_sections", "mic_cells")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class VocalSimError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load():
    """Loads libvocalsim.so and declares the ABI.  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VocalSimError(-3, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_void, c_size = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t
    L.vs_last_error.restype = ctypes.c_char_p
    L.vs_abi_version.restype = c_int
    L.vs_create.argtypes = [ctypes.POINTER(TubeSpecC), c_int, c_int, ctypes.POINTER(c_void)]
    L.vs_destroy.argtypes = [c_void]
    L.vs_num_tubes.argtypes = [c_void, ctypes.POINTER(c_int)]
    L.vs_get_dims.argtypes = [c_void, c_int, ctypes.POINTER(Dims)]
    L.vs_get_field.argtypes = [c_void, c_int, c_int, c_void, c_size]
    L.vs_run.argtypes = [c_void, c_int]
    L.vs_sync.argtypes = [c_void]
    L.vs_steps_done.argtypes = [c_void, ctypes.POINTER(c_int)]
    L.vs_reset.argtypes = [c_void]
    L.vs_get_probe.argtypes = [c_void, c_int, _D, c_size]
    L.vs_get_source_trace.argtypes = [c_void, c_int, _D, c_size]
    L.vs_set_option.argtypes = [c_void, c_int, c_int]
    L.vs_get_option.argtypes = [c_void, c_int, ctypes.POINTER(c_int)]
    L.vs_field_ptr.argtypes = [c_void, c_int, c_int, ctypes.POINTER(c_void), ctypes.POINTER(c_size)]
    L.vs_stream.argtypes = [c_void, ctypes.POINTER(c_void)]
    L.vs_last_run_timing.argtypes = [c_void, _D, ctypes.POINTER(c_int), _D]
    L.vs_cells_per_step.argtypes = [c_void, ctypes.POINTER(ctypes.c_longlong)]
    L.vs_dtft.argtypes = [_D, c_int, ctypes.c_double, c_int, _D, c_int, _D, _D]
    L.vs_probe_dtft.argtypes = [c_void, c_int, c_int, c_int, _D, c_int, _D, _D]
    L.vs_autotune.argtypes = [c_void, c_int, ctypes.c_char_p, c_size]
    L.vs_adjoint_forward.argtypes = [c_void, c_int, _D, c_size]
    L.vs_adjoint_backward.argtypes = [c_void, _D, c_size, _D, _D, _D, c_size]
    L.vs_probe_audio.argtypes = [c_void, c_int, _F, c_size, ctypes.POINTER(c_size)]
    L.vs_butter2_lowpass.argtypes = [ctypes.c_double, ctypes.c_double, _D, _D]
    L.vs_write_wav.argtypes = [ctypes.c_char_p, _F, c_size, c_int]
    L.vs_copy_to_host.argtypes = [c_void, c_void, c_size]
    L.vs_measure_lds_bandwidth.argtypes = [c_int, _D]
    L.vs_dtft_noalloc.argtypes = [_D, c_int, ctypes.c_double, c_int, _D, c_int, _D, _D]
    L.vs_adjoint_timing.argtypes = [c_void, _D, _D]
    L.vs_release_cached_memory.argtypes = []
    for name in EXPORTS:
        if name not in ("vs_last_error", "vs_abi_version"):
            getattr(L, name).restype = c_int
    _lib = L
    return L


def check(rc: int):
    if rc != VS_OK:
        raise VocalSimError(rc, load().vs_last_error().decode())


def _fptr(a):
    return None if a is None else a.ctypes.data_as(_F)


def spec_to_c(spec):
    """Converts a vocalsim.tubes.TubeSpec to the C struct (keeps arrays alive
    through the returned tuple)."""
    keep = (spec.lengths_cm, spec.areas_cm2, spec.aspect, spec.depth_scale, spec.mu3d)
    c = TubeSpecC(spec.solver_type, spec.shape, spec.mouth_dirichlet, spec.srate_mul, spec.pml,
                  spec.source_type, spec.num_steps, spec.audio_dur, spec.num_sections,
                  _fptr(spec.lengths_cm), _fptr(spec.areas_cm2), _fptr(spec.aspect),
                  _fptr(spec.depth_scale), _fptr(spec.mu3d), spec.mu3d_default,
                  spec.gauss_amp, spec.gauss_t0, spec.gauss_width, spec.tm_ps, spec.tm_q,
                  spec.engine_variant, spec.open_space)
    return c, keep
