"""ctypes loader of the CPU oracle (oracle/liboracle_vs.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke() check and by the
cpu_baseline leg of bench.py, always as the checker or as the timed CPU
baseline, never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VSO_ORACLE_LIB names another build of the same source (the mutation tests
# load copies built with a planted mistake)
LIB_PATH = os.environ.get("VSO_ORACLE_LIB") or os.path.join(HERE, "liboracle_vs.so")

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)


class VsoSpec(ctypes.Structure):
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


DIMS = ["frameH", "frameW", "domainH", "domainW", "listenerX", "listenerY",
        "excitationX", "excitationY", "num_steps", "tube_start_x", "tube_end_x",
        "tube_len_cells", "midY", "num_sections", "mic_cells"]
SCALARS = ["dt", "ds", "rho", "c", "kappa", "K1", "K2", "ds_f32", "z_default"]
FIELDS = {"cell_type": 0, "boundary_seg": 1, "depth_p": 2, "depth_x": 3, "depth_y": 4,
          "z_inv": 5, "source": 6, "upper_wall": 7, "lower_wall": 8, "segment": 9,
          "diam_cells": 10, "minor_axis": 11, "z_section": 12}
COEFFS = ["N_out", "N_in", "minVxBeta", "minVyBeta", "maxVxSigmaPrimedt",
          "maxVySigmaPrimedt", "betaVxSqr_dt_inv_rho_inv_ds", "betaVySqr_dt_inv_rho_inv_ds",
          "pressureSigmaPrimedt", "areWeNotExcitationVx", "areWeNotExcitationVy",
          "excitationWeightVx", "excitationWeightVy", "xor_val1", "xor_val2", "xor_val3", "xor_val4"]

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.vso_create.restype = ctypes.c_void_p
        L.vso_create.argtypes = [ctypes.POINTER(VsoSpec), ctypes.c_int]
        L.vso_destroy.argtypes = [ctypes.c_void_p]
        L.vso_last_error.restype = ctypes.c_char_p
        L.vso_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.vso_scalars.argtypes = [ctypes.c_void_p, _D]
        L.vso_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_long]
        L.vso_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, _D]
        L.vso_reset.argtypes = [ctypes.c_void_p]
        L.vso_get_state.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.vso_set_state.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.vso_run_structured.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D]
        L.vso_run_threads.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _D]
        L.vso_float_num_steps.restype = ctypes.c_int
        L.vso_float_num_steps.argtypes = [ctypes.c_float, ctypes.c_int]
        L.vso_twomass_run.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      _D, ctypes.c_int, _D]
        L.vso_coeff_triple.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]
        L.vso_plane_contour.argtypes = [ctypes.c_float, _F, _F, ctypes.c_int, _D]
        L.vso_polygon_area.restype = ctypes.c_float
        L.vso_polygon_area.argtypes = [_F, _F, ctypes.c_int]
        L.vso_vertical_chord.argtypes = [_F, _F, ctypes.c_int, ctypes.c_float, _F, _F]
        L.vso_segment_cross.argtypes = [_F, _F, _F]
        _lib = L
    return _lib


def _fptr(a):
    return None if a is None else a.ctypes.data_as(_F)


def _dptr(a):
    return None if a is None else a.ctypes.data_as(_D)


class OracleError(RuntimeError):
    pass


class OracleTube:
    """One tube on the CPU oracle.  ``spec`` is any object with the attributes
    of vocalsim.tubes.TubeSpec."""

    def __init__(self, spec, precision: int = 64):
        self._keep = [spec.lengths_cm, spec.areas_cm2, spec.aspect, spec.depth_scale, spec.mu3d]
        s = VsoSpec(
            spec.solver_type, spec.shape, spec.mouth_dirichlet, spec.srate_mul, spec.pml,
            spec.source_type, spec.num_steps, spec.audio_dur, spec.num_sections,
            _fptr(spec.lengths_cm), _fptr(spec.areas_cm2), _fptr(spec.aspect),
            _fptr(spec.depth_scale), _fptr(spec.mu3d), spec.mu3d_default,
            spec.gauss_amp, spec.gauss_t0, spec.gauss_width, spec.tm_ps, spec.tm_q,
            spec.engine_variant, spec.open_space)
        self.precision = precision
        self.h = lib().vso_create(ctypes.byref(s), precision)
        if not self.h:
            raise OracleError(lib().vso_last_error().decode())
        d = (ctypes.c_int * len(DIMS))()
        lib().vso_dims(self.h, d)
        self.dims = dict(zip(DIMS, list(d)))
        sc = (ctypes.c_double * len(SCALARS))()
        lib().vso_scalars(self.h, sc)
        self.scalars = dict(zip(SCALARS, list(sc)))
        self.H, self.W = self.dims["frameH"], self.dims["frameW"]
        self.num_steps = self.dims["num_steps"]

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().vso_destroy(h)
            self.h = None

    def field(self, name: str) -> np.ndarray:
        H, W, L = self.H, self.W, self.dims["tube_len_cells"]
        ns = self.dims["num_sections"]
        if name in COEFFS:
            fid, shape = 100 + COEFFS.index(name), (H - 2, W - 2)
        else:
            fid = FIELDS[name]
            shape = {"source": (self.num_steps,), "upper_wall": (L,), "lower_wall": (L,),
                     "segment": (L,), "diam_cells": (ns,), "minor_axis": (ns,),
                     "z_section": (ns,)}.get(name, (H, W))
        out = np.zeros(int(np.prod(shape)), dtype=np.float64)
        if lib().vso_get_field(self.h, fid, _dptr(out), out.size) != 0:
            raise OracleError(f"field {name} unavailable")
        return out.reshape(shape)

    def run(self, n_steps=None, first_step: int = 0):
        n = self.num_steps if n_steps is None else int(n_steps)
        probe = np.zeros(first_step + n, dtype=np.float64)
        src = np.zeros(first_step + n, dtype=np.float64)
        lib().vso_run(self.h, first_step, n, _dptr(probe), _dptr(src))
        return probe[first_step:], src[first_step:]

    def reset(self):
        lib().vso_reset(self.h)

    def state(self):
        P = np.zeros(self.H * self.W)
        Vx = np.zeros_like(P)
        Vy = np.zeros_like(P)
        lib().vso_get_state(self.h, _dptr(P), _dptr(Vx), _dptr(Vy))
        return P.reshape(self.H, self.W), Vx.reshape(self.H, self.W), Vy.reshape(self.H, self.W)

    def set_state(self, P, Vx, Vy):
        P = np.ascontiguousarray(P, dtype=np.float64).ravel()
        Vx = np.ascontiguousarray(Vx, dtype=np.float64).ravel()
        Vy = np.ascontiguousarray(Vy, dtype=np.float64).ravel()
        lib().vso_set_state(self.h, _dptr(P), _dptr(Vx), _dptr(Vy))

    def run_threads(self, n_steps: int, threads: int, first_step: int = 0):
        """run() with each phase's rows over `threads` OpenMP threads (same
        results); returns (probe, seconds of the step loop)."""
        probe = np.zeros(first_step + n_steps)
        secs = ctypes.c_double(0.0)
        if lib().vso_run_threads(self.h, first_step, n_steps, _dptr(probe), threads, ctypes.byref(secs)) != 0:
            raise OracleError("vso_run_threads: bad arguments")
        return probe[first_step:], secs.value

    def run_structured(self, n_steps: int):
        probe = np.zeros(n_steps)
        secs = ctypes.c_double(0.0)
        lib().vso_run_structured(self.h, n_steps, _dptr(probe), ctypes.byref(secs))
        return probe, secs.value


def float_num_steps(audio_dur: float, srate_mul: int = 15) -> int:
    return int(lib().vso_float_num_steps(audio_dur, srate_mul))


def twomass_run(p1, precision: int = 32, q: float = 1.0, ps: float = 686.46, srate_mul: int = 15,
                abs_double: bool = False) -> np.ndarray:
    """Volume velocity of the two-mass recurrence fed the supraglottal
    pressures p1 (one per step)."""
    p1 = np.ascontiguousarray(p1, dtype=np.float64)
    ug = np.zeros(len(p1))
    if lib().vso_twomass_run(precision, q, ps, srate_mul, int(abs_double), _dptr(p1), len(p1), _dptr(ug)) != 0:
        raise OracleError("vso_twomass_run: bad arguments")
    return ug


def coeff_triple(cur: int, right: int, top: int, precision: int = 64, srate_mul: int = 15) -> np.ndarray:
    out = np.zeros(17)
    if lib().vso_coeff_triple(precision, srate_mul, cur, right, top, _dptr(out)) != 0:
        raise OracleError("bad cell type")
    return out


def plane_contour(area_cm2: float):
    """Aligned single-plane-symmetric contour of one section: (y, z, info) with
    info = dict(minY, minZ, maxZ, diameter, area, steps)."""
    info = np.zeros(6)
    n = lib().vso_plane_contour(area_cm2, None, None, 0, _dptr(info))
    if n < 0:
        raise OracleError(lib().vso_last_error().decode())
    y = np.zeros(n, dtype=np.float32)
    z = np.zeros(n, dtype=np.float32)
    lib().vso_plane_contour(area_cm2, _fptr(y), _fptr(z), n, _dptr(info))
    keys = ["minY", "minZ", "maxZ", "diameter", "area", "steps"]
    return y, z, dict(zip(keys, info.tolist()))


def polygon_area(y, z) -> float:
    y = np.ascontiguousarray(y, dtype=np.float32)
    z = np.ascontiguousarray(z, dtype=np.float32)
    return float(lib().vso_polygon_area(_fptr(y), _fptr(z), y.size))


def vertical_chord(y, z, y0: float):
    """(lowest, highest) crossing of the vertical line y = y0 with the
    polyline, or None."""
    y = np.ascontiguousarray(y, dtype=np.float32)
    z = np.ascontiguousarray(z, dtype=np.float32)
    lo = np.zeros(1, dtype=np.float32)
    hi = np.zeros(1, dtype=np.float32)
    if not lib().vso_vertical_chord(_fptr(y), _fptr(z), y.size, y0, _fptr(lo), _fptr(hi)):
        return None
    return float(lo[0]), float(hi[0])


def segment_cross(p, q):
    p = np.ascontiguousarray(p, dtype=np.float32)
    q = np.ascontiguousarray(q, dtype=np.float32)
    out = np.zeros(2, dtype=np.float32)
    if not lib().vso_segment_cross(_fptr(p), _fptr(q), _fptr(out)):
        return None
    return float(out[0]), float(out[1])
