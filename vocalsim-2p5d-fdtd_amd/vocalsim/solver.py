"""Batch solver handle over the C ABI (one handle = n >= 1 tubes on the GPU)."""
from __future__ import annotations

import ctypes
from typing import Dict, List, Sequence

import numpy as np

from . import _abi
from ._abi import check, load


def _check_device_f64(t, n: int, what: str):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
            and t.numel() == n):
        raise ValueError(f"{what} must be a contiguous float64 GPU tensor of {n} values")


class Solver:
    """Owns a vs_handle: a batch of tubes with their device state.

    ``precision`` is 64 (the plain FP64 path) or 32 (the batched FP32 path).
    Setup (walls, cell classes, depth maps, z_inv, sources) runs on the GPU in
    the constructor.
    """

    def __init__(self, specs, precision: int = 64, threads: int = 256,
                 steps_per_launch: int = 256, kernel: int = _abi.K_EXACT):
        if not isinstance(specs, (list, tuple)):
            specs = [specs]
        L = load()
        arr = (_abi.TubeSpecC * len(specs))()
        self._keep = []
        for i, s in enumerate(specs):
            c, keep = _abi.spec_to_c(s)
            arr[i] = c
            self._keep.append(keep)
        h = ctypes.c_void_p()
        check(L.vs_create(arr, len(specs), precision, ctypes.byref(h)))
        self.h = h
        self.n = len(specs)
        self.precision = precision
        self.dims: List[Dict[str, int]] = []
        for i in range(self.n):
            d = _abi.Dims()
            check(L.vs_get_dims(self.h, i, ctypes.byref(d)))
            self.dims.append(d.as_dict())
        self.set_option(_abi.OPT_THREADS, threads)
        self.set_option(_abi.OPT_STEPS_PER_LAUNCH, steps_per_launch)
        self.set_option(_abi.OPT_KERNEL, kernel)

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            load().vs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- options -----------------------------------------------------------
    def set_option(self, opt: int, value: int):
        check(load().vs_set_option(self.h, opt, int(value)))

    def get_option(self, opt: int) -> int:
        v = ctypes.c_int()
        check(load().vs_get_option(self.h, opt, ctypes.byref(v)))
        return v.value

    def autotune(self, timed_steps: int = 64) -> dict:
        """Picks the fastest work-group shape / steps per launch for this batch
        (vs_autotune) and resets the state.  Returns the JSON report."""
        import json
        buf = ctypes.create_string_buffer(16384)
        check(load().vs_autotune(self.h, int(timed_steps), buf, len(buf)))
        return json.loads(buf.value.decode())

    # -- stepping ----------------------------------------------------------
    def run(self, n_steps: int = None, sync: bool = True):
        if n_steps is None:
            n_steps = max(d["num_steps"] for d in self.dims)
        check(load().vs_run(self.h, int(n_steps)))
        if sync:
            self.sync()

    def sync(self):
        check(load().vs_sync(self.h))

    def reset(self):
        check(load().vs_reset(self.h))

    @property
    def steps_done(self) -> int:
        v = ctypes.c_int()
        check(load().vs_steps_done(self.h, ctypes.byref(v)))
        return v.value

    def timing(self):
        ms, mean = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int()
        check(load().vs_last_run_timing(self.h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(mean)))
        return ms.value, n.value, mean.value

    @property
    def cells_per_step(self) -> int:
        v = ctypes.c_longlong()
        check(load().vs_cells_per_step(self.h, ctypes.byref(v)))
        return v.value

    # -- reverse-time gradients --------------------------------------------
    def adjoint_forward(self, checkpoint_every: int = 512) -> List[np.ndarray]:
        """Differentiable forward pass from the zero state (vs_adjoint_forward):
        returns each tube's probe trace and keeps checkpoints for
        :meth:`adjoint_backward`.  The ordinary state (run/reset) is untouched."""
        total = sum(d["num_steps"] for d in self.dims)
        out = np.zeros(total, dtype=np.float64)
        check(load().vs_adjoint_forward(self.h, int(checkpoint_every), out.ctypes.data_as(_abi._D), total))
        self._adj_total = total
        return self._split_steps(out)

    def adjoint_backward(self, gprobe: Sequence[np.ndarray]):
        """Gradients of J = sum_tubes sum_n gprobe[t][n] probe[t][n] with respect to
        every section's area_cm2, depth_scale and mu3d (vs_adjoint_backward).
        Returns three lists of per-tube arrays."""
        if len(gprobe) != self.n:
            raise ValueError(f"need {self.n} gradient traces, got {len(gprobe)}")
        parts = []
        for t, g in enumerate(gprobe):
            g = np.asarray(g, dtype=np.float64).ravel()
            if g.size != self.dims[t]["num_steps"]:
                raise ValueError(f"tube {t}: gradient trace needs {self.dims[t]['num_steps']} values")
            parts.append(g)
        g = np.ascontiguousarray(np.concatenate(parts))
        ns = sum(d["num_sections"] for d in self.dims)
        ga, gs, gm = np.zeros(ns), np.zeros(ns), np.zeros(ns)
        check(load().vs_adjoint_backward(self.h, g.ctypes.data_as(_abi._D), g.size,
                                         ga.ctypes.data_as(_abi._D), gs.ctypes.data_as(_abi._D),
                                         gm.ctypes.data_as(_abi._D), ns))
        split = np.cumsum([d["num_sections"] for d in self.dims])[:-1]
        return np.split(ga, split), np.split(gs, split), np.split(gm, split)

    def adjoint_forward_to(self, out, checkpoint_every: int = 512):
        """:meth:`adjoint_forward` writing the tube-major probe traces straight
        into ``out``, a contiguous float64 torch tensor on the GPU with one
        value per tube step (no host copy)."""
        total = sum(d["num_steps"] for d in self.dims)
        _check_device_f64(out, total, "probe output")
        check(load().vs_adjoint_forward(self.h, int(checkpoint_every), ctypes.cast(out.data_ptr(), _abi._D), total))
        self._adj_total = total
        return out

    def adjoint_backward_from(self, gprobe):
        """:meth:`adjoint_backward` reading the tube-major gradient traces from
        ``gprobe``, a contiguous float64 torch tensor on the GPU (read by the
        library on its own stream once the caller's current stream is done)."""
        import torch
        total = sum(d["num_steps"] for d in self.dims)
        _check_device_f64(gprobe, total, "gradient traces")
        torch.cuda.current_stream(gprobe.device).synchronize()
        ns = sum(d["num_sections"] for d in self.dims)
        ga, gs, gm = np.zeros(ns), np.zeros(ns), np.zeros(ns)
        check(load().vs_adjoint_backward(self.h, ctypes.cast(gprobe.data_ptr(), _abi._D), total,
                                         ga.ctypes.data_as(_abi._D), gs.ctypes.data_as(_abi._D),
                                         gm.ctypes.data_as(_abi._D), ns))
        split = np.cumsum([d["num_sections"] for d in self.dims])[:-1]
        return np.split(ga, split), np.split(gs, split), np.split(gm, split)

    def adjoint_timing(self):
        """(device reverse pass ms, host chain ms) of the last adjoint_backward."""
        r, c = ctypes.c_double(), ctypes.c_double()
        check(load().vs_adjoint_timing(self.h, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    def _split_steps(self, flat: np.ndarray) -> List[np.ndarray]:
        split = np.cumsum([d["num_steps"] for d in self.dims])[:-1]
        return np.split(flat, split)

    # -- outputs -----------------------------------------------------------
    def probe(self, tube: int = 0) -> np.ndarray:
        n = self.dims[tube]["num_steps"]
        out = np.zeros(n, dtype=np.float64)
        check(load().vs_get_probe(self.h, tube, out.ctypes.data_as(_abi._D), n))
        return out[: min(n, self.steps_done)]

    def audio(self, tube: int = 0) -> np.ndarray:
        """44.1 kHz float32 audio of the probe (vs_probe_audio): normalised,
        clipped, Butterworth-filtered at 5050 Hz and decimated on the GPU, as
        results/generate_audio.py does.  tube = -1: all tubes concatenated."""
        n = ctypes.c_size_t()
        check(load().vs_probe_audio(self.h, int(tube), None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.float32)
        check(load().vs_probe_audio(self.h, int(tube), out.ctypes.data_as(_abi._F), out.size, ctypes.byref(n)))
        return out

    def write_wav(self, path: str, tube: int = 0):
        """Writes :meth:`audio` as a 44.1 kHz float32 WAV file."""
        write_wav(path, self.audio(tube))

    def source_trace(self, tube: int = 0) -> np.ndarray:
        n = self.dims[tube]["num_steps"]
        out = np.zeros(n, dtype=np.float64)
        check(load().vs_get_source_trace(self.h, tube, out.ctypes.data_as(_abi._D), n))
        return out[: min(n, self.steps_done)]

    def field(self, name: str, tube: int = 0) -> np.ndarray:
        d = self.dims[tube]
        H, W, L = d["frameH"], d["frameW"], d["tube_len_cells"]
        real = np.float64 if self.precision == 64 else np.float32
        table = {
            "cell_type": (_abi.F_CELL_TYPE, (H, W), np.int32),
            "boundary_seg": (_abi.F_BOUNDARY_SEG, (H, W), np.int32),
            "cell_class": (_abi.F_CELL_CLASS, (H, W), np.uint8),
            "depth_p": (_abi.F_DEPTH_P, (H, W), real),
            "depth_x": (_abi.F_DEPTH_X, (H, W), real),
            "depth_y": (_abi.F_DEPTH_Y, (H, W), real),
            "z_inv": (_abi.F_Z_INV, (H, W), real),
            "pressure": (_abi.F_PRESSURE, (H, W), real),
            "vx": (_abi.F_VX, (H, W), real),
            "vy": (_abi.F_VY, (H, W), real),
            "source": (_abi.F_SOURCE, (d["num_steps"],), real),
            "upper_wall": (_abi.F_UPPER_WALL, (L,), np.int32),
            "lower_wall": (_abi.F_LOWER_WALL, (L,), np.int32),
            "segment": (_abi.F_SEGMENT, (L,), np.int32),
        }
        if name in _abi.COEFFS:
            fid, shape, dt = _abi.F_COEFF0 + _abi.COEFFS.index(name), (H - 2, W - 2), real
        else:
            fid, shape, dt = table[name]
        out = np.zeros(shape, dtype=dt)
        check(load().vs_get_field(self.h, tube, fid, out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out


def butter2_lowpass(cutoff_hz: float, fs: float):
    """(b, a) of the second-order Butterworth low-pass (vs_butter2_lowpass)."""
    b = np.zeros(3)
    a = np.zeros(3)
    check(load().vs_butter2_lowpass(float(cutoff_hz), float(fs), b.ctypes.data_as(_abi._D),
                                    a.ctypes.data_as(_abi._D)))
    return b, a


def write_wav(path: str, samples: np.ndarray, rate: int = 44100):
    """Mono float32 WAV (vs_write_wav)."""
    x = np.ascontiguousarray(samples, dtype=np.float32)
    check(load().vs_write_wav(str(path).encode(), x.ctypes.data_as(_abi._F), x.size, int(rate)))
