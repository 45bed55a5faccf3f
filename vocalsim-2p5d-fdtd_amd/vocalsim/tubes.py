"""Tube descriptions and the seeded area-function generators of the workloads.

A :class:`TubeSpec` carries what the reference reads from its compile-time
flags (src/main.cu:68-102) and its area-function file
(src/TubeGeometry.cu:53-107), plus the inputs the workloads add: in-memory
area functions, per-section ellipse aspect, depth scale and wall admittance,
Gaussian pulse parameters, two-mass subglottal pressure and pitch factor, and
an explicit step count.

Workload families (sizes quoted in the docstrings of the generators):

* ``uniform_tube``        17.5 cm circular tube, 2.84 cm^2, Gaussian pulse
* ``vowel_tube(seed)``     44 elliptical sections, smooth seeded log-area walk
* ``batch_tubes(n, seed)`` ragged batch: 13-19 cm, 30-50 sections, per-tube
                           admittance in [0.002, 0.01], half Gaussian / half
                           two-mass sources
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

SRATE_MUL = 15
FS = 44100 * SRATE_MUL

SOURCE_SINE = 0
SOURCE_GAUSSIAN = 1
SOURCE_TWO_MASS = 2

SHAPE_CIRCULAR = 1
SHAPE_ELLIPTICAL = 2
SHAPE_PLANE_SYMMETRIC = 3


@dataclasses.dataclass
class TubeSpec:
    lengths_cm: np.ndarray
    areas_cm2: np.ndarray
    aspect: Optional[np.ndarray] = None
    depth_scale: Optional[np.ndarray] = None
    mu3d: Optional[np.ndarray] = None
    solver_type: int = 1          # 0: 2D, 1: 2.5D
    shape: int = SHAPE_CIRCULAR
    mouth_dirichlet: int = 1
    srate_mul: int = SRATE_MUL
    pml: int = 1
    source_type: int = SOURCE_GAUSSIAN
    num_steps: int = 0            # 0: float-accumulated from audio_dur
    audio_dur: float = 0.02
    mu3d_default: float = 0.0     # 0: reference default
    gauss_amp: float = 0.0        # 0: reference Gaussian
    gauss_t0: float = 0.0
    gauss_width: float = 0.0
    tm_ps: float = 0.0            # 0: 686.46 Pa
    tm_q: float = 0.0             # 0: 1.0
    engine_variant: int = 0       # 0: FDTDEngineCPU, 1: FDTDEngineEigen
    open_space: int = 0           # 1: open-space domain (simulation_type 0), sections ignored

    def __post_init__(self):
        self.lengths_cm = np.ascontiguousarray(self.lengths_cm, dtype=np.float32)
        self.areas_cm2 = np.ascontiguousarray(self.areas_cm2, dtype=np.float32)
        for name in ("aspect", "depth_scale", "mu3d"):
            v = getattr(self, name)
            if v is not None:
                setattr(self, name, np.ascontiguousarray(v, dtype=np.float32))
        if self.lengths_cm.shape != self.areas_cm2.shape:
            raise ValueError("lengths and areas must have the same number of sections")

    @property
    def num_sections(self) -> int:
        return int(self.areas_cm2.size)

    def replace(self, **kw) -> "TubeSpec":
        d = dataclasses.asdict(self)
        d.update(kw)
        return TubeSpec(**d)


def open_space(num_steps: int = 1000, pml: int = 1, solver_type: int = 1,
               source_type: int = SOURCE_GAUSSIAN, **kw) -> TubeSpec:
    """Open-space domain of the reference's simulation_type 0
    (FDTDSolver.cu:479-550): 43 x 341 cells of air, a point source at the
    centre and the listener 5 cells down and right."""
    return TubeSpec(lengths_cm=np.array([1.0]), areas_cm2=np.array([1.0]), open_space=1,
                    num_steps=num_steps, pml=pml, solver_type=solver_type, source_type=source_type,
                    **kw)


def read_area_file(path: str) -> TubeSpec:
    """Reads the reference's area-function format: one ``<segment length cm>
    <area cm^2>`` pair per line, whitespace separated (TubeGeometry.cu:83-88)."""
    lens: List[float] = []
    areas: List[float] = []
    with open(path) as f:
        toks = f.read().split()
    for i in range(0, len(toks) - 1, 2):
        try:
            a, b = float(toks[i]), float(toks[i + 1])
        except ValueError:
            break
        lens.append(a)
        areas.append(b)
    if not lens:
        raise ValueError(f"no sections in {path}")
    return TubeSpec(lengths_cm=np.array(lens), areas_cm2=np.array(areas))


# --------------------------------------------------------------------------
# Seeded generator (SplitMix64), identical wherever it is used.
# --------------------------------------------------------------------------
class SplitMix64:
    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


def smooth_log_areas(rng: SplitMix64, n: int, lo: float = 0.25, hi: float = 8.0) -> np.ndarray:
    """log A(x) = mu + sum_{k=1..4} a_k cos(k pi x + phi_k), x = section centre
    in [0, 1], clipped to [lo, hi] cm^2."""
    mu = math.log(2.0) + rng.uniform(-0.3, 0.3)
    amps = [rng.uniform(-0.9, 0.9) / k for k in range(1, 5)]
    phis = [rng.uniform(0.0, 2.0 * math.pi) for _ in range(4)]
    x = (np.arange(n) + 0.5) / n
    la = mu + sum(a * np.cos((k + 1) * math.pi * x + p) for k, (a, p) in enumerate(zip(amps, phis)))
    return np.clip(np.exp(la), lo, hi).astype(np.float32)


def uniform_tube(num_steps: int = 13230, pml: int = 1, sigma_s: float = 0.2e-3,
                 mu3d: float = 0.005) -> TubeSpec:
    """Straight 17.5 cm circular tube of 2.84 cm^2 (frame 250 x 41 with PML),
    Gaussian velocity pulse of standard deviation ``sigma_s`` (width sqrt(2)
    sigma, centre 4 sigma, amplitude 0.025), wall admittance 0.005."""
    return TubeSpec(lengths_cm=np.array([17.5]), areas_cm2=np.array([2.84]),
                    shape=SHAPE_CIRCULAR, pml=pml, source_type=SOURCE_GAUSSIAN,
                    num_steps=num_steps, mu3d_default=mu3d,
                    gauss_amp=0.025, gauss_t0=4.0 * sigma_s,
                    gauss_width=math.sqrt(2.0) * sigma_s)


# Pitch factor that puts the two-mass oscillation near 120 Hz at 800 Pa on the
# 44-section vowel tube (measured with the FP64 oracle: q = 0.72 gives 122 Hz).
TWO_MASS_Q_120HZ = 0.70


def vowel_tube(seed: int = 0, num_steps: int = 33075, source: int = SOURCE_TWO_MASS,
               n_sections: int = 44, length_cm: float = 17.5,
               sigma_s: float = 0.2e-3) -> TubeSpec:
    """44 elliptical sections over 17.5 cm; log-areas from a seeded smooth
    4-term cosine series clipped to [0.25, 8] cm^2; aspect ratio U[1, 2];
    two-mass source at 800 Pa (or a Gaussian pulse)."""
    rng = SplitMix64(0x5EED0000 + seed)
    areas = smooth_log_areas(rng, n_sections)
    aspect = np.array([rng.uniform(1.0, 2.0) for _ in range(n_sections)], dtype=np.float32)
    spec = TubeSpec(lengths_cm=np.full(n_sections, length_cm / n_sections),
                    areas_cm2=areas, aspect=aspect, shape=SHAPE_ELLIPTICAL,
                    num_steps=num_steps, source_type=source, mu3d_default=0.005)
    if source == SOURCE_TWO_MASS:
        spec.tm_ps = 800.0
        spec.tm_q = TWO_MASS_Q_120HZ
    elif source == SOURCE_GAUSSIAN:
        spec.gauss_amp = 0.025
        spec.gauss_t0 = 4.0 * sigma_s
        spec.gauss_width = math.sqrt(2.0) * sigma_s
    return spec


def batch_tubes(n: int = 256, seed: int = 0, num_steps: int = 19845,
                sigma_s: float = 0.2e-3) -> List[TubeSpec]:
    """Ragged batch: tube length U[13, 19] cm, 30-50 sections, areas as in
    ``vowel_tube``, per-section depth scale U[0.8, 1.2], per-tube wall
    admittance U[0.002, 0.01].  The first half uses Gaussian pulses, the second
    half the two-mass source at 800 Pa."""
    out = []
    rng = SplitMix64(0xBA7C0000 + seed)
    for b in range(n):
        length = rng.uniform(13.0, 19.0)
        nsec = 30 + int(rng.uniform(0, 21)) % 21
        areas = smooth_log_areas(rng, nsec)
        aspect = np.array([rng.uniform(1.0, 2.0) for _ in range(nsec)], dtype=np.float32)
        dscale = np.array([rng.uniform(0.8, 1.2) for _ in range(nsec)], dtype=np.float32)
        mu = rng.uniform(0.002, 0.01)
        spec = TubeSpec(lengths_cm=np.full(nsec, length / nsec), areas_cm2=areas,
                        aspect=aspect, depth_scale=dscale,
                        mu3d=np.full(nsec, mu, dtype=np.float32), shape=SHAPE_ELLIPTICAL,
                        num_steps=num_steps)
        if b < n // 2:
            spec.source_type = SOURCE_GAUSSIAN
            spec.gauss_amp = 0.025
            spec.gauss_t0 = 4.0 * sigma_s
            spec.gauss_width = math.sqrt(2.0) * sigma_s
        else:
            spec.source_type = SOURCE_TWO_MASS
            spec.tm_ps = 800.0
            spec.tm_q = TWO_MASS_Q_120HZ
        out.append(spec)
    return out
