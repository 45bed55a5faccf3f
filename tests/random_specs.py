"""Seeded random tube specifications over the input ranges the library
accepts, for sweeps that compare its set-up and step with the oracle.

Each spec draws: 1-60 sections of uneven lengths over 8-19 cm, areas
log-uniform in [0.1, 12] cm^2, a circular or elliptical cross-section
(aspect 1-3), 2D or 2.5D, PML on or off, a Dirichlet or open mouth (open
only with PML), the sample-rate multiplier (8, 10, 15 or 20), the source
(sine, Gaussian of random width, or two-mass of random pressure and pitch
factor), the engine variant, and optional per-section depth scale and wall
admittance.
"""
from __future__ import annotations

import numpy as np

from vocalsim import tubes


def random_spec(rng: np.random.Generator, num_steps: int) -> tubes.TubeSpec:
    nsec = int(rng.integers(1, 61))
    w = rng.uniform(0.5, 1.5, nsec)
    lengths = rng.uniform(8.0, 19.0) * w / w.sum()
    areas = np.exp(rng.uniform(np.log(0.1), np.log(12.0), nsec))
    shape = int(rng.choice([tubes.SHAPE_CIRCULAR, tubes.SHAPE_ELLIPTICAL]))
    pml = int(rng.random() < 0.75)
    d = dict(lengths_cm=lengths, areas_cm2=areas, shape=shape,
             solver_type=int(rng.random() < 0.8), pml=pml,
             mouth_dirichlet=1 if not pml else int(rng.random() < 0.7),
             srate_mul=int(rng.choice([8, 10, 15, 15, 15, 20])),
             source_type=int(rng.choice([tubes.SOURCE_SINE, tubes.SOURCE_GAUSSIAN, tubes.SOURCE_GAUSSIAN,
                                         tubes.SOURCE_TWO_MASS])),
             num_steps=num_steps, engine_variant=int(rng.random() < 0.2))
    if shape == tubes.SHAPE_ELLIPTICAL:
        d["aspect"] = rng.uniform(1.0, 3.0, nsec)
    if rng.random() < 0.6:
        d["depth_scale"] = rng.uniform(0.6, 1.4, nsec)
    if rng.random() < 0.6:
        d["mu3d"] = rng.uniform(0.001, 0.02, nsec)
    if d["source_type"] == tubes.SOURCE_GAUSSIAN:
        sigma = rng.uniform(0.1e-3, 0.4e-3)
        d.update(gauss_amp=float(rng.uniform(0.01, 0.05)), gauss_t0=4.0 * sigma, gauss_width=np.sqrt(2.0) * sigma)
    elif d["source_type"] == tubes.SOURCE_TWO_MASS:
        d.update(tm_ps=float(rng.uniform(500.0, 1000.0)), tm_q=float(rng.uniform(0.6, 1.4)))
    return tubes.TubeSpec(**d)


def random_specs(n: int, seed: int, num_steps: int = 40):
    rng = np.random.default_rng(seed)
    return [random_spec(rng, num_steps) for _ in range(n)]
