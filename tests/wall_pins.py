"""Measurements of the oracle's wall-admittance term that have a closed form
or a sign fixed by physics.  Shared by tests/test_oracle_walls.py, which runs
them on the oracle and, in a child process, on copies of the oracle built
with a planted mistake (VSO_ORACLE_LIB names the copy).

The wall term is FDTDEngineCPU.cpp:391-401: on a face whose current cell is a
wall carrying a boundary segment, the face velocity is
  Vb = (xor2 P N_out - xor1 P_right N_in) * notExc * z_inv(current cell)
with z_inv = 1 / (rho c (1 + sqrt(1 - alpha)) / (1 - sqrt(1 - alpha))),
alpha = 1 / (0.5 + 0.25 (mu2D + 1/mu2D)), mu2D = mu3D * pi / 1.84
(FDTDSolver.cu:613-617).  Since alpha = 4 mu2D / (1 + mu2D)^2, rho c z_inv
equals mu2D, and a plane wave meeting such a wall at normal incidence
reflects with R = (1 - rho c z_inv) / (1 + rho c z_inv) = (1 - mu2D) / (1 + mu2D).
"""
from __future__ import annotations

import math

import numpy as np

import oracle as orc
from vocalsim import tubes

AIR = 1


def mu2d(mu3d: float) -> float:
    return mu3d * math.pi / 1.84


def normal_incidence_decay(mu3d: float, length_cm: float = 60.0, steps: int = 560):
    """Transverse standing wave between the two walls of a long straight 2D
    tube (PML off, source off): returns (measured, expected) decay rates of
    its amplitude per step.

    The first transverse mode P ~ cos(pi (y - y_top) / D) of a strip of
    D air rows bounces between the walls once per round trip 2 D ds / c.
    Only the lower wall absorbs: the upper wall's face belongs to the air cell
    below it, where z_inv is zero (z_inv is sampled at the current cell,
    FDTDEngineCPU.cpp:397-398).  The amplitude therefore falls by R per round
    trip: rate = ln R / (2 D ds / (c dt)).  The amplitude is the projection of
    the middle columns' pressure on the mode shape; the tube is long enough
    that the disturbances from its two ends do not reach the middle within
    ``steps``."""
    spec = tubes.TubeSpec(lengths_cm=np.array([length_cm]), areas_cm2=np.array([2.84]),
                          shape=tubes.SHAPE_CIRCULAR, pml=0, solver_type=0,
                          source_type=tubes.SOURCE_GAUSSIAN, num_steps=steps,
                          gauss_amp=1e-300, mu3d_default=mu3d)
    t = orc.OracleTube(spec, 64)
    H, W = t.H, t.W
    d = t.dims
    up, lo = t.field("upper_wall").astype(int), t.field("lower_wall").astype(int)
    mid = (d["tube_start_x"] + d["tube_end_x"]) // 2
    j = mid - d["tube_start_x"]
    ru, rl = up[j], lo[j]
    D = rl - ru - 1
    rows = np.arange(H)
    shape = np.where((rows > ru) & (rows < rl), np.cos(math.pi * (rows - ru - 0.5) / D), 0.0)
    ct = t.field("cell_type").astype(int)
    P0 = np.repeat(shape[:, None], W, axis=1)
    P0[ct != AIR] = 0
    t.set_state(P0, np.zeros_like(P0), np.zeros_like(P0))
    amp = np.empty(steps)
    for s in range(steps):
        t.run(1, first_step=s)
        P, _, _ = t.state()
        amp[s] = np.dot(P[:, mid - 5:mid + 6].mean(axis=1), shape) / np.dot(shape, shape)
    # envelope: local maxima of |amp| (one per half period)
    a = np.abs(amp)
    pk = np.array([i for i in range(1, steps - 1) if a[i] >= a[i - 1] and a[i] > a[i + 1]])
    measured = float(np.polyfit(pk, np.log(a[pk]), 1)[0])
    sc = t.scalars
    m2 = mu2d(mu3d)
    R = (1 - m2) / (1 + m2)
    round_trip = 2 * D * sc["ds"] / sc["c"] / sc["dt"]
    return measured, math.log(R) / round_trip


def wall_energy_history(mu3d: float, checkpoints: int = 8, every: int = 300, seed: int = 0):
    """Depth-weighted acoustic energy of a 2.5D vowel tube with PML off and the
    source off, from a pressure blob inside the tube, at ``checkpoints``
    times ``every`` steps.  The Dirichlet mouth, the rigid excitation face and
    the walls' z_inv are the only boundaries; with mu3d > 0 the walls are
    passive, so the energy must fall."""
    spec = tubes.vowel_tube(seed, num_steps=checkpoints * every, source=tubes.SOURCE_GAUSSIAN).replace(
        pml=0, gauss_amp=1e-300, mu3d_default=mu3d)
    t = orc.OracleTube(spec, 64)
    H, W = t.H, t.W
    rr, cc = np.mgrid[0:H, 0:W]
    m = t.dims["midY"]
    P0 = np.exp(-(((rr - m) / 3.0) ** 2 + ((cc - 120) / 3.0) ** 2))
    P0[t.field("cell_type").astype(int) != AIR] = 0
    t.set_state(P0, np.zeros_like(P0), np.zeros_like(P0))
    X, Y, Dp = t.field("depth_x"), t.field("depth_y"), t.field("depth_p")
    sc = t.scalars

    def energy():
        P, Vx, Vy = t.state()
        return 0.5 * np.sum(P ** 2 * Dp / sc["kappa"]) + 0.5 * sc["rho"] * np.sum(Vx ** 2 * X + Vy ** 2 * Y)

    es = [energy()]
    for i in range(checkpoints):
        t.run(every, first_step=i * every)
        es.append(energy())
    return np.array(es)


def report() -> dict:
    """Every measurement the wall pins use (JSON-serialisable)."""
    out = {}
    for mu in (0.005, 0.025):
        m, e = normal_incidence_decay(mu)
        out[f"normal_{mu}"] = [m, e]
    for mu in (0.005, 0.025):
        out[f"energy_{mu}"] = wall_energy_history(mu).tolist()
    return out


def check(rep: dict) -> list:
    """The pins: a list of failures (empty when every pin holds)."""
    fails = []
    for mu in (0.005, 0.025):
        m, e = rep[f"normal_{mu}"]
        if not (0.97 <= m / e <= 1.03):
            fails.append(f"normal incidence mu={mu}: measured {m:.4e}/step, closed form {e:.4e}/step")
    e1 = np.array(rep["energy_0.005"])
    e2 = np.array(rep["energy_0.025"])
    for mu, e in ((0.005, e1), (0.025, e2)):
        if not np.all(np.diff(e) < 0):
            fails.append(f"energy not strictly falling at mu={mu}: {np.round(e / e[0], 4).tolist()}")
    if not np.all(e2[1:] < e1[1:]):
        fails.append("energy does not fall faster with the larger admittance")
    return fails
