"""Seeded random tube specifications (tests/random_specs.py) against the
oracle, on the exact kernels through the C ABI.

One batch of 24 random tubes (1-60 sections, 8-19 cm, areas 0.1-12 cm^2,
circle or ellipse, 2D or 2.5D, PML on or off, Dirichlet or open mouth,
sample-rate multipliers 8-20, sine, Gaussian or two-mass sources, both
engine variants, random depth scales and admittances):
  * FP64 and FP32: every set-up field and every coefficient grid of every
    tube equals the oracle's at the same precision, bit for bit;
  * FP64: after 800 steps the probe, the source trace and the final state of
    every tube equal the oracle's (the bar is 1e-10 relative L2; the exact
    kernel follows the oracle's operation order, so max |diff| is 0);
  * FP32 on the on-chip kernel, the tubes it holds in one batch: the probe
    within 1e-4 relative L2 of the FP64 oracle's after 800 steps.
"""
import numpy as np
import pytest

import oracle as orc
import random_specs
from vocalsim import Solver, _abi

pytestmark = pytest.mark.gpu

SPECS = random_specs.random_specs(24, seed=77, num_steps=800)
SETUP_FIELDS = ["cell_type", "boundary_seg", "depth_x", "depth_y", "depth_p", "z_inv",
                "upper_wall", "lower_wall", "segment", "source"]


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / (nb if nb > 0 else 1.0))


@pytest.mark.parametrize("precision", [64, 32])
def test_random_specs_setup_equals_oracle(hip_device, precision):
    print(f"[random specs] set-up, precision {precision}, {len(SPECS)} tubes", flush=True)
    with Solver(SPECS, precision=precision) as s:
        for i, spec in enumerate(SPECS):
            ref = orc.OracleTube(spec, precision)
            assert s.dims[i] == {k: ref.dims[k] for k in s.dims[i]}, f"tube {i}: frame dimensions differ"
            for f in SETUP_FIELDS + list(orc.COEFFS):
                g = s.field(f, i).astype(np.float64)
                o = ref.field(f)
                assert g.shape == o.shape and np.array_equal(g, o), f"tube {i}: {f} differs"


def test_random_specs_fp64_steps_equal_oracle(hip_device):
    print(f"[random specs] FP64 steps, {len(SPECS)} tubes x 800 steps", flush=True)
    with Solver(SPECS, precision=64) as s:
        s.run()
        got = [(s.probe(i), s.source_trace(i), s.field("pressure", i), s.field("vx", i), s.field("vy", i))
               for i in range(len(SPECS))]
    worst = 0.0
    for i, spec in enumerate(SPECS):
        ref = orc.OracleTube(spec, 64)
        p_ref, s_ref = ref.run()
        rP, rVx, rVy = ref.state()
        p, src, P, Vx, Vy = got[i]
        assert np.abs(p_ref).max() > 0, f"tube {i}: silent probe"
        for name, a, b in (("probe", p, p_ref), ("source", src, s_ref), ("P", P, rP), ("Vx", Vx, rVx),
                           ("Vy", Vy, rVy)):
            err = rel_l2(a, b)
            worst = max(worst, float(np.max(np.abs(a - b))))
            assert err <= 1e-10, f"tube {i} (source {spec.source_type}, srate {spec.srate_mul}): {name} {err:.3e}"
    print(f"  max |diff| over all tubes and fields: {worst:.3e}", flush=True)


def test_random_specs_fp32_on_chip_near_fp64_oracle(hip_device):
    # the random tubes the on-chip kernel holds (at most 18 000 padded cells
    # at 1024 threads), in one FP32 batch; FP32 arithmetic in a reassociated
    # form, so the bar is a tolerance on the probe against the FP64 oracle
    fits = [sp for sp in SPECS if (lambda d: (d["frameH"] - 2) * (d["frameW"] - 1))(
        orc.OracleTube(sp, 32).dims) <= 18000]
    assert len(fits) >= 12
    print(f"[random specs] FP32 on chip, {len(fits)} tubes x 800 steps", flush=True)
    with Solver(fits, precision=32, threads=1024, kernel=_abi.K_FAST) as s:
        s.run()
        got = [s.probe(i) for i in range(len(fits))]
    for i, spec in enumerate(fits):
        p_ref, _ = orc.OracleTube(spec, 64).run()
        err = rel_l2(got[i], p_ref)
        print(f"  tube {i}: source {spec.source_type}, srate {spec.srate_mul}, 2.5D {spec.solver_type}, "
              f"probe rel L2 {err:.2e}", flush=True)
        assert err <= 1e-4, f"tube {i}: probe rel L2 {err:.3e}"
