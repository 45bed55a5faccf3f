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
    within 1e-4 relative L2 of the FP64 oracle's after 800 steps;
  * gradients of six random tubes with table sources in one batch of mixed
    sample rates: the FP32 on-chip reverse pass within 2e-3 of the FP64
    adjoint, and the FP64 adjoint of one input per tube within 1e-4 of
    central finite differences; inputs the depth does not read continuously
    (the area of circles and of 2D tubes, the depth scale of 2D tubes) have
    zero gradients.
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


def test_random_specs_gradients(hip_device):
    # random tubes with table sources (the reverse pass needs one), 600 steps,
    # in one batch of mixed sample rates: the FP32 on-chip reverse pass
    # against the FP64 adjoint (rel L2 <= 2e-3 per kind and tube), and the
    # FP64 adjoint of one area per tube (a depth scale or an admittance
    # where the area enters only integer geometry) against
    # central finite differences of the FP64 forward (h = 1e-4 relative,
    # integer geometry unchanged)
    pool = random_specs.random_specs(64, seed=91, num_steps=600)
    chosen = [sp for sp in pool if sp.source_type != 2 and (lambda d: (d["frameH"] - 2) * (d["frameW"] - 1))(
        orc.OracleTube(sp, 32).dims) <= 18000][:6]
    chosen = [sp.replace(depth_scale=np.ones(sp.num_sections) if sp.depth_scale is None else sp.depth_scale,
                         mu3d=np.full(sp.num_sections, 0.005) if sp.mu3d is None else sp.mu3d) for sp in chosen]
    assert len(chosen) == 6
    print(f"[random specs] gradients, {len(chosen)} tubes, srate {[sp.srate_mul for sp in chosen]}", flush=True)
    rng = np.random.default_rng(5)
    gs = [rng.standard_normal(sp.num_steps) * np.hanning(sp.num_steps) for sp in chosen]
    grads = {}
    for precision in (64, 32):
        with Solver(chosen, precision=precision) as s:
            s.adjoint_forward(256)
            grads[precision] = s.adjoint_backward(gs)
    for i, sp in enumerate(chosen):
        for k, kind in enumerate(("area", "scale", "mu")):
            r = rel_l2(grads[32][k][i], grads[64][k][i])
            print(f"  tube {i} ({sp.num_sections} sections, srate {sp.srate_mul}, source {sp.source_type}): "
                  f"{kind} FP32 vs FP64 adjoint rel L2 {r:.2e}", flush=True)
            assert r <= 2e-3, (i, kind, r)
    # the depth reads the area only through an ellipse's minor axis, and 2D
    # tubes read no depth: for circles (radius from the integer air rows) and
    # in 2D the area sets only integer geometry, and in 2D the depth scale
    # nothing, so those gradients are zero and the check takes the next kind
    for i, sp in enumerate(chosen):
        smooth_area = sp.solver_type == 1 and sp.shape == 2
        for k, zero in ((0, not smooth_area), (1, sp.solver_type == 0)):
            if zero:
                assert not np.any(grads[64][k][i]) and not np.any(grads[32][k][i]), f"tube {i}: kind {k} not zero"
        kind, field = (("area", "areas_cm2") if smooth_area else ("scale", "depth_scale") if sp.solver_type == 1
                       else ("mu", "mu3d"))
        ga = grads[64][{"area": 0, "scale": 1, "mu": 2}[kind]][i]
        big = np.abs(ga).max()
        assert big > 0, (i, kind)
        with Solver([sp], precision=64) as s:
            cls0 = s.field("cell_class")
        for sec in np.argsort(-np.abs(ga))[:3]:
            arr_p = np.array(getattr(sp, field), np.float32)
            arr_m = arr_p.copy()
            arr_p[sec] = np.float32(float(arr_p[sec]) * (1 + 1e-4))
            arr_m[sec] = np.float32(float(arr_m[sec]) * (1 - 1e-4))
            J = {}
            for name, arr in (("p", arr_p), ("m", arr_m)):
                with Solver([sp.replace(**{field: arr})], precision=64) as s:
                    s.run()
                    J[name] = float(np.dot(gs[i], s.probe()))
                    moved = not np.array_equal(s.field("cell_class"), cls0)
                if moved:
                    break
            if moved:
                continue   # this section's perturbation moves a wall: try the next one
            fd = (J["p"] - J["m"]) / (float(arr_p[sec]) - float(arr_m[sec]))
            print(f"  tube {i} ({'2.5D' if sp.solver_type else '2D'}, shape {sp.shape}): {kind}[{sec}] adjoint {ga[sec]:.8e} "
                  f"fd {fd:.8e}", flush=True)
            assert abs(ga[sec] - fd) <= 1e-4 * max(abs(fd), 1e-3 * big), (i, sec, ga[sec], fd)
            break
        else:
            raise AssertionError(f"tube {i}: every tried perturbation moved a wall")
