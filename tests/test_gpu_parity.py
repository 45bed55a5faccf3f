"""GPU parity of the HIP library against the CPU oracle, through the C ABI.

Bar: integer maps bit-exact; FP64 depths, coefficients and probe waveforms
within 1e-10 relative L2 (the FP64 kernels follow the oracle's operation
order, so they are expected to match bit for bit and the max |diff| is
printed); FP32 against the FP32 oracle within 1e-6 relative L2 for table
sources (two-mass runs in double on the device, so 1e-3 there).
"""
import numpy as np
import pytest

import oracle as orc
from conftest import AREA_FILES
from vocalsim import Solver, tubes

pytestmark = pytest.mark.gpu

REL_FP64 = 1e-10


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def _cases():
    u = tubes.uniform_tube(num_steps=600)
    cases = {
        "uniform_pml": u,
        "uniform_nopml": u.replace(pml=0),
        "uniform_open_mouth": u.replace(mouth_dirichlet=0),
        "uniform_eigen_variant": u.replace(engine_variant=1),
        "uniform_2d": u.replace(solver_type=0),
        "uniform_default_gauss": u.replace(gauss_amp=0.0, gauss_t0=0.0, gauss_width=0.0),
        "vowel0_twomass": tubes.vowel_tube(0, num_steps=600),
        "vowel3_gauss": tubes.vowel_tube(3, num_steps=600, source=tubes.SOURCE_GAUSSIAN),
        "batch_tube7": tubes.batch_tubes(8, seed=1, num_steps=600)[7],
        "batch_tube2": tubes.batch_tubes(8, seed=1, num_steps=600)[2],
        "vowel_a_file_ellipse": tubes.read_area_file(AREA_FILES + "/vocaltract_a.txt").replace(
            shape=tubes.SHAPE_ELLIPTICAL, num_steps=600),
        "cylinder_file_circle": tubes.read_area_file(AREA_FILES + "/cylindrical-tube.txt").replace(
            num_steps=600),
        "vowel5_plane_symmetric": tubes.vowel_tube(5, num_steps=600, source=tubes.SOURCE_GAUSSIAN,
                                                   n_sections=12).replace(shape=tubes.SHAPE_PLANE_SYMMETRIC,
                                                                          aspect=None),
        "open_space_pml": tubes.open_space(num_steps=500),
        "open_space_2d_nopml_twomass": tubes.open_space(num_steps=500, pml=0, solver_type=0,
                                                        source_type=tubes.SOURCE_TWO_MASS),
        "vowel_u_file_plane_symmetric": tubes.read_area_file(AREA_FILES + "/vocaltract_u.txt").replace(
            shape=tubes.SHAPE_PLANE_SYMMETRIC, num_steps=400),
    }
    return cases


CASES = _cases()

SETUP_FIELDS = ["cell_type", "boundary_seg", "depth_x", "depth_y", "depth_p", "z_inv",
                "upper_wall", "lower_wall", "segment", "source"]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", list(CASES))
def test_setup_parity(hip_device, name, precision):
    spec = CASES[name]
    print(f"[setup parity] case={name} precision={precision}", flush=True)
    ref = orc.OracleTube(spec, precision)
    with Solver(spec, precision=precision) as s:
        assert s.dims[0] == {k: ref.dims[k] for k in s.dims[0]}, "frame dimensions differ"
        for f in SETUP_FIELDS:
            g = s.field(f).astype(np.float64)
            o = ref.field(f)
            assert g.shape == o.shape, f
            assert np.array_equal(g, o), f"{f}: max |diff| {np.max(np.abs(g - o))}"
        for k in orc.COEFFS:
            g = s.field(k).astype(np.float64)
            o = ref.field(k)
            assert np.array_equal(g, o), f"coefficient {k} differs"


@pytest.mark.parametrize("name", list(CASES))
def test_step_parity_fp64(hip_device, name):
    spec = CASES[name]
    print(f"[step parity fp64] case={name}", flush=True)
    ref = orc.OracleTube(spec, 64)
    p_ref, s_ref = ref.run()
    with Solver(spec, precision=64) as s:
        s.run()
        p = s.probe()
        src = s.source_trace()
        P, Vx, Vy = s.field("pressure"), s.field("vx"), s.field("vy")
    rP, rVx, rVy = ref.state()
    err = rel_l2(p, p_ref)
    print(f"  probe rel L2 {err:.3e}  max|diff| {np.max(np.abs(p - p_ref)):.3e}  "
          f"state max|diff| {max(np.max(np.abs(P - rP)), np.max(np.abs(Vx - rVx)), np.max(np.abs(Vy - rVy))):.3e}")
    assert np.abs(p_ref).max() > 0
    assert err <= REL_FP64
    assert rel_l2(src, s_ref) <= REL_FP64
    assert rel_l2(P, rP) <= REL_FP64 and rel_l2(Vx, rVx) <= REL_FP64 and rel_l2(Vy, rVy) <= REL_FP64


@pytest.mark.parametrize("name", ["uniform_pml", "vowel3_gauss", "batch_tube7", "vowel0_twomass"])
def test_step_fp32_vs_oracle32(hip_device, name):
    spec = CASES[name]
    print(f"[step fp32] case={name}", flush=True)
    p_ref, _ = orc.OracleTube(spec, 32).run()
    with Solver(spec, precision=32) as s:
        s.run()
        p = s.probe()
    err = rel_l2(p, p_ref)
    tol = 1e-3 if spec.source_type == tubes.SOURCE_TWO_MASS else 1e-6
    print(f"  probe rel L2 vs FP32 oracle {err:.3e}")
    assert err <= tol


def test_batch_matches_single_tubes_fp64(hip_device):
    specs = tubes.batch_tubes(6, seed=3, num_steps=500)
    specs[1] = tubes.uniform_tube(num_steps=450)
    print("[batch fp64] 6 ragged tubes, 500 steps", flush=True)
    with Solver(specs, precision=64) as s:
        s.run()
        probes = [s.probe(i) for i in range(len(specs))]
    for i, sp in enumerate(specs):
        p_ref, _ = orc.OracleTube(sp, 64).run()
        assert probes[i].shape == p_ref.shape
        assert rel_l2(probes[i], p_ref) <= REL_FP64, f"tube {i}"


def test_run_in_blocks_equals_one_run(hip_device):
    spec = tubes.vowel_tube(1, num_steps=700)
    print("[blocks] uneven blocks vs one run", flush=True)
    with Solver(spec, precision=64, steps_per_launch=37) as s:
        for n in (1, 100, 263, 336):
            s.run(n)
        a = s.probe()
        s.reset()
        s.set_option(0, 700)
        s.run()
        b = s.probe()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("threads", [64, 192, 1024])
def test_thread_counts(hip_device, threads):
    spec = tubes.vowel_tube(2, num_steps=300, source=tubes.SOURCE_GAUSSIAN)
    print(f"[threads] {threads}", flush=True)
    p_ref, _ = orc.OracleTube(spec, 64).run()
    with Solver(spec, precision=64, threads=threads) as s:
        s.run()
        assert rel_l2(s.probe(), p_ref) <= REL_FP64
