"""GPU checks of the FP32 flux-form kernel (VS_K_FAST) through the C ABI.

The flux form reassociates the reference update (coefficients precomputed in
double, rounded once to float), so it is compared within FP32 tolerances:
against the FP32 oracle (relative L2 of the probe <= 1e-4) and against the
FP64 oracle (relative L2 <= 1e-4 over the runs below).
"""
import numpy as np
import pytest

import oracle as orc
import random_specs
from vocalsim import Solver, tubes
from vocalsim import _abi

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


CASES = {
    "uniform": tubes.uniform_tube(num_steps=1500),
    "vowel3_gauss": tubes.vowel_tube(3, num_steps=1500, source=tubes.SOURCE_GAUSSIAN),
    "vowel0_twomass": tubes.vowel_tube(0, num_steps=1500),
    "batch_tube5": tubes.batch_tubes(8, seed=2, num_steps=1500)[5],
    "open_mouth": tubes.uniform_tube(num_steps=1500).replace(mouth_dirichlet=0),
    "no_pml": tubes.vowel_tube(1, num_steps=1500, source=tubes.SOURCE_GAUSSIAN).replace(pml=0),
}


@pytest.mark.parametrize("threads", [1024, 768])
@pytest.mark.parametrize("name", list(CASES))
def test_fast_vs_oracles(hip_device, name, threads):
    spec = CASES[name]
    print(f"[fast] case={name} threads={threads}", flush=True)
    p64, s64 = orc.OracleTube(spec, 64).run()
    p32, _ = orc.OracleTube(spec, 32).run()
    with Solver(spec, precision=32, threads=threads) as s:
        d = s.dims[0]
        if (d["frameH"] - 2) * (d["frameW"] - 2) > 24 * (threads - 64):
            pytest.skip("tube larger than the on-chip kernel holds at this work-group size")
        s.set_option(_abi.OPT_KERNEL, _abi.K_FAST)
        s.run()
        p = s.probe()
        src = s.source_trace()
    e32, e64 = rel_l2(p, p32), rel_l2(p, p64)
    print(f"  rel L2 vs FP32 oracle {e32:.2e}, vs FP64 oracle {e64:.2e}, FP32 oracle vs FP64 {rel_l2(p32, p64):.2e}")
    assert np.all(np.isfinite(p))
    # the FP32 oracle runs the two-mass recurrence in float; the device runs it in
    # double, so the FP32 oracle can be the farther one from the FP64 oracle
    assert e64 <= 1e-4 and e32 <= max(1e-4, 2 * rel_l2(p32, p64))
    assert rel_l2(src, s64) <= 1e-4


def test_fast_batch_matches_single(hip_device):
    specs = tubes.batch_tubes(16, seed=4, num_steps=800)
    print("[fast batch] 16 tubes", flush=True)
    with Solver(specs, precision=32, threads=1024, kernel=_abi.K_FAST) as s:
        s.run()
        probes = [s.probe(i) for i in range(len(specs))]
    for i in (0, 7, 8, 15):
        with Solver(specs[i], precision=32, threads=1024, kernel=_abi.K_FAST) as s1:
            s1.run()
            assert np.array_equal(s1.probe(), probes[i]), i
        p64, _ = orc.OracleTube(specs[i], 64).run()
        assert rel_l2(probes[i], p64) <= 1e-4


def test_fast_blocks_equal_one_run(hip_device):
    spec = tubes.vowel_tube(5, num_steps=900)
    with Solver(spec, precision=32, threads=1024, kernel=_abi.K_FAST, steps_per_launch=900) as s:
        s.run()
        a = s.probe()
    with Solver(spec, precision=32, threads=1024, kernel=_abi.K_FAST, steps_per_launch=64) as s:
        for n in (100, 300, 500):
            s.run(n)
        b = s.probe()
    assert np.array_equal(a, b)


def test_fast_state_export_is_velocity(hip_device):
    spec = tubes.vowel_tube(2, num_steps=400, source=tubes.SOURCE_GAUSSIAN)
    o = orc.OracleTube(spec, 64)
    o.run()
    P, Vx, Vy = o.state()
    with Solver(spec, precision=32, threads=1024, kernel=_abi.K_FAST) as s:
        s.run()
        gP, gVx, gVy = s.field("pressure"), s.field("vx"), s.field("vy")
    assert rel_l2(gP, P) < 1e-4 and rel_l2(gVx, Vx) < 1e-4 and rel_l2(gVy, Vy) < 1e-4


def test_fast_rejects_oversized_tube(hip_device):
    spec = tubes.vowel_tube(0, num_steps=10)
    with Solver(spec, precision=32, threads=256, kernel=_abi.K_FAST) as s:
        with pytest.raises(_abi.VocalSimError):
            s.run()


def test_fast_rejects_fp64(hip_device):
    spec = tubes.uniform_tube(num_steps=10)
    with Solver(spec, precision=64) as s:
        with pytest.raises(_abi.VocalSimError):
            s.set_option(_abi.OPT_KERNEL, _abi.K_FAST)


def test_fast_odd_cells_per_lane(hip_device):
    # tube 31 of the bench batch: 16 185 padded cells, so K = 17, 19, 21 and
    # 23 at 1024, 960, 896 and 832 threads (K is rounded up to odd; K = 17 is
    # the instantiation built in its own unit under another scheduler).  The
    # flux kernel's results do not depend on the work-group shape, and all
    # track the FP64 oracle
    spec = tubes.batch_tubes(256, seed=0, num_steps=700)[31]
    want = {1024: 17, 960: 19, 896: 21, 832: 23}
    runs = {}
    for threads, K in want.items():
        print(f"[fast odd K] threads={threads} K={K}", flush=True)
        with Solver(spec, precision=32, threads=threads, kernel=_abi.K_FAST, steps_per_launch=256) as s:
            d = s.dims[0]
            padded = (d["frameH"] - 2) * (d["frameW"] - 1)
            k = -(-padded // (threads - 64))
            assert k + (k % 2 == 0) == K
            s.run()
            runs[threads] = (s.probe(), s.field("pressure"), s.field("vx"), s.field("vy"))
    for threads in (960, 896, 832):
        for a, b in zip(runs[1024], runs[threads]):
            assert np.array_equal(a, b)
    p64, _ = orc.OracleTube(spec, 64).run()
    assert rel_l2(runs[1024][0], p64) <= 1e-4


def _full_batch_gaussian():
    # the batch workload at full size (256 ragged tubes, 19 845 steps = 30 ms)
    # with every tube on the Gaussian source, so the probes are linear in its
    # amplitude
    base = tubes.batch_tubes(256, seed=0, num_steps=19845)
    g = base[0]
    return [sp.replace(source_type=tubes.SOURCE_GAUSSIAN, gauss_amp=g.gauss_amp, gauss_t0=g.gauss_t0,
                       gauss_width=g.gauss_width) for sp in base]


def _batch_probes(specs, precision, kernel, amp):
    run = [sp.replace(gauss_amp=sp.gauss_amp * amp) for sp in specs]
    with Solver(run, precision=precision, threads=1024, kernel=kernel, steps_per_launch=256) as s:
        s.set_option(_abi.OPT_USE_GRAPH, 1)
        s.run()
        assert s.steps_done == 19845
        return [s.probe(i) for i in range(len(run))]


def test_full_batch_fp32_determinism_linearity_parity(hip_device):
    # The FP32 on-chip kernel on the full batch, 256-step launches replayed
    # as graphs.  A second identical run is bit-identical (no races).  The
    # probes are linear in the source amplitude only up to FP32 rounding:
    # where the wave front passes through subnormal magnitudes, products
    # round differently for amplitudes a and 2a, after which the two runs
    # carry independent FP32 rounding noise (the same 1e-4..1e-3 as FP32
    # against FP64; the FP32 oracle behaves the same way).  Two tubes are
    # checked against the FP64 oracle over the whole run.
    specs = _full_batch_gaussian()
    print("[full batch FP32] amplitude x1, twice, and x2", flush=True)
    a = _batch_probes(specs, 32, _abi.K_FAST, 1.0)
    a2 = _batch_probes(specs, 32, _abi.K_FAST, 1.0)
    b = _batch_probes(specs, 32, _abi.K_FAST, 2.0)
    worst = 0.0
    for x, y, z in zip(a, a2, b):
        assert np.array_equal(x, y)
        assert np.isfinite(x).all() and np.abs(x).max() > 0
        worst = max(worst, float(np.max(np.abs(z - 2 * x)) / np.max(np.abs(x))))
    print(f"[full batch FP32] max |p(2a) - 2 p(a)| / max |p(a)| = {worst:.3e}", flush=True)
    assert worst <= 5e-3
    for i in (0, 255):
        p64, _ = orc.OracleTube(specs[i], 64).run()
        e = rel_l2(a[i], p64)
        print(f"[full batch FP32] tube {i}: FP32 on chip vs FP64 oracle rel L2 = {e:.2e}", flush=True)
        assert e <= 1e-3


def test_full_batch_fp64_exact_linearity_and_parity(hip_device):
    # The FP64 exact kernel on the same full batch: no value of these runs
    # comes near the FP64 subnormal range, so doubling the amplitude doubles
    # every probe sample exactly, and two tubes equal the FP64 oracle's
    # FDTDEngineCPU restatement over the whole run (1e-10 relative L2 bar)
    specs = _full_batch_gaussian()
    print("[full batch FP64] amplitude x1 and x2", flush=True)
    a = _batch_probes(specs, 64, _abi.K_EXACT, 1.0)
    b = _batch_probes(specs, 64, _abi.K_EXACT, 2.0)
    for x, z in zip(a, b):
        assert np.abs(x).max() > 0
        assert np.array_equal(z, 2 * x)
    for i in (0, 255):
        p64, _ = orc.OracleTube(specs[i], 64).run()
        e = rel_l2(a[i], p64)
        print(f"[full batch FP64] tube {i}: exact kernel vs FP64 oracle rel L2 = {e:.2e}", flush=True)
        assert e <= 1e-10


WIDE_CASES = {
    "batch16_mixed": tubes.batch_tubes(16, seed=4, num_steps=700),
    "open_mouth": [tubes.uniform_tube(num_steps=700).replace(mouth_dirichlet=0)],
    "no_pml_vowel": [tubes.vowel_tube(1, num_steps=700, source=tubes.SOURCE_GAUSSIAN).replace(pml=0)],
    "vowel_twomass": [tubes.vowel_tube(0, num_steps=900)],
    "sine": [tubes.vowel_tube(2, num_steps=700, source=tubes.SOURCE_SINE)],
    # 4 cm of 0.5 cm^2: about 1 700 padded cells, the 4-cells-per-lane shape
    "short_narrow": [tubes.uniform_tube(num_steps=700).replace(lengths_cm=np.array([4.0]),
                                                                areas_cm2=np.array([0.5]))],
    # random tubes at four sample rates, 2D and 2.5D, that the on-chip kernel
    # holds (about 100 pair classes); built when the case runs
    "random_mixed_rates": None,
    # source columns in different groups of four run slots (PML on and off):
    # every group handles the source
    "mixed_pml": [tubes.vowel_tube(1, num_steps=700, source=tubes.SOURCE_GAUSSIAN).replace(pml=0),
                  tubes.vowel_tube(2, num_steps=700, source=tubes.SOURCE_GAUSSIAN)],
}


def _random_mixed_rates():
    def padded(sp):
        d = orc.OracleTube(sp, 32).dims
        return (d["frameH"] - 2) * (d["frameW"] - 1)
    return [sp for sp in random_specs.random_specs(24, 5, num_steps=500) if padded(sp) <= 18000]


@pytest.mark.parametrize("threads", [1024, 512])
@pytest.mark.parametrize("name", list(WIDE_CASES))
def test_wide_layout_equals_odd_layout(hip_device, name, threads):
    # the wide-access form (rows padded to 4 cells, 4 (mod 8) cells per lane,
    # 16-byte LDS accesses, select-free face update) gives the values of the
    # odd-run form: probes, source traces and final states compare equal
    specs = WIDE_CASES[name] or _random_mixed_rates()
    runs = {}
    for wide in (1, 0):
        with Solver(specs, precision=32, threads=threads, kernel=_abi.K_FAST, steps_per_launch=256) as s:
            s.set_option(_abi.OPT_FAST_WIDE, wide)
            K = s.get_option(_abi.OPT_FAST_LAYOUT_K)
            if wide:
                if K == 0 or K % 8 != 4:
                    pytest.skip(f"the wide form does not hold this batch at {threads} threads (K = {K})")
                if name == "short_narrow" and threads == 1024:
                    assert K == 4
            else:
                assert K % 2 == 1
            print(f"[wide vs odd] case={name} threads={threads} wide={wide} K={K}", flush=True)
            s.run()
            # the layout that ran: the wide form reads its class tables by
            # cell pair and holds at most 128 pair classes
            npairs = s.get_option(_abi.OPT_FAST_PAIR_CLASSES)
            q = s.get_option(_abi.OPT_FAST_ROW_QUANTUM)
            print(f"[wide vs odd]   pair classes {npairs}, rows padded to a multiple of {q}", flush=True)
            assert 1 <= npairs <= 128
            assert s.get_option(_abi.OPT_FAST_LAYOUT_K) == K
            assert q == (K if wide else 0)
            runs[wide] = [(s.probe(i), s.source_trace(i), s.field("pressure", i), s.field("vx", i), s.field("vy", i))
                          for i in range(len(specs))]
    for a, b in zip(runs[1], runs[0]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    p64, _ = orc.OracleTube(specs[0], 64).run()
    assert rel_l2(runs[1][0][0], p64) <= 1e-3


def test_wide_compact_rows_equal_odd_layout(hip_device):
    # a 20 cm tube (49 x 281 interior cells) at 768 threads: its rows padded
    # to a multiple of 20 cells need 14 700 slots, more than the 14 080 of
    # 20 cells on 704 lanes, so the wide form pads them to a multiple of 4
    # (the source handled in every group of four cells); it equals the
    # odd-run form (K = 21)
    base = tubes.vowel_tube(1, num_steps=700, source=tubes.SOURCE_GAUSSIAN)
    spec = base.replace(lengths_cm=base.lengths_cm * (20.0 / base.lengths_cm.sum()))
    runs = {}
    for wide in (1, 0):
        with Solver([spec], precision=32, threads=768, kernel=_abi.K_FAST, steps_per_launch=256) as s:
            s.set_option(_abi.OPT_FAST_WIDE, wide)
            s.run()
            K, q = s.get_option(_abi.OPT_FAST_LAYOUT_K), s.get_option(_abi.OPT_FAST_ROW_QUANTUM)
            print(f"[compact rows] wide={wide} K={K} row quantum={q}", flush=True)
            assert (K, q) == ((20, 4) if wide else (21, 0))
            runs[wide] = (s.probe(0), s.source_trace(0), s.field("pressure", 0), s.field("vx", 0), s.field("vy", 0))
    for x, y in zip(runs[1], runs[0]):
        assert np.array_equal(x, y)


def test_wide_layout_graph_replays(hip_device):
    specs = tubes.batch_tubes(8, seed=6, num_steps=1100)
    out = []
    for graph in (0, 1):
        with Solver(specs, precision=32, threads=1024, kernel=_abi.K_FAST, steps_per_launch=128) as s:
            s.set_option(_abi.OPT_USE_GRAPH, graph)
            assert s.get_option(_abi.OPT_FAST_LAYOUT_K) % 8 == 4
            s.run()
            out.append([s.probe(i) for i in range(len(specs))])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
