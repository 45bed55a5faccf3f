"""hipGraph-captured step blocks and the autotuner, through the C ABI.

Graph replays and ordinary launches run the same kernels on the same state,
so their probes must agree bit for bit; so must runs at different work-group
shapes (every cell's update is the same arithmetic whatever lane owns it)."""
import numpy as np
import pytest

from vocalsim import Solver, tubes
from vocalsim import _abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,kernel", [(32, _abi.K_FAST), (64, _abi.K_EXACT), (32, _abi.K_EXACT)])
def test_graph_replay_equals_launches(hip_device, precision, kernel):
    specs = tubes.batch_tubes(12, seed=5, num_steps=1500)
    print(f"[graph] precision={precision} kernel={kernel}", flush=True)
    with Solver(specs, precision=precision, threads=1024, steps_per_launch=128, kernel=kernel) as a:
        a.run(1300)
        pa = [a.probe(i) for i in range(len(specs))]
    with Solver(specs, precision=precision, threads=1024, steps_per_launch=128, kernel=kernel) as b:
        b.set_option(_abi.OPT_USE_GRAPH, 1)
        b.run(700)   # one replay of 4 x 128 steps + remainder
        b.run(600)
        pb = [b.probe(i) for i in range(len(specs))]
        assert b.steps_done == 1300
    for x, y in zip(pa, pb):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("graph", [0, 1])
def test_run_that_launches_nothing_syncs(hip_device, graph):
    # a zero-step run on a fresh handle (and a graph-mode run with no step
    # left) records no events: vs_sync succeeds and reports zero time
    spec = tubes.uniform_tube(num_steps=300)
    print(f"[zero-step run] graph={graph}", flush=True)
    with Solver(spec, precision=32, threads=256, steps_per_launch=128) as s:
        s.set_option(_abi.OPT_USE_GRAPH, graph)
        s.run(0)
        assert s.timing()[0] == 0.0 and s.steps_done == 0
        s.run(300)
        assert s.steps_done == 300
        s.run(50)   # nothing left to run
        assert s.steps_done == 300 and s.timing()[0] == 0.0


def test_fast_kernel_independent_of_workgroup_shape(hip_device):
    specs = tubes.batch_tubes(6, seed=6, num_steps=500)
    probes = []
    for t in (1024, 896, 768):
        with Solver(specs, precision=32, threads=t, kernel=_abi.K_FAST) as s:
            s.run()
            probes.append([s.probe(i) for i in range(len(specs))])
    for other in probes[1:]:
        for x, y in zip(probes[0], other):
            assert np.array_equal(x, y)


def test_autotune_picks_a_valid_shape_and_resets(hip_device):
    specs = tubes.batch_tubes(32, seed=7, num_steps=600)
    with Solver(specs, precision=32, threads=1024, kernel=_abi.K_FAST) as s:
        rep = s.autotune(64)
        print("[autotune]", rep, flush=True)
        assert s.steps_done == 0
        assert rep["threads"] % 64 == 0 and 128 <= rep["threads"] <= 1024
        assert s.get_option(_abi.OPT_THREADS) == rep["threads"]
        s.run()
        tuned = [s.probe(i) for i in range(len(specs))]
        assert len(rep["candidates"]) >= 2
        rep2 = s.autotune(64) if s.steps_done == 0 else None
    with Solver(specs, precision=32, threads=1024, kernel=_abi.K_FAST) as s2:
        s2.run()
        for i in range(len(specs)):
            assert np.array_equal(s2.probe(i), tuned[i])


def test_autotune_refuses_after_steps(hip_device):
    spec = tubes.uniform_tube(num_steps=100)
    with Solver(spec, precision=64) as s:
        s.run(10)
        with pytest.raises(_abi.VocalSimError):
            s.autotune()
