"""Batches whose tubes stop at different steps.

A tube stops at its own num_steps while the rest of the batch goes on
(tend = min(t0 + n, num_steps) per tube).  Stepped in pieces that cross
launch and graph-replay boundaries, every tube of such a batch must give
the probe and final state it gives alone, bit for bit: on the FP32 on-chip
kernel (whose cells-per-lane layout follows the batch's largest tube and
does not change any value) and on the FP64 exact kernel (a batch steps on
the per-step kernel, a single tube on the cooperative one; both follow the
reference's operation order).
"""
import numpy as np
import pytest

from vocalsim import Solver, _abi, tubes

pytestmark = pytest.mark.gpu

COUNTS = [1100, 700, 257, 1]


def specs():
    base = tubes.batch_tubes(4, seed=12, num_steps=1)
    # one two-mass tube among them (the batch builder makes the second half two-mass)
    return [sp.replace(num_steps=n) for sp, n in zip(base, COUNTS)]


def shape(precision):
    return (_abi.K_FAST, 1024) if precision == 32 else (_abi.K_EXACT, 256)


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("precision", [32, 64])
def test_tubes_stop_at_their_own_step_count(hip_device, precision, graph):
    kernel, threads = shape(precision)
    sp = specs()
    print(f"[mixed step counts] precision={precision} graph={graph} counts={COUNTS}", flush=True)
    # 64-step launches in graphs of 4 launches: a replay covers 256 steps, so
    # with graphs on, the pieces below run whole replays (from steps 0, 300
    # and 429) followed by plain launches for the rest of each piece, and the
    # last piece runs past every tube's end
    with Solver(sp, precision=precision, threads=threads, kernel=kernel, steps_per_launch=64) as s:
        s.set_option(_abi.OPT_USE_GRAPH, graph)
        for n in (300, 129, 511, 5000):
            s.run(n)
        assert s.steps_done == max(COUNTS)
        got = [(s.probe(i), s.field("pressure", i), s.field("vx", i), s.field("vy", i)) for i in range(len(sp))]
    for i, one in enumerate(sp):
        with Solver([one], precision=precision, threads=threads, kernel=kernel, steps_per_launch=128) as s1:
            s1.run()
            ref = (s1.probe(0), s1.field("pressure", 0), s1.field("vx", 0), s1.field("vy", 0))
        assert got[i][0].size == COUNTS[i]
        for name, a, b in zip(("probe", "P", "Vx", "Vy"), got[i], ref):
            assert np.array_equal(a, b), f"tube {i} ({COUNTS[i]} steps): {name} differs from the tube run alone"


def test_batch_larger_than_the_chip(hip_device):
    # 300 tubes, one work-group each on a 256-CU device: the on-chip kernel's
    # work-groups are independent (no grid-wide wait), so those that do not
    # fit the first wave run after the largest tubes (the launch order is
    # largest first) and give the same values
    # the bench batch's first 150 geometries (all within the on-chip
    # kernel's 1024-thread capacity), twice
    half = tubes.batch_tubes(150, seed=0, num_steps=300)
    sp = half + half[::-1]
    print("[batch 300] FP32 on chip with graphs", flush=True)
    with Solver(sp, precision=32, threads=1024, kernel=_abi.K_FAST, steps_per_launch=64) as s:
        s.set_option(_abi.OPT_USE_GRAPH, 1)
        s.run()
        got = {i: s.probe(i) for i in (0, 149, 150, 299)}
    for i, p in got.items():
        with Solver([sp[i]], precision=32, threads=1024, kernel=_abi.K_FAST, steps_per_launch=64) as s1:
            s1.run()
            assert np.array_equal(p, s1.probe(0)), f"tube {i}"
