"""Exact step of a single tube spread over several work-groups of one
cooperative launch (VS_OPT_EXACT_GROUPS), several steps per grid barrier
(VS_OPT_EXACT_STEPS) with recomputed halo columns: the per-cell code is the
same as the one-work-group kernel, so probes and whole fields must be
bit-identical for any group and step count, in FP64 and FP32, with Gaussian
and two-mass sources, with launches that end inside a block of steps, and
with the source in the middle of the domain; the FP64 run must equal the
CPU oracle exactly."""
import numpy as np
import pytest

import oracle as orc
from vocalsim import Solver, _abi, tubes

pytestmark = pytest.mark.gpu


def run(spec, precision, groups, threads=1024, steps=None, blocks=1, per_barrier=0):
    with Solver(spec, precision=precision, threads=threads, steps_per_launch=max(1, (steps or spec.num_steps) // blocks)) as s:
        s.set_option(_abi.OPT_EXACT_GROUPS, groups)
        s.set_option(_abi.OPT_EXACT_STEPS, per_barrier)
        s.run(steps)
        return s.probe(), s.field("pressure"), s.field("vx"), s.field("vy")


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("source", [tubes.SOURCE_GAUSSIAN, tubes.SOURCE_TWO_MASS])
def test_groups_are_bit_identical(hip_device, precision, source):
    spec = tubes.vowel_tube(2, num_steps=900, source=source)
    ref = run(spec, precision, 1)
    for groups in (2, 3, 8, 16):
        got = run(spec, precision, groups, blocks=3)
        print(f"[coop] precision={precision} source={source} groups={groups}", flush=True)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("precision,source", [(64, tubes.SOURCE_TWO_MASS), (32, tubes.SOURCE_GAUSSIAN)])
def test_steps_per_barrier_are_bit_identical(hip_device, precision, source):
    # 901 steps in launches of 300: every launch ends inside a block of
    # steps for 4 and 8, and the last launch is a single step
    spec = tubes.vowel_tube(4, num_steps=901, source=source)
    ref = run(spec, precision, 1)
    for groups, per_barrier in ((2, 2), (8, 3), (14, 4), (16, 8), (24, 1), (0, 0), (0, 6)):
        got = run(spec, precision, groups, blocks=3, per_barrier=per_barrier)
        print(f"[coop] precision={precision} source={source} groups={groups} steps={per_barrier}", flush=True)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("source", [tubes.SOURCE_GAUSSIAN, tubes.SOURCE_TWO_MASS])
def test_source_in_the_middle(hip_device, source):
    # open space: the point source sits in column 176 of 355, so some group
    # counts put a column boundary next to it and lower the steps per barrier
    spec = tubes.open_space(num_steps=400, source_type=source)
    ref = run(spec, 64, 1)
    for groups in (2, 7, 8, 15, 16, 32):
        got = run(spec, 64, groups, blocks=2, per_barrier=4)
        print(f"[coop] open space source={source} groups={groups}", flush=True)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [64, 32])
def test_graph_replays_mix_with_cooperative_launches(hip_device, precision):
    # graph replays keep the one-work-group kernel, the remainders between
    # them run cooperatively; the source state handed from one to the other
    # (excitation pressure, two-mass state) must keep the run bit-identical
    spec = tubes.vowel_tube(6, num_steps=1500)
    ref = run(spec, precision, 1)
    print(f"[coop] graph mix precision={precision}", flush=True)
    with Solver(spec, precision=precision, threads=1024, steps_per_launch=128) as s:
        s.set_option(_abi.OPT_USE_GRAPH, 1)
        for n in (300, 700, 500):   # cooperative; a 512-step replay + 188 cooperative; cooperative
            s.run(n)
        assert s.steps_done == 1500
        got = s.probe(), s.field("pressure"), s.field("vx"), s.field("vy")
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_barrier_timeout_is_reported_and_recoverable(hip_device):
    # the grid barrier's wait is bounded: with a limit of one poll it gives
    # up at once (every index is independent of the data, so the launch
    # still runs to its end), vs_sync reports the timeout, and after a reset
    # at the default limit the handle reproduces the reference bit for bit
    spec = tubes.vowel_tube(2, num_steps=600, source=tubes.SOURCE_GAUSSIAN)
    ref = run(spec, 64, 1)
    with Solver(spec, precision=64, threads=1024, steps_per_launch=300) as s:
        assert s.get_option(_abi.OPT_BARRIER_SPIN) == 0
        with pytest.raises(Exception):
            s.set_option(_abi.OPT_BARRIER_SPIN, -1)
        s.set_option(_abi.OPT_BARRIER_SPIN, 1)
        print("[coop] barrier spin limit 1", flush=True)
        with pytest.raises(_abi.VocalSimError):
            s.run()
        s.set_option(_abi.OPT_BARRIER_SPIN, 0)
        s.reset()
        s.run()
        got = s.probe(), s.field("pressure"), s.field("vx"), s.field("vy")
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_groups_match_the_oracle_fp64(hip_device):
    spec = tubes.vowel_tube(5, num_steps=600, source=tubes.SOURCE_TWO_MASS)
    probe, P, _, _ = run(spec, 64, 8, threads=512)
    p_ref, _ = orc.OracleTube(spec, 64).run()
    print(f"[coop] vs oracle max |diff| = {np.abs(probe - p_ref).max():.3e}", flush=True)
    assert np.array_equal(probe, p_ref)


def test_batches_keep_one_group_per_tube(hip_device):
    a = tubes.vowel_tube(1, num_steps=400, source=tubes.SOURCE_GAUSSIAN)
    b = tubes.vowel_tube(3, num_steps=400, source=tubes.SOURCE_GAUSSIAN)
    with Solver([a, b], precision=64) as s:
        s.set_option(_abi.OPT_EXACT_GROUPS, 8)
        s.run()
        pa, pb = s.probe(0), s.probe(1)
    assert np.array_equal(pa, run(a, 64, 1, threads=256)[0])
    assert np.array_equal(pb, run(b, 64, 1, threads=256)[0])


def test_option_range(hip_device):
    spec = tubes.vowel_tube(0, num_steps=10)
    with Solver(spec, precision=64) as s:
        assert s.get_option(_abi.OPT_EXACT_GROUPS) == 0
        for bad in (-1, 65):
            with pytest.raises(Exception):
                s.set_option(_abi.OPT_EXACT_GROUPS, bad)
        assert s.get_option(_abi.OPT_EXACT_STEPS) == 0
        for bad in (-1, 33):
            with pytest.raises(Exception):
                s.set_option(_abi.OPT_EXACT_STEPS, bad)
        s.set_option(_abi.OPT_EXACT_STEPS, 32)
        assert s.get_option(_abi.OPT_EXACT_STEPS) == 32
