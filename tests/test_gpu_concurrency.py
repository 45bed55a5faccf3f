"""Handles stepped from several host threads at once.

Each handle owns its stream, buffers and captured graphs, and the library's
shared state (the autotuner's and the adjoint forward's shape caches, the
per-kernel LDS limits, the device memory kept for reuse after a handle is
destroyed, the error string) is locked, atomic or thread-local.  So batches
created and stepped concurrently must give the probes they give one after
another, bit for bit.  The set covers graph replays of the on-chip FP32
kernel, the cooperative FP64 kernel (whose grid barrier has a bounded wait),
a single-plane tube (its max-chord search runs on the handle's stream during
creation), a small FP64 batch on the per-step exact kernel and an FP32
gradient pass on the on-chip reverse kernels (record slabs from the kept
memory, taken and returned by every thread).
"""
import threading

import numpy as np
import pytest

from vocalsim import Solver, _abi, tubes

pytestmark = pytest.mark.gpu


def jobs():
    plane = tubes.vowel_tube(5, num_steps=900, source=tubes.SOURCE_GAUSSIAN, n_sections=6).replace(shape=3,
                                                                                                   aspect=None)
    return {
        "fp32 batch, graphs": (tubes.batch_tubes(6, seed=9, num_steps=1200), 32, _abi.K_FAST, 1024, True),
        "fp64 two-mass, cooperative": ([tubes.vowel_tube(0, num_steps=1200)], 64, _abi.K_EXACT, 1024, False),
        "fp32 single-plane": ([plane], 32, _abi.K_FAST, 1024, False),
        "fp64 batch": (tubes.batch_tubes(3, seed=4, num_steps=700), 64, _abi.K_EXACT, 256, False),
        "fp32 gradient, on chip": ([tubes.vowel_tube(k, num_steps=700, source=tubes.SOURCE_GAUSSIAN)
                                    for k in (2, 3)], 32, None, 1024, False),
    }


def run(job):
    specs, precision, kernel, threads, graph = job
    if kernel is None:
        # probes and the three gradients of J = sum_n w[n] probe[n]
        rng = np.random.default_rng(11)
        w = [rng.standard_normal(sp.num_steps) for sp in specs]
        with Solver(specs, precision=precision) as s:
            probes = s.adjoint_forward(256)
            grads = s.adjoint_backward(w)
        return list(probes) + [g for gs in grads for g in gs]
    with Solver(specs, precision=precision, threads=threads, kernel=kernel) as s:
        if graph:
            s.set_option(_abi.OPT_USE_GRAPH, 1)
        s.run()
        return [s.probe(i) for i in range(len(specs))]


def test_handles_on_threads_match_sequential_runs(hip_device):
    todo = jobs()
    print("[concurrency] sequential runs", flush=True)
    ref = {name: run(job) for name, job in todo.items()}
    out, errors = {}, []

    def worker(name, job, key):
        try:
            out[key] = run(job)
        except Exception as e:   # reported below, in the main thread
            errors.append((key, repr(e)))

    for round_ in range(3):
        print(f"[concurrency] round {round_}: {2 * len(todo)} threads, every job twice", flush=True)
        threads = [threading.Thread(target=worker, args=(name, job, (name, 2 * round_ + r)))
                   for r in range(2) for name, job in todo.items()]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=300)
            assert not t.is_alive(), "a thread did not finish"
        assert not errors, errors
    for (name, r), probes in out.items():
        assert len(probes) == len(ref[name])
        for i, (a, b) in enumerate(zip(probes, ref[name])):
            assert np.array_equal(a, b), f"{name} (run {r}), tube {i}: probe differs from the sequential run"


def test_small_handles_take_small_chunks(hip_device):
    """A handle takes its field memory in chunks sized from its batch (4 MiB
    for one tube), so many live one-tube handles hold a few MB each; the
    memory stays with the library after they close until it is released."""
    import torch
    import vocalsim
    vocalsim.release_cached_memory()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    spec = tubes.vowel_tube(1, num_steps=2000)
    handles = [Solver([spec], precision=64) for _ in range(16)]
    torch.cuda.synchronize()
    held = free0 - torch.cuda.mem_get_info()[0]
    print(f"[small handles] 16 one-tube FP64 handles hold {held / 2**20:.0f} MiB", flush=True)
    for h in handles:
        h.close()
    vocalsim.release_cached_memory()
    assert held <= 16 * (32 << 20)
