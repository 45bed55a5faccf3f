"""Benchmark: grid-cell-updates/s of the batched FP32 forward (256 ragged tubes,
13-19 cm, 30 ms) against the plain FP64 single-tube GPU path, on one MI355X.

One bench "step" is one 256-time-step block advanced over the whole batch
(the unit the step kernel / graph replays).  Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vocalsim-2p5d-fdtd_amd"))

BYTES_FP32 = 37   # P, Vx, Vy read+write, dP, dX, dY read (4 B each), 1 B class
BYTES_FP64 = 73
HBM_PEAK_GBS = 8000.0
# LDS bytes per cell and step of the on-chip FP32 step, from its compiled
# code.  Odd-run form: reads P, Uy, Uy below, aP, P above, P right (4 B
# each), (ax, ay) (8 B), the old Uy (4 B); writes P', Uy' (4 B each).  Wide
# form (K = 4 mod 8): reads P, Uy, Uy below, P, P above, Uy (4 B each, as
# 16-byte vectors), per pair of cells (aP, aP) (8 B) and (ax, ay, ax, ay)
# (16 B), a run end per K cells; writes P', Uy' (4 B each): 57 600 B per
# wave and step at K = 20, i.e. 45 B per cell slot.  Rated on real cells
# only; the peak is measured live (vs_measure_lds_bandwidth).
LDS_BYTES_PER_CELL = {"odd": 44, "wide": 45}
# algorithmic flops per cell-update (pressure 15, Vx 18, Vy 18; counted from
# FDTDEngineCPU.cpp:320-405) and MI355X's published dense FP32 vector peak
FLOPS_PER_CELL = 51
FP32_PEAK_TFLOPS = 157.3
# HBM bytes the on-chip kernel must move per real cell and launch: loads P,
# Ux, Uy, kP, gX, gY (4 B each) and the class byte; stores P, Ux, Uy, Vx, Vy
# (4 B each) and reads dX, dY for the velocity export
COMPULSORY_BYTES_PER_CELL = 6 * 4 + 1 + 5 * 4 + 2 * 4
BLOCK = 256
GRAPH_BLOCKS = 4   # 256-step launches per captured step graph (the library's default)
# HBM bytes per 256-step launch of k_fast_wide<20> on the 256-tube batch at
# 1024 threads (the autotuned shape), from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE passes, each scaled by the same counter over a plain copy kernel
# of known size (FETCH x2.000, WRITE x1.000): 102.4 MB read, 62.7 MB written;
# profiles/k_fast_wide_batch256_fp32_pmc.md.
PMC_TRAFFIC_BYTES = 165.06e6   # k_fast_wide<20, 1>
PMC_THREADS = 1024
PMC_LAYOUT_K = 20
# Fraction of the time each pipe of k_fast_wide<20> is busy on this batch, from
# rocprofv3 SQ counters (profiles/k_fast_wide_batch256_fp32_sq.md): VALU =
# SQ_ACTIVE_INST_VALU x 4 waves per SIMD / SQ_WAVE_CYCLES, LDS =
# SQ_LDS_IDX_ACTIVE / cycles per CU.
PIPES_BUSY = {"valu": 0.77, "lds": 0.69}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed 256-step blocks")
    # 24 untimed blocks: the device's clocks ramp over the first 20-25
    # launches after it idles (the per-launch time falls by about 10 %,
    # whatever the data), so the timed blocks run at the sustained rate
    ap.add_argument("--warmup", type=int, default=24,
                    help="untimed 256-step blocks (4 or more also captures the step graph before timing)")
    ap.add_argument("--tubes", type=int, default=256)
    ap.add_argument("--threads", type=int, default=0, help="0: autotuned / default")
    ap.add_argument("--kernel", type=int, default=-1, help="-1: best available")
    ap.add_argument("--fp64-blocks", type=int, default=8, help="0: skip the plain FP64 leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-tubes", type=int, default=64)
    ap.add_argument("--adj-steps", type=int, default=13230)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="counter calibration only: 5 launches of a plain FP32 copy kernel over this many MiB")
    return ap.parse_args()


def host_cpu():
    """Model name and core counts of the host the CPU baselines run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 0
    return {"model": model, "nproc": usable, "logical_cpus": os.cpu_count() or 0,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def copy_calibration(mib):
    """Plain elementwise copy (y = x * 1) of a known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters of the step kernels."""
    import torch
    torch.cuda.set_device(0)
    x = torch.rand(mib * (1 << 18), device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(5):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel_bytes_read": mib << 20, "copy_kernel_bytes_written": mib << 20,
                      "launches": 5}), flush=True)


def adjoint_leg(args, np, torch, vocalsim, tubes, _abi):
    """Loss = mean squared difference of ln|Y(f)| (0-5 kHz, 1024 bins, FP64 DFT
    of the probe) against targets from other seeds; FP32 forward with 512-step
    checkpoints, then the reverse-time pass to area, depth scale and
    admittance of every section."""
    from vocalsim.adjoint import dft_basis, log_magnitude, tube_probe
    n, B = args.adj_steps, args.adj_tubes
    fs = 44100.0 * tubes.SRATE_MUL

    def spec_of(seed):
        sp = tubes.vowel_tube(seed, num_steps=n, source=tubes.SOURCE_GAUSSIAN)
        return sp.replace(depth_scale=np.ones(sp.num_sections, np.float32),
                          mu3d=np.full(sp.num_sections, 0.005, np.float32))

    specs = [spec_of(k) for k in range(B)]
    basis = dft_basis(n, fs, np.linspace(0.0, 5000.0, 1024))
    with vocalsim.Solver([spec_of(64 + k) for k in range(B)], precision=32, threads=1024,
                         kernel=_abi.K_FAST) as s:
        s.run()
        target = log_magnitude(torch.tensor(np.stack([s.probe(i) for i in range(B)]), device="cuda"), basis)
        cells = s.cells_per_step
    dev = "cuda"

    def one_pass():
        A = torch.tensor(np.stack([sp.areas_cm2 for sp in specs]), dtype=torch.float64, device=dev).requires_grad_(True)
        D = torch.tensor(np.stack([sp.depth_scale for sp in specs]), dtype=torch.float64, device=dev).requires_grad_(True)
        M = torch.tensor(np.stack([sp.mu3d for sp in specs]), dtype=torch.float64, device=dev).requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p = tube_probe(A, D, M, specs, 32, 512)
        loss = ((log_magnitude(p, basis) - target) ** 2).mean()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return A, loss, t0, t1, t2

    one_pass()   # untimed: first-call costs (kernel loading, library GEMM selection)
    A, loss, t0, t1, t2 = one_pass()
    from vocalsim.adjoint import last_backward_timing
    rev_ms, chain_ms = last_backward_timing()
    out = {"workload": f"{B} seeded 44-section vowel tubes, Gaussian pulse, {n} steps; loss = mean squared "
                       f"ln|Y(f)| difference over 1024 bins 0-5 kHz against {B} other seeds",
           "precision": "f32 state, f64 gradient accumulation", "checkpoint_every": 512,
           "timing": "second of two identical passes",
           "cells_per_step": cells, "forward_ms": (t1 - t0) * 1e3, "backward_ms": (t2 - t1) * 1e3,
           "forward_cell_updates_per_s": cells * n / (t1 - t0),
           "gradient_cell_updates_per_s": cells * n / (t2 - t0),
           "backward_split_ms": {"reverse_pass_on_device": rev_ms, "host_chain_to_sections": chain_ms},
           "loss": float(loss.detach()), "grad_norm_area": float(A.grad.norm())}
    log(f"[bench] gradient workload: forward {out['forward_ms']:.0f} ms, backward {out['backward_ms']:.0f} ms "
        f"(reverse pass {rev_ms:.0f} ms, host chain {chain_ms:.1f} ms; {B} tubes x {n} steps)")
    return out


def main():
    args = parse()
    if args.copy_mib:
        copy_calibration(args.copy_mib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(tubes would shard by batch).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import vocalsim
    from vocalsim import _abi, tubes

    torch.cuda.set_device(0)
    vocalsim.load()

    # ---------------- batched FP32 (256 ragged tubes) ----------------
    total_blocks = args.warmup + args.steps
    # 19 845 steps (0.03 s at 661.5 kHz), or more when warm-up, timed blocks
    # and the 4 per-launch timed blocks after them need it
    nsteps = max(19845, (total_blocks + 4) * BLOCK)
    specs = tubes.batch_tubes(args.tubes, seed=0, num_steps=nsteps)
    kernel = args.kernel if args.kernel >= 0 else _abi.K_FAST
    threads = args.threads or 1024
    log(f"[bench] building {args.tubes} tubes (FP32), kernel={kernel}")
    s = vocalsim.Solver(specs, precision=32, threads=threads, steps_per_launch=BLOCK, kernel=kernel)
    if not args.threads:
        # work-group shape from the library's autotuner (timed on this batch);
        # the launch length stays at the 256-step block
        rep = s.autotune(256)
        threads = s.get_option(_abi.OPT_THREADS)
        s.set_option(_abi.OPT_STEPS_PER_LAUNCH, BLOCK)
        log("[bench] autotune: " + ", ".join(f"{c['threads']}/{c['steps_per_launch']}: {c['us_per_step']:.2f} us"
                                            for c in rep.get("candidates", [])) + f" -> threads {threads}")
    cells = s.cells_per_step
    dims = s.dims
    log(f"[bench] interior cells per step: {cells}, threads {threads}")
    # 256-step blocks replayed as hipGraphs of GRAPH_BLOCKS launches (the
    # remainder, if any, as ordinary launches); the warm-up captures the graph
    s.set_option(_abi.OPT_USE_GRAPH, 1)
    s.run(args.warmup * BLOCK)
    torch.cuda.synchronize()
    s.set_option(_abi.OPT_TIMING, 0)
    t0 = time.perf_counter()
    s.run(args.steps * BLOCK, sync=False)
    s.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    ev_ms, launches, _ = s.timing()
    value = cells * args.steps * BLOCK / wall
    log(f"[bench] batched FP32: {wall*1e3:.1f} ms wall ({ev_ms:.1f} ms events) for {args.steps} blocks "
        f"-> {value:.3e} cell-updates/s")
    # the kernel's mean launch duration over the timed region: HIP events on
    # the library's stream around the timed launches (graph replays, one
    # 256-step launch per block), divided by the launch count
    if launches != args.steps or not ev_ms > 0:
        raise RuntimeError(f"timed region ran {launches} launches ({ev_ms} ms of events)")
    mean_launch_ms = ev_ms / launches
    # beside it: 4 more launches, each timed alone with HIP events
    s.set_option(_abi.OPT_TIMING, 1)
    s.run(4 * BLOCK)
    _, nl, single_launch_ms = s.timing()
    if nl != 4 or not single_launch_ms > 0:
        raise RuntimeError(f"per-launch timing ran {nl} launches ({single_launch_ms} ms)")
    # the form that ran (the pair-class count is known once the kernel has run)
    layout_k = s.get_option(_abi.OPT_FAST_LAYOUT_K) if kernel == _abi.K_FAST else 0
    layout = "wide" if layout_k % 8 == 4 else "odd"
    pair_classes = s.get_option(_abi.OPT_FAST_PAIR_CLASSES) if kernel == _abi.K_FAST else 0
    log(f"[bench] on-chip layout: {layout} form, K = {layout_k}, {pair_classes} pair classes")
    s.close()
    alg_bytes = cells * BLOCK * BYTES_FP32
    hbm_alg = alg_bytes / (mean_launch_ms * 1e-3) / 1e9 if mean_launch_ms > 0 else 0.0
    # the bound the kernel works against: LDS bytes of real cells over the
    # launch, against the device's LDS read bandwidth measured here
    lds_bytes = cells * BLOCK * LDS_BYTES_PER_CELL[layout]
    lds_achieved = lds_bytes / (mean_launch_ms * 1e-3) / 1e9 if mean_launch_ms > 0 else 0.0
    lds_peaks = {}
    for w in (4, 8, 16):
        bw = ctypes.c_double(0.0)
        _abi.check(vocalsim.load().vs_measure_lds_bandwidth(w, ctypes.byref(bw)))
        lds_peaks[f"b{8 * w}"] = bw.value / 1e9
    lds_peak = max(lds_peaks.values())
    log(f"[bench] LDS read bandwidth (GB/s): " + ", ".join(f"{k} {v:.0f}" for k, v in lds_peaks.items()))
    flops = cells * BLOCK * FLOPS_PER_CELL / (mean_launch_ms * 1e-3) / 1e12 if mean_launch_ms > 0 else 0.0
    # every work-group steps layout_k slots on each of its threads - 64 cell
    # lanes, the count the batch's largest tube needs; the same LDS traffic
    # counted over those slots is what each CU actually moves
    slots = args.tubes * layout_k * (threads - 64) if layout_k else 0
    slots_info = None
    if slots and mean_launch_ms > 0:
        slot_gbs = slots * BLOCK * LDS_BYTES_PER_CELL[layout] / (mean_launch_ms * 1e-3) / 1e9
        slots_info = {"slots_per_step": slots, "real_cell_fraction": cells / slots,
                      "achieved_gbs_over_slots": slot_gbs, "frac_over_slots": slot_gbs / lds_peak,
                      "note": "frac above counts real cells only; all work-groups step the largest tube's "
                              "slot count, so frac = frac_over_slots x real_cell_fraction"}

    # ---------------- plain FP64 single tube ----------------
    spec64 = tubes.vowel_tube(0, num_steps=33075)
    fp64_rate, c64 = None, None
    if args.fp64_blocks > 0:
        s64 = vocalsim.Solver(spec64, precision=64, threads=1024, steps_per_launch=BLOCK)
        c64 = s64.cells_per_step
        s64.run(BLOCK)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s64.run(args.fp64_blocks * BLOCK)
        torch.cuda.synchronize()
        w64 = time.perf_counter() - t0
        s64.close()
        fp64_rate = c64 * args.fp64_blocks * BLOCK / w64
        log(f"[bench] plain FP64 single tube ({c64} cells): {fp64_rate:.3e} cell-updates/s")

    # ---------------- FP32 precision: F1-F4 vs FP64 (17 tubes, 50 ms) ----------------
    precision = None
    if not args.no_precision:
        from vocalsim import analysis
        pspecs = [tubes.vowel_tube(k, num_steps=33075, source=tubes.SOURCE_GAUSSIAN) for k in range(17)]
        with vocalsim.Solver(pspecs, precision=64, threads=1024) as a64:
            a64.run()
            q64 = [(a64.probe(i), a64.source_trace(i)) for i in range(17)]
        with vocalsim.Solver(pspecs, precision=32, threads=1024, kernel=_abi.K_FAST) as a32:
            a32.run()
            q32 = [(a32.probe(i), a32.source_trace(i)) for i in range(17)]
        fs = 661500.0
        dF, dB, rl2 = [], [], []
        for (y64, u64), (y32, u32) in zip(q64, q32):
            F64, B64 = analysis.formants(y64, u64, fs)
            F32, B32 = analysis.formants(y32, u32, fs)
            dF.append(np.abs(F32 - F64))
            dB.append(np.abs(B32 - B64))
            rl2.append(float(np.linalg.norm(y32 - y64) / np.linalg.norm(y64)))
        dF, dB = np.array(dF), np.array(dB)
        precision = {"workload": "17 seeded 44-section vowel tubes, Gaussian pulse, 50 ms (33075 steps)",
                     "max_abs_dF1_4_hz": float(np.nanmax(dF)), "mean_abs_dF1_4_hz": float(np.nanmean(dF)),
                     "max_abs_dB1_4_hz": float(np.nanmax(dB)), "max_probe_rel_l2": float(max(rl2)),
                     "reference": "FP64 exact GPU kernel (bit-exact to the CPU oracle)"}
        log(f"[bench] FP32 vs FP64: max |dF1-4| = {precision['max_abs_dF1_4_hz']:.4f} Hz, "
            f"max probe rel L2 = {precision['max_probe_rel_l2']:.2e}")

    # ---------------- gradients: log-spectrum loss over a batch of tubes ----------------
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(args, np, torch, vocalsim, tubes, _abi)

    # ---------------- CPU baseline (oracle, one core) ----------------
    cpu = cpu_structured = cpu_threads = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        t = orc.OracleTube(spec64, 64)
        c64 = c64 or (t.H - 2) * (t.W - 2)
        n_full = t.num_steps
        n_cpu, dt, runs = 0, 0.0, 0
        while dt < args.cpu_seconds and runs < 20:
            n = min(n_full, max(200, int(n_full if runs else 400)))
            t.reset()
            t0 = time.perf_counter()
            t.run(n)
            dt += time.perf_counter() - t0
            n_cpu += n
            runs += 1
        cpu = {"value": c64 * n_cpu / dt, "unit": "grid-cell-updates/s", "cores": 1, "kind": "port",
               "host": host_cpu(),
               "sample": f"FP64 oracle (flat arrays, -O2, 1 thread), 44-section vowel tube "
                         f"{t.H}x{t.W}: {n_cpu} steps in {runs} runs of up to {n_full} steps, {dt:.1f} s"}
        log(f"[bench] CPU oracle: {cpu['value']:.3e} cell-updates/s over {n_cpu} steps")
        # the same FP64 engine with each phase's rows over the host's cores
        # (bit-identical), on the CPU share of the box
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        tt = orc.OracleTube(spec64, 64)
        n_t, dt_t, runs_t = 0, 0.0, 0
        while dt_t < args.cpu_seconds / 2 and runs_t < 20:
            tt.reset()
            _, secs = tt.run_threads(n_full, nthr)
            dt_t += secs
            n_t += n_full
            runs_t += 1
        cpu_threads = {"value": c64 * n_t / dt_t, "unit": "grid-cell-updates/s", "cores": nthr, "kind": "port",
                       "sample": f"FP64 oracle (flat arrays, -O2), each phase's rows over {nthr} OpenMP threads, "
                                 f"same tube: {n_t} steps in {runs_t} runs, {dt_t:.1f} s of step loop"}
        log(f"[bench] CPU oracle, {nthr} threads: {cpu_threads['value']:.3e} cell-updates/s over {n_t} steps")
        # the reference engine's own layout and precision: vector<vector<cell>>
        # state, a deep copy per step and the inert wall sweep, FP32, timed
        # over the step loop only (FDTDEngineCPU.cpp:254-292)
        ts = orc.OracleTube(spec64, 32)
        n_s, dt_s, runs_s = 0, 0.0, 0
        while dt_s < args.cpu_seconds / 2 and runs_s < 20:
            n = min(n_full, 400 if runs_s == 0 else n_full)
            ts.reset()
            _, secs = ts.run_structured(n)
            dt_s += secs
            n_s += n
            runs_s += 1
        cpu_structured = {"value": c64 * n_s / dt_s, "unit": "grid-cell-updates/s", "cores": 1,
                          "sample": f"reference-structured FP32 engine (vector<vector<cell>>, per-step copy, "
                                    f"wall sweep, -O2, 1 thread), same tube: {n_s} steps in {runs_s} runs, "
                                    f"{dt_s:.1f} s of step loop"}
        log(f"[bench] CPU reference-structured FP32: {cpu_structured['value']:.3e} cell-updates/s over {n_s} steps")

    out = {
        "metric": "grid-cell-updates/s (batched FP32 vs plain FP64) at <=1 Hz F1-F4 error",
        "value": value,
        "unit": "grid-cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "step_unit": f"one {BLOCK}-time-step block over the whole batch",
        "ms_per_time_step": wall * 1e3 / (args.steps * BLOCK),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batch of {args.tubes} ragged elliptical tubes 13-19 cm, 30-50 sections, "
                               f"half Gaussian / half two-mass, FP32 state, 256-step blocks replayed as "
                               f"hipGraphs of {GRAPH_BLOCKS}",
                   "tubes": args.tubes, "cells_per_step": cells, "block_steps": BLOCK,
                   "graph_blocks": GRAPH_BLOCKS, "kernel": kernel, "threads": threads,
                   "cells_per_lane": layout_k, "pair_classes": pair_classes},
        "fp64_plain": {"value": fp64_rate, "cells_per_step": c64,
                       "workload": "44-section vowel tube, two-mass 800 Pa, FP64, one tube"},
        "ratio_batched_fp32_over_plain_fp64": value / fp64_rate if fp64_rate else None,
        "precision_fp32_vs_fp64": precision,
        "roofline": {"bound": "valu", "achieved": flops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": flops / FP32_PEAK_TFLOPS,
                     "traffic": PMC_TRAFFIC_BYTES if (kernel == _abi.K_FAST and threads == PMC_THREADS
                                                      and layout_k == PMC_LAYOUT_K and args.tubes == 256) else None,
                     "kernel": ("k_fast_wide<K>" if layout == "wide" else "k_fast_step<K>") + " (batched FP32)",
                     "cells_per_lane": layout_k, "mean_launch_ms": mean_launch_ms,
                     "mean_launch_basis": "HIP events around the timed region on the library's stream / launches",
                     "single_launch_ms": single_launch_ms,
                     "basis": f"{FLOPS_PER_CELL} algorithmic flops per cell-update x {cells} cells x {BLOCK} steps "
                              f"per launch; peak = MI355X's published FP32 vector peak (it counts packed FP32 "
                              f"FMAs, which this build does not emit: the unpacked issue limit is half of it)",
                     "pipes_busy": PIPES_BUSY if (kernel == _abi.K_FAST and layout_k == PMC_LAYOUT_K) else None,
                     "lds": {"achieved_gbs": lds_achieved, "peak_gbs": lds_peak,
                             "frac": lds_achieved / lds_peak if lds_peak > 0 else None,
                             "basis": f"{LDS_BYTES_PER_CELL[layout]} B of LDS traffic per real cell and step; peak = "
                                      f"LDS read bandwidth measured on this device (vs_measure_lds_bandwidth, "
                                      f"widest read)",
                             "peak_by_width_gbs": lds_peaks, "slots": slots_info},
                     "hbm": {"algorithmic_gbs_if_streamed": hbm_alg, "peak_gbs": HBM_PEAK_GBS,
                             "compulsory_bytes_per_launch": cells * COMPULSORY_BYTES_PER_CELL,
                             "measured_bytes_per_launch": PMC_TRAFFIC_BYTES if layout_k == PMC_LAYOUT_K else None,
                             "note": "the state stays on chip for the whole launch: HBM carries only the "
                                     "launch's load and write-back"}},
        "cpu_baseline": cpu,
        "cpu_reference_structured": cpu_structured,
        "cpu_oracle_threads": cpu_threads,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
