"""FP32 precision against FP64 on the GPU: probe relative L2 and F1-F4 formant
and bandwidth errors (Hz) over the 17-tube, 50 ms Gaussian-pulse workload.

The FP64 reference is the exact FP64 kernel, which the parity tests hold bit
for bit to the CPU oracle (and which test_fp64_full_run_matches_oracle checks
again here over a full 33 075-step run).  Bar: |dF1..4| <= 1 Hz per tube."""
import json
import os

import numpy as np
import pytest

import oracle as orc
from vocalsim import Solver, analysis, tubes
from vocalsim import _abi

pytestmark = pytest.mark.gpu
FS = 661500.0
N_STEPS = 33075


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def test_gpu_dtft_matches_numpy_fft(hip_device):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(33075)
    X = np.fft.fft(x)
    ks = np.array([0, 1, 37, 500, 2501, 16000])
    got = analysis.dtft(x, FS, ks * FS / x.size)
    np.testing.assert_allclose(got, X[ks], rtol=1e-11, atol=1e-9)
    w = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(x.size) / (x.size - 1))
    got_h = analysis.dtft(x, FS, ks * FS / x.size, hann=True)
    np.testing.assert_allclose(got_h, np.fft.fft(x * w)[ks], rtol=1e-11, atol=1e-9)


def precision_specs():
    return [tubes.vowel_tube(s, num_steps=N_STEPS, source=tubes.SOURCE_GAUSSIAN) for s in range(17)]


@pytest.fixture(scope="module")
def runs(hip_device):
    specs = precision_specs()
    print("[precision] 17 tubes x 33075 steps: FP64 exact and FP32 fast", flush=True)
    with Solver(specs, precision=64, threads=1024) as s64:
        s64.run()
        p64 = [s64.probe(i) for i in range(len(specs))]
        src64 = [s64.source_trace(i) for i in range(len(specs))]
    with Solver(specs, precision=32, threads=1024, kernel=_abi.K_FAST) as s32:
        s32.run()
        p32 = [s32.probe(i) for i in range(len(specs))]
        src32 = [s32.source_trace(i) for i in range(len(specs))]
    return specs, p64, src64, p32, src32


def test_fp64_full_run_matches_oracle(runs):
    specs, p64, _, _, _ = runs
    ref, _ = orc.OracleTube(specs[0], 64).run()
    assert rel_l2(p64[0], ref) <= 1e-10


def test_fp32_formant_error_below_1hz(runs):
    specs, p64, src64, p32, src32 = runs
    rows = []
    for i in range(len(specs)):
        F64, B64 = analysis.formants(p64[i], src64[i], FS)
        F32, B32 = analysis.formants(p32[i], src32[i], FS)
        rows.append({"tube": i, "rel_l2": rel_l2(p32[i], p64[i]),
                     "F64": F64.tolist(), "dF": (F32 - F64).tolist(),
                     "B64": B64.tolist(), "dB": (B32 - B64).tolist()})
        print(f"  tube {i:2d} relL2 {rows[-1]['rel_l2']:.2e} F {np.round(F64, 1)} "
              f"dF {np.round(F32 - F64, 4)} dB {np.round(B32 - B64, 3)}", flush=True)
    out_dir = os.environ.get("VOCALSIM_RESULTS_DIR")   # optional per-tube table
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, "precision_f1f4.json"), "w") as f:
            json.dump(rows, f, indent=1)
    dF = np.array([r["dF"] for r in rows])
    assert np.all(np.isfinite(dF))
    assert np.max(np.abs(dF)) <= 1.0
    assert max(r["rel_l2"] for r in rows) <= 1e-3
