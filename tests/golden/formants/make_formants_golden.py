"""Writes golden vectors of the reference's formant/bandwidth picker.

get_formants_bw (results/formants_bandwidths.py:5-80) is the reference's own
post-processing routine; its exact behaviour (MATLAB-style frequency axis
df = fs / (N - 1), the i = k advance, np.interp on the descending right-hand
pair) is pinned only by running it, so this script imports it from the
reference checkout (numpy is its only dependency; the module's command-line
code is under a __main__ guard) and records its outputs.

Inputs are sums of exponentially damped cosines (formants F with bandwidths
B), stored as float32 samples:
  vowel44k    N = 4096 at 44.1 kHz, F = 500/1500/2500/3500 Hz, B = 60..150 Hz
  vowel_noisy the same plus seeded noise at -40 dB
  probe_like  N = 13230 at 661.5 kHz (a 20 ms probe at srate_mul 15),
              F = 620/1180/2410/3300 Hz
  odd_length  N = 5001 at 44.1 kHz (the float arange axis has N + 1 points)
Each is evaluated as float64 (how probes are passed) and, for vowel44k, also
as float32 (the reference keeps the caller's dtype, and numpy's FFT then
runs in single precision).
Output: formants_ref.npz next to this script.
Run from the repository root: python tests/golden/formants/make_formants_golden.py
"""
import importlib.util
import os
import sys

import numpy as np

sys.dont_write_bytecode = True   # importing the reference must not write a __pycache__ beside it

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/results/formants_bandwidths.py"


def resonators(F, B, n, fs, amp=None):
    t = np.arange(n) / fs
    amp = np.ones(len(F)) if amp is None else amp
    y = np.zeros(n)
    for f, b, a in zip(F, B, amp):
        y += a * np.exp(-np.pi * b * t) * np.cos(2 * np.pi * f * t)
    return y


def cases():
    v = resonators([500, 1500, 2500, 3500], [60, 90, 120, 150], 4096, 44100.0, [1.0, 0.6, 0.4, 0.3])
    noisy = v + 1e-2 * np.random.default_rng(3).standard_normal(v.size)
    probe = resonators([620, 1180, 2410, 3300], [70, 85, 110, 140], 13230, 661500.0, [1.0, 0.8, 0.5, 0.35])
    odd = resonators([700, 1200, 2600], [50, 80, 100], 5001, 44100.0)
    return {
        "vowel44k": (v.astype(np.float32), 44100.0, (1, 4, 10)),
        "vowel_noisy": (noisy.astype(np.float32), 44100.0, (4, 10)),
        "probe_like": (probe.astype(np.float32), 661500.0, (4, 10)),
        "odd_length": (odd.astype(np.float32), 44100.0, (3, 6)),
    }


def main():
    spec = importlib.util.spec_from_file_location("ref_formants", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = {}
    for name, (x, fs, nfs) in cases().items():
        out[f"{name}__x"] = x
        out[f"{name}__fs"] = np.array(fs)
        for nf in nfs:
            out[f"{name}__f64__{nf}"] = mod.get_formants_bw(x.astype(np.float64), fs, nf)
        if name == "vowel44k":
            out[f"{name}__f32__4"] = mod.get_formants_bw(x, fs, 4)
    np.savez_compressed(os.path.join(HERE, "formants_ref.npz"), **out)


if __name__ == "__main__":
    main()
