"""Spectral analysis: the sub-bin formant estimator on signals whose formants
and bandwidths are known in closed form, and the restated reference picker.
CPU tests drive the estimator with a numpy DTFT; the GPU DTFT is checked
against numpy's FFT in tests/test_gpu_analysis.py."""
import numpy as np
import pytest

from vocalsim import analysis

FS = 22050.0   # low rate keeps the numpy DTFT of these tests fast


def np_dtft(x, fs, freqs):
    n = np.arange(x.size)
    out = np.empty(freqs.size, dtype=complex)
    for i in range(0, freqs.size, 512):
        f = freqs[i:i + 512]
        out[i:i + 512] = np.exp(-2j * np.pi * np.outer(f, n) / fs) @ x
    return out


def resonator_bank(F, B, n, fs=FS):
    """Impulse response of parallel resonators: mode k decays as
    exp(-pi B_k t) (-3 dB bandwidth B_k) and rings at F_k."""
    t = np.arange(n) / fs
    y = np.zeros(n)
    for f, b in zip(F, B):
        y += np.exp(-np.pi * b * t) * np.cos(2 * np.pi * f * t)
    src = np.zeros(n)
    src[0] = 1.0
    return y, src


def test_formants_of_known_resonators():
    F = np.array([520.0, 1480.0, 2510.0, 3490.0])
    B = np.array([40.0, 60.0, 80.0, 100.0])
    y, s = resonator_bank(F, B, int(0.05 * FS))
    Fe, Be = analysis.formants(y, s, FS, n=4, dtft_fn=np_dtft)
    # the tails of neighbouring modes tilt each peak, which moves |H| maxima by
    # a fraction of a percent from the pole frequencies (the F1-F4 metric
    # compares two runs through the same estimator, where this bias cancels)
    np.testing.assert_allclose(Fe, F, rtol=5e-3)
    np.testing.assert_allclose(Be, B, rtol=0.25)


def test_formant_estimate_is_stable_to_small_perturbations():
    F = np.array([610.0, 1720.0, 2650.0, 3800.0])
    B = np.array([50.0, 70.0, 90.0, 120.0])
    y, s = resonator_bank(F, B, int(0.05 * FS))
    rng = np.random.default_rng(1)
    yp = y * (1 + 1e-5 * rng.standard_normal(y.size))
    F0, B0 = analysis.formants(y, s, FS, dtft_fn=np_dtft)
    F1, B1 = analysis.formants(yp, s, FS, dtft_fn=np_dtft)
    assert np.max(np.abs(F1 - F0)) < 0.05
    assert np.max(np.abs(B1 - B0)) < 0.5


def test_reference_picker_on_a_single_resonance():
    fs = 44100.0
    y, _ = resonator_bank([1000.0], [80.0], 4096, fs=fs)
    out = analysis.get_formants_bw(y, fs, 1)
    df = fs / (y.size - 1)
    assert abs(out[1, 0] - 1000.0) <= df
    assert out[0, 0] == pytest.approx(80.0, rel=0.35)


def test_get_formants_bw_matches_reference_golden_vectors():
    # outputs of the reference's own get_formants_bw (results/formants_bandwidths.py:5-80),
    # written by tests/golden/formants/make_formants_golden.py: damped-resonator
    # signals at 44.1 and 661.5 kHz, an odd length, 1 to 10 formants, float64
    # and float32 input; the restatement reproduces them exactly
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "formants", "formants_ref.npz"))
    n = 0
    for key in g.files:
        name, kind, *rest = key.split("__")
        if kind not in ("f64", "f32"):
            continue
        x = g[f"{name}__x"]
        fs = float(g[f"{name}__fs"])
        got = analysis.get_formants_bw(x.astype(np.float64) if kind == "f64" else x, fs, int(rest[0]))
        assert np.array_equal(got, g[key]), key
        n += 1
    assert n == 10
