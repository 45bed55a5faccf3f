"""Spectral analysis of probe waveforms: transfer function, formants and
bandwidths.

``get_formants_bw`` restates the reference's FFT peak picker
(results/formants_bandwidths.py:5-80: FFT power in dB on a MATLAB-style axis
f = 0:fs/(N-1):fs, local-maximum scan, -3 dB points by linear interpolation,
advance past each peak).

``formants`` is the sub-bin estimator used for the F1-F4 error metric.  The
transfer function H(f) = Y(f) / S(f) (probe over injected source velocity) is
evaluated with an FP64 DTFT on the GPU (vs_dtft) at arbitrary frequencies,
after an exponential window exp(-alpha n) on both signals: that evaluates H
on the circle |z| = e^alpha, which leaves isolated resonance frequencies in
place, adds alpha*fs/pi Hz to every -3 dB bandwidth (subtracted again), and
makes the truncation of a still-ringing response negligible.  Peaks are found
on a 1 Hz grid, refined on a 0.01 Hz grid with a parabolic fit, and their
-3 dB points interpolated on a 0.25 Hz grid.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional, Tuple

import numpy as np

from . import _abi


# ---------------------------------------------------------------------------
# Reference picker (results/formants_bandwidths.py:5-80)
# ---------------------------------------------------------------------------
def get_formants_bw(audio_source, fs: float, num_formants: int) -> np.ndarray:
    """Returns a (3, num_formants) array: bandwidths, centre frequencies and
    peak levels (dB), as the reference's post-processing script does."""
    if isinstance(audio_source, str):
        x = np.loadtxt(audio_source)
    else:
        x = np.asarray(audio_source)   # the caller's dtype (a float32 signal gets a float32 FFT)
    N = x.size
    P_db = 10 * np.log10(np.abs(np.fft.fft(x)) ** 2)
    df = fs / (N - 1)
    f = np.arange(0, fs + df, df)
    out = np.zeros((3, num_formants))
    count, i = 0, 1
    while count < num_formants and i < N - 1:
        if P_db[i] > P_db[i - 1] and P_db[i] > P_db[i + 1]:
            A0 = P_db[i]
            A3 = A0 - 3
            j = i
            while j > 0 and P_db[j] > A3:
                j -= 1
            f_L = np.interp(A3, P_db[j:j + 2], f[j:j + 2])
            k = i
            while k < N - 1 and P_db[k] > A3:
                k += 1
            # np.interp on the descending pair, as the reference calls it
            f_R = np.interp(A3, P_db[k - 1:k + 1], f[k - 1:k + 1])
            out[0, count] = f_R - f_L
            out[1, count] = f[i]
            out[2, count] = A0
            count += 1
            i = k
        else:
            i += 1
    return out


# ---------------------------------------------------------------------------
# FP64 DTFT on the GPU
# ---------------------------------------------------------------------------
def dtft(x: np.ndarray, fs: float, freqs: np.ndarray, hann: bool = False) -> np.ndarray:
    """X(f) = sum_n x[n] exp(-i 2 pi f n / fs) in FP64 on the GPU."""
    L = _abi.load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    freqs = np.ascontiguousarray(freqs, dtype=np.float64)
    re = np.zeros(freqs.size)
    im = np.zeros(freqs.size)
    _abi.check(L.vs_dtft(x.ctypes.data_as(_abi._D), x.size, float(fs), int(hann),
                         freqs.ctypes.data_as(_abi._D), freqs.size,
                         re.ctypes.data_as(_abi._D), im.ctypes.data_as(_abi._D)))
    return re + 1j * im


def transfer_function(probe: np.ndarray, source: np.ndarray, fs: float, freqs: np.ndarray,
                      alpha: float = 0.0, dtft_fn: Optional[Callable] = None) -> np.ndarray:
    """H(f) = DTFT(probe w) / DTFT(source w) with w[n] = exp(-alpha n)."""
    f = dtft if dtft_fn is None else dtft_fn
    w = np.exp(-alpha * np.arange(probe.size))
    return f(probe * w, fs, freqs) / f(source[: probe.size] * w, fs, freqs)


def _peaks(mag_db: np.ndarray, min_prominence_db: float):
    idx = np.where((mag_db[1:-1] > mag_db[:-2]) & (mag_db[1:-1] >= mag_db[2:]))[0] + 1
    keep = []
    for i in idx:
        lo = mag_db[max(0, i - 400):i].min() if i > 0 else mag_db[i]
        hi = mag_db[i + 1:i + 401].min() if i + 1 < mag_db.size else mag_db[i]
        if mag_db[i] - max(lo, hi) >= min_prominence_db:
            keep.append(i)
    return keep


def formants(probe: np.ndarray, source: np.ndarray, fs: float, n: int = 4, fmin: float = 80.0,
             fmax: float = 5000.0, decay: float = 8.0, min_prominence_db: float = 3.0,
             dtft_fn: Optional[Callable] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Formant frequencies and -3 dB bandwidths [Hz] of the first ``n``
    resonances of H = probe / source (see the module docstring)."""
    alpha = decay / probe.size
    broaden = alpha * fs / np.pi
    tf = lambda fr: np.abs(transfer_function(probe, source, fs, fr, alpha, dtft_fn))
    grid = np.arange(fmin, fmax + 1.0, 1.0)
    mag = 20 * np.log10(tf(grid) + 1e-300)
    pk = _peaks(mag, min_prominence_db)[:n]
    F = np.full(n, np.nan)
    B = np.full(n, np.nan)
    for m, i in enumerate(pk):
        fine = np.arange(grid[i] - 1.5, grid[i] + 1.5 + 1e-9, 0.01)
        y = 20 * np.log10(tf(fine))
        j = int(np.clip(np.argmax(y), 1, y.size - 2))
        d = (y[j - 1] - y[j + 1]) / (2 * (y[j - 1] - 2 * y[j] + y[j + 1]))
        F[m] = fine[j] + d * 0.01
        peak_db = y[j] - 0.25 * (y[j - 1] - y[j + 1]) * d
        half = max(broaden * 4 + 150.0, 300.0)
        band = np.arange(F[m] - half, F[m] + half, 0.25)
        yb = 20 * np.log10(tf(band))
        c = int(np.argmin(np.abs(band - F[m])))
        lvl = peak_db - 3.0
        l = c
        while l > 0 and yb[l] > lvl:
            l -= 1
        r = c
        while r < yb.size - 1 and yb[r] > lvl:
            r += 1
        if l == 0 or r == yb.size - 1:
            continue
        fl = np.interp(lvl, [yb[l], yb[l + 1]], [band[l], band[l + 1]])
        fr = np.interp(lvl, [yb[r], yb[r - 1]], [band[r], band[r - 1]])
        B[m] = (fr - fl) - broaden
    return F, B
