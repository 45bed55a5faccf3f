"""Host-side audio export helpers (results/generate_audio.py:9-42) against the
library routines the reference script calls: scipy.signal.butter for the
filter coefficients and scipy.io.wavfile.write for the file bytes."""
import numpy as np
import pytest
from scipy.io.wavfile import write as scipy_write
from scipy.signal import butter

import vocalsim
from vocalsim.solver import butter2_lowpass, write_wav


@pytest.mark.parametrize("srate_mul", [2, 4, 8, 15, 30])
def test_butterworth_coefficients_match_scipy(srate_mul):
    fs = 44100.0 * srate_mul
    b, a = butter2_lowpass(5050.0, fs)
    bs, as_ = butter(2, 5050.0 / (fs / 2.0), btype="low")
    np.testing.assert_allclose(b, bs, rtol=4e-15, atol=0)
    np.testing.assert_allclose(a, as_, rtol=4e-15, atol=1e-17)


def test_butterworth_rejects_cutoff_above_nyquist():
    with pytest.raises(vocalsim.VocalSimError):
        butter2_lowpass(30000.0, 44100.0)


@pytest.mark.parametrize("n", [0, 1, 17, 4410])
def test_wav_bytes_match_scipy(tmp_path, n):
    x = np.random.default_rng(n).uniform(-1, 1, n).astype(np.float32)
    scipy_write(tmp_path / "a.wav", 44100, x)
    write_wav(tmp_path / "b.wav", x)
    assert (tmp_path / "a.wav").read_bytes() == (tmp_path / "b.wav").read_bytes()


def test_audio_entry_point_rejects_null_handle():
    L = vocalsim.load()
    assert L.vs_probe_audio(None, 0, None, 0, None) == -1
