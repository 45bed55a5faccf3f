"""Audio export on the GPU (vs_probe_audio) against the reference script's own
pipeline (results/generate_audio.py:9-42) run with scipy on the same probe:
normalise by max |p|, clip, scipy.signal.butter(2, 5050 / (fs / 2)) +
scipy.signal.lfilter, decimate by srate_mul, float32.  Bar: equal length and
max |diff| <= 2e-7 (float32 output of FP64 filtering)."""
import numpy as np
import pytest
from scipy.signal import butter, lfilter

from vocalsim import Solver, tubes

pytestmark = pytest.mark.gpu


def reference_audio(p, srate_mul):
    audio = np.clip(p / np.max(np.abs(p)), -1.0, 1.0)
    if 44100 * srate_mul != 44100:
        b, a = butter(2, 5050.0 / (44100.0 * srate_mul / 2.0), btype="low")
        audio = lfilter(b, a, audio)[::srate_mul]
    return audio.astype(np.float32)


@pytest.mark.parametrize("srate_mul,steps", [(15, 3000), (4, 1500), (1, 400)])
def test_audio_matches_reference_pipeline(hip_device, srate_mul, steps):
    spec = tubes.vowel_tube(2, num_steps=steps, source=tubes.SOURCE_GAUSSIAN).replace(srate_mul=srate_mul)
    print(f"[audio] srate_mul={srate_mul} steps={steps}", flush=True)
    with Solver(spec, precision=64) as s:
        s.run()
        got = s.audio()
        want = reference_audio(s.probe(), srate_mul)
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) <= 2e-7


def test_batch_audio_is_the_concatenation(hip_device, tmp_path):
    specs = tubes.batch_tubes(3, seed=2, num_steps=900)
    with Solver(specs, precision=32, threads=256) as s:
        s.run()
        allv = s.audio(-1)
        parts = [s.audio(i) for i in range(3)]
        s.write_wav(str(tmp_path / "t0.wav"), 0)
    assert np.array_equal(allv, np.concatenate(parts))
    from scipy.io.wavfile import read
    rate, data = read(tmp_path / "t0.wav")
    assert rate == 44100 and np.array_equal(data, parts[0])
