"""vs_dtft_noalloc (the allocation-free variant of vs_dtft) against vs_dtft,
through the C ABI: two calls in a row, the second with a smaller input, each
bit-identical to vs_dtft on the same input (both run the same FP64 kernel on
the same values); sizes past the static array's limits are rejected."""
import ctypes

import numpy as np
import pytest

import vocalsim
from vocalsim import _abi

pytestmark = pytest.mark.gpu

FS = 661500.0
_D = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_D)


def _dtft(fn, x, freqs, hann):
    re = np.zeros(freqs.size)
    im = np.zeros(freqs.size)
    rc = fn(_p(x), x.size, FS, hann, _p(freqs), freqs.size, _p(re), _p(im))
    return rc, re, im


def test_noalloc_dtft_equals_dtft_twice(hip_device):
    L = vocalsim.load()
    rng = np.random.default_rng(11)
    for n, nf, hann in ((33075, 4096, 0), (1200, 37, 1)):
        print(f"[no-alloc dtft] n={n} nf={nf} hann={hann}", flush=True)
        x = np.ascontiguousarray(rng.standard_normal(n))
        freqs = np.ascontiguousarray(np.sort(rng.uniform(0.0, 5000.0, nf)))
        rc0, re0, im0 = _dtft(L.vs_dtft, x, freqs, hann)
        rc1, re1, im1 = _dtft(L.vs_dtft_noalloc, x, freqs, hann)
        assert rc0 == 0 and rc1 == 0, _abi.load().vs_last_error()
        assert np.array_equal(re1, re0) and np.array_equal(im1, im0)
        # and the transform itself: numpy's FFT at the bin frequencies
        if not hann:
            ks = np.array([0, 1, 7, 100])
            got = _dtft(L.vs_dtft_noalloc, x, np.ascontiguousarray(ks * FS / n, dtype=np.float64), 0)
            np.testing.assert_allclose(got[1] + 1j * got[2], np.fft.fft(x)[ks], rtol=1e-10, atol=1e-8)


def test_noalloc_dtft_rejects_oversized_inputs(hip_device):
    L = vocalsim.load()
    x = np.zeros((1 << 18) + 1)
    f = np.zeros(4)
    rc, _, _ = _dtft(L.vs_dtft_noalloc, x, f, 0)
    assert rc == -4 and b"at most" in L.vs_last_error()
    rc, _, _ = _dtft(L.vs_dtft_noalloc, np.zeros(10), np.zeros((1 << 14) + 1), 0)
    assert rc == -4
