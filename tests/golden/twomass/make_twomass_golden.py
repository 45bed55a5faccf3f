"""Writes the two-mass golden vectors from the reference's own compiled model.

Builds oracle/_ref/libref_twomass.so (`make -C oracle ref`: the reference's
src/TwoMassModel.cpp compiled with g++ where it lies) and records its volume
velocity for two supraglottal-pressure feeds, each 16 000 steps at
srate_mul = 15 with the reference's own parameters (ps = 686.46 Pa, q = 1).
16 000 steps (24 ms) take the folds through their first closure (near step
9 200, where the flow is exactly 0) and reopening, so both branches of every
condition of the model are exercised.
  zero   p1 = 0
  noise  p1 = 200 * standard normal, seed 7 (float32)
Output: twomass_ref_fp32.npz next to this script (p1_noise, ug_zero, ug_noise).
Run from the repository root: python tests/golden/twomass/make_twomass_golden.py
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", "..", ".."))
N = 16000


def feeds():
    return {
        "zero": np.zeros(N, np.float32),
        "noise": (200 * np.random.default_rng(7).standard_normal(N)).astype(np.float32),
    }


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_twomass.so"))
    out = {}
    for name, p1 in feeds().items():
        ug = np.zeros(N, np.float32)
        rc = lib.ref_twomass_run(15, p1.ctypes.data_as(ctypes.c_void_p), N, ug.ctypes.data_as(ctypes.c_void_p))
        assert rc == 0
        if name != "zero":
            out["p1_" + name] = p1
        out["ug_" + name] = ug
    np.savez_compressed(os.path.join(HERE, "twomass_ref_fp32.npz"), **out)


if __name__ == "__main__":
    main()
