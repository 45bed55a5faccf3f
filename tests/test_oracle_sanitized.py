"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer,
driven through its index-heavy parts: wall rows and the gap fill, the dead
and PML rings, circle / ellipse / single-plane-symmetric depth maps (contour
growth and vertical chords included), open space, 2D mode, the coefficient
triples, the two-mass recurrence, and the flat, threaded and
reference-structured step engines at both precisions.  The run must finish
with no sanitizer report.  A copy with a planted one-past-the-end read
(VSO_MUTATION=4) shows that the instrumented build does report such errors.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

WORKLOAD = r"""
import sys
sys.path[:0] = {paths!r}
import numpy as np
import oracle as orc
from vocalsim import tubes
specs = [tubes.uniform_tube(num_steps=50), tubes.vowel_tube(0, num_steps=50),
         tubes.vowel_tube(3, num_steps=50, source=tubes.SOURCE_GAUSSIAN).replace(pml=0),
         tubes.uniform_tube(num_steps=50).replace(mouth_dirichlet=0),
         tubes.uniform_tube(num_steps=50).replace(solver_type=0),
         tubes.vowel_tube(1, num_steps=40, source=tubes.SOURCE_SINE).replace(engine_variant=1),
         tubes.open_space(num_steps=30), tubes.open_space(num_steps=30, pml=0, solver_type=0, source_type=2),
         tubes.vowel_tube(5, num_steps=30, source=tubes.SOURCE_GAUSSIAN, n_sections=6).replace(shape=3, aspect=None)]
specs += tubes.batch_tubes(6, seed=2, num_steps=40)
for sp in specs:
    for prec in (32, 64):
        t = orc.OracleTube(sp, prec)
        t.run()
        for f in ("cell_type", "boundary_seg", "depth_p", "depth_x", "depth_y", "z_inv", "upper_wall",
                  "lower_wall", "segment", "source", "N_out", "xor_val4"):
            t.field(f)
        P, Vx, Vy = t.state()
        t.set_state(P, Vx, Vy)
        t.run_structured(10)
        t.run_threads(10, 2)
for cur in range(13):
    for right in range(13):
        orc.coeff_triple(cur, right, (cur + right) % 13)
orc.twomass_run(np.linspace(0, 50, 3000), precision=32)
orc.twomass_run(np.linspace(0, 50, 3000), precision=64, q=0.7, ps=800.0)
for a in (0.4, 2.84, 7.5):
    orc.plane_contour(a)
print("sanitized oracle workload ok", len(specs))
"""


def _sanitizer_libs():
    libs = []
    for name in ("libasan.so", "libubsan.so"):
        p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
        if not p or not os.path.isabs(p) or not os.path.exists(p):
            pytest.skip(f"{name} not available")
        libs.append(p)
    return libs


def _build(tmp_path, mutation=0):
    lib = tmp_path / f"liboracle_san{mutation}.so"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp",
                    "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined",
                    f"-DVSO_MUTATION={mutation}", "-shared", "-o", str(lib),
                    os.path.join(ROOT, "oracle", "vs_oracle.cpp")], check=True)
    return lib


def _run(lib, code):
    env = dict(os.environ, VSO_ORACLE_LIB=str(lib), LD_PRELOAD=":".join(_sanitizer_libs()),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               OMP_NUM_THREADS="2")
    return subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=900)


def _code():
    return WORKLOAD.format(paths=[os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"),
                                  os.path.join(ROOT, "vocalsim-2p5d-fdtd_amd")])


def test_oracle_under_address_and_ub_sanitizers(tmp_path):
    r = _run(_build(tmp_path), _code())
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sanitized oracle workload ok" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr


def test_sanitizer_build_reports_a_planted_overflow(tmp_path):
    r = _run(_build(tmp_path, mutation=4), _code())
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr
