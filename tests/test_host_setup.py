"""The library's host-only set-up code (vocalsim-2p5d-fdtd_amd/csrc/vs_host.cpp)
on the CPU: the float32 prelude of every case, and of 200 seeded random
specs, against the oracle's frame dimensions and per-section values, the
single-plane contours against the
oracle's, the cell-class enumeration of both step forms, and the refusal of
bad specs.  The same checks run again on a build with AddressSanitizer and
UndefinedBehaviorSanitizer, which must finish with no report; a planted
one-past-the-end read in the harness shows that build reports such errors.
"""
import os
import subprocess
import sys

import pytest

import host_setup_checks as hsc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def host_lib(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("host") / "libhost_setup.so")
    return hsc.load(hsc.build(path))


def test_prelude_matches_oracle(host_lib):
    assert hsc.check_prelude(host_lib) == 2 * len(hsc.cases())


def test_prelude_matches_oracle_on_random_specs(host_lib):
    assert hsc.check_prelude_random(host_lib) == 400


def test_prelude_refuses_what_the_oracle_refuses(host_lib):
    hsc.check_bad_specs(host_lib)


def test_plane_contours_match_oracle(host_lib):
    hsc.check_plane_contours(host_lib, areas=(0.4, 1.0, 2.84, 5.2, 7.5))


def test_exact_classes(host_lib):
    hsc.check_exact_classes(host_lib)


def test_fast_classes(host_lib):
    hsc.check_fast_classes(host_lib)


def test_fast_pair_classes(host_lib):
    hsc.check_fast_pairs(host_lib)


def _sanitizer_libs():
    libs = []
    for name in ("libasan.so", "libubsan.so"):
        p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
        if not p or not os.path.isabs(p) or not os.path.exists(p):
            pytest.skip(f"{name} not available")
        libs.append(p)
    return libs


def _run_sanitized(lib, *args):
    env = dict(os.environ, LD_PRELOAD=":".join(_sanitizer_libs()),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    return subprocess.run([sys.executable, os.path.join(ROOT, "tests", "host_setup_checks.py"), str(lib), *args],
                          env=env, capture_output=True, text=True, timeout=900)


@pytest.fixture(scope="module")
def sanitized_lib(tmp_path_factory):
    return hsc.build(str(tmp_path_factory.mktemp("host_san") / "libhost_setup_san.so"), sanitize=True)


def test_host_setup_under_address_and_ub_sanitizers(sanitized_lib):
    r = _run_sanitized(sanitized_lib)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "host set-up checks ok" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr


def test_sanitized_build_reports_a_planted_overflow(sanitized_lib):
    r = _run_sanitized(sanitized_lib, "overflow")
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr
