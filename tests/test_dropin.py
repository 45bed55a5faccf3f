"""The C++ drop-in API (include/vocalsim/*.h, libvocalsim_dropin.so) driven by
vs_demo, a driver written like the reference's main() (src/main.cu:65-146).

CPU tests cover the host-side error paths; GPU tests compare the probe file
written by FDTDEngineCPU / FDTDEngineEigen::startSolverEngine (6 significant
digits, FDTDEngineCPU.cpp:435-450) with the oracle on the reference's own
area-function files."""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from conftest import AREA_FILES
from vocalsim import tubes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "vocalsim-2p5d-fdtd_amd", "vs_demo")
DATA = AREA_FILES + "/"


def run_demo(args, tmp_path, env_extra=None):
    env = dict(os.environ, VOCALSIM_DATA_DIR=DATA)
    if env_extra:
        env.update(env_extra)
    out = tmp_path / "recorded-pressure.txt"
    cmd = [DEMO] + [str(a) for a in args] + [str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=300)
    return r, out


def test_demo_built():
    assert os.path.exists(DEMO), "vs_demo missing: run __graft_entry__.build()"


def test_regular_tube_flag_selects_no_file(tmp_path):
    # TubeGeometry.cu:58-71: with tube_geometry = 0 no area file is chosen
    r, _ = run_demo([1, 1, 0, 1, 1, 1, 15, 0.02, 1, 1, 0, 64, 100], tmp_path)
    assert r.returncode == 1 and "Unable to open file" in r.stderr


@pytest.mark.gpu
def test_open_space_simulation(hip_device, tmp_path):
    # simulation_type 0 (FDTDSolver.cu:479-550): no area file is read
    r, out = run_demo([1, 0, -1, 1, -1, 1, 15, 0.02, 1, 1, 0, 64, 400], tmp_path,
                      env_extra={"VOCALSIM_DATA_DIR": "/nonexistent/"})
    assert r.returncode == 0, r.stderr
    assert "frame 57 x 355, listener (32, 181)" in r.stdout
    got = np.loadtxt(out)
    ref = orc.OracleTube(tubes.open_space(num_steps=400), 64).run()[0]
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=1e-6 * np.abs(ref).max())


def _oracle_probe(vowel_file, shape, source, steps, variant=0, prec=64):
    spec = tubes.read_area_file(DATA + vowel_file).replace(
        shape=shape, source_type=source, num_steps=steps, pml=1, engine_variant=variant)
    return orc.OracleTube(spec, prec).run()[0]


@pytest.mark.gpu
@pytest.mark.parametrize("vowel,fname,shape,source,engine", [
    (1, "vocaltract_a.txt", 1, 1, 0),
    (2, "vocaltract_i.txt", 2, 1, 1),
    (3, "vocaltract_u.txt", 1, 2, 0),
    (7, "cylindrical-tube.txt", 2, 0, 0),
])
def test_dropin_probe_matches_oracle(hip_device, tmp_path, vowel, fname, shape, source, engine):
    steps = 800
    print(f"[dropin] vowel={vowel} shape={shape} source={source} engine={engine}", flush=True)
    r, out = run_demo([1, 1, 1, vowel, shape, 1, 15, 0.02, source, 1, engine, 64, steps], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "Elapsed time" in r.stdout
    got = np.loadtxt(out)
    ref = _oracle_probe(fname, shape, source, steps, variant=engine)
    assert got.shape == ref.shape
    # the text file keeps 6 significant digits
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=1e-6 * np.abs(ref).max())


@pytest.mark.gpu
def test_dropin_fp32(hip_device, tmp_path):
    r, out = run_demo([1, 1, 1, 1, 2, 1, 15, 0.02, 1, 1, 0, 32, 600], tmp_path)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(out)
    ref = _oracle_probe("vocaltract_a.txt", 2, 1, 600, prec=32)
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)


@pytest.mark.gpu
def test_dropin_float_step_count(hip_device, tmp_path):
    # num_steps from the float-accumulated time vector (SolverParams.cu:59-66)
    r, out = run_demo([1, 1, 1, 1, 1, 1, 15, 0.002, 1, 1, 0, 64, 0], tmp_path)
    assert r.returncode == 0, r.stderr
    assert np.loadtxt(out).size == orc.float_num_steps(0.002)


@pytest.mark.gpu
def test_setup_csv_dumps_match_cell_types(hip_device, tmp_path):
    # FDTDSolver.cu:664-712: setupWaveSolver writes the cell types, beta and
    # sigma'dt of every interior cell and its right / top neighbours into
    # ../../data/ relative to the working directory
    run_dir = tmp_path / "build" / "bin"
    run_dir.mkdir(parents=True)
    (tmp_path / "data").mkdir()
    env = dict(os.environ, VOCALSIM_DATA_DIR=DATA)
    out = tmp_path / "p.txt"
    r = subprocess.run([DEMO] + [str(a) for a in (1, 1, 1, 1, 1, 1, 15, 0.02, 1, 1, 0, 64, 50)] + [str(out)],
                       capture_output=True, text=True, env=env, cwd=run_dir, timeout=300)
    assert r.returncode == 0, r.stderr
    spec = tubes.read_area_file(DATA + "vocaltract_a.txt").replace(num_steps=50)
    ct = orc.OracleTube(spec, 64).field("cell_type").astype(int)
    lines = (tmp_path / "data" / "cell_type_vals.csv").read_text().splitlines()
    assert lines[0] == "current, right, top"
    got = np.array([[int(x) for x in l.split(",")] for l in lines[1:]])
    want = np.stack([ct[1:-1, 1:-1].ravel(), ct[1:-1, 2:].ravel(), ct[:-2, 1:-1].ravel()], axis=1)
    assert np.array_equal(got, want)
    beta = (tmp_path / "data" / "beta_vals.csv").read_text().splitlines()
    assert beta[0] == "cuurent, right, top" and len(beta) == len(lines)
    sig = np.loadtxt(tmp_path / "data" / "sigma_prime_dt_vals.csv", delimiter=",", skiprows=1)
    assert sig.shape == (len(lines) - 1, 3) and np.isclose(sig.max(), 1e6)


# the reference driver's include set (src/main.cu:10-15), by the reference's
# header names, with the drop-in directory on the include path
MAIN_INCLUDES = ["SimulationUnits.h", "SimulationParams.h", "FDTDSolver.h", "FDTDEngineCPU.h",
                 "FDTDEngineEigen.h", "TestTwoMassModel.h"]


def _compile_driver(tmp_path, body):
    src = tmp_path / "driver.cpp"
    src.write_text("".join(f'#include "{h}"\n' for h in MAIN_INCLUDES) + body)
    exe = tmp_path / "driver"
    pkg = os.path.join(ROOT, "vocalsim-2p5d-fdtd_amd")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include", "vocalsim"), str(src), "-o",
                    str(exe), "-L" + pkg, "-lvocalsim_dropin", "-lvocalsim", "-Wl,-rpath," + pkg], check=True)
    return exe


def test_reference_main_include_set_compiles(tmp_path):
    exe = _compile_driver(tmp_path, "int main() { return 0; }\n")
    assert subprocess.run([str(exe)]).returncode == 0


def test_test_two_mass_model_writes_norm_ug(tmp_path):
    # include/TestTwoMassModel.h:9-54: the two-mass model with no supraglottal
    # feedback, normalised by its maximum, "<t ms>\t<norm ug>" per line
    # (2 ms here; the reference's driver runs 1 s into ../../results/norm_ug.txt)
    out = tmp_path / "norm_ug.txt"
    exe = _compile_driver(tmp_path, "#include <cstdlib>\n"
                                    "int main(int c, char** v) { return writeTwoMassNormUg(v[1], 15, 0.05f) > 0 ? 0 : 1; }\n")
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(out)
    n = orc.float_num_steps(0.05)
    assert got.shape == (n, 2)
    ug = orc.twomass_run(np.zeros(n), precision=32)
    dt = np.float32(1.0) / np.float32(661500)
    t = np.zeros(n, np.float32)
    for i in range(1, n):
        t[i] = np.float32(t[i - 1] + dt)
    np.testing.assert_allclose(got[:, 0], t.astype(np.float64) * 1000.0, rtol=1e-5, atol=1e-6)
    want = (ug.astype(np.float32) / np.float32(ug.max())).astype(np.float64)
    # six significant digits of the text stream
    np.testing.assert_allclose(got[:, 1], want, rtol=1e-5, atol=1e-6)
    assert got[:, 1].max() == 1.0 and ug.max() > 0


def test_test_two_mass_model_ends_the_program(tmp_path):
    # the reference's testTwoMassModel() finishes with exit(EXIT_SUCCESS) after
    # writing ../../results/norm_ug.txt relative to the working directory
    run_dir = tmp_path / "build" / "bin"
    run_dir.mkdir(parents=True)
    (tmp_path / "results").mkdir()
    exe = _compile_driver(tmp_path, "int main() { testTwoMassModel(); return 3; }\n")
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=run_dir, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"Completed writing {orc.float_num_steps(1.0)} samples to norm_ug.txt" in r.stdout
    assert (tmp_path / "results" / "norm_ug.txt").exists()
