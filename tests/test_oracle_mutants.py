"""The closed-form step tests catch the mistakes they are written against.

Builds the oracle with each planted mistake (ORACLE_PLANT in
oracle/onco_oracle_real.c) and runs the test aimed at it on that build, in a
child process that loads it through ONCO_ORACLE_PLANTED_LIB; the test must
fail. The same tests pass on the oracle itself (test_oracle_step_pins.py).
CPU only; needs gcc.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
PINS = os.path.join(ROOT, "tests", "test_oracle_step_pins.py")
FLAGS = ["-O1", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-w"]

PLANTS = {
    1: "test_uptake_removes_k_times_uptake_per_occupied_voxel",
    2: "test_uptake_precedes_the_sweep",
    3: "test_displacement_capped_at_max_disp_along_the_pair_axis",
    4: "test_cells_pushed_past_a_face_land_on_it",
    5: "test_drug_kill_is_strict_first_word_of_the_cell_draw",
    6: "test_drug_kill_is_strict_first_word_of_the_cell_draw",
    7: "test_drug_kill_is_strict_first_word_of_the_cell_draw",
}


def _cc(args):
    subprocess.run(["gcc", *FLAGS, *args], check=True, capture_output=True)


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    d = tmp_path_factory.mktemp("planted")
    int_o = str(d / "int.o")
    _cc(["-c", os.path.join(ORACLE, "onco_oracle_int.c"), "-o", int_o])
    real = os.path.join(ORACLE, "onco_oracle_real.c")

    def build(p):
        f32 = str(d / f"f32_{p}.o")
        f64 = str(d / f"f64_{p}.o")
        _cc([f"-DORACLE_PLANT={p}", "-DREAL=float", "-DSFX=f32", "-DORACLE_F32", "-c", real, "-o", f32])
        _cc([f"-DORACLE_PLANT={p}", "-DREAL=double", "-DSFX=f64", "-c", real, "-o", f64])
        so = str(d / f"liboracle_plant{p}.so")
        subprocess.run(["gcc", "-shared", "-fopenmp", "-o", so, int_o, f32, f64, "-lm"], check=True,
                       capture_output=True)
        return p, so

    with ThreadPoolExecutor(4) as ex:
        return dict(ex.map(build, list(PLANTS) + [0]))


def _run_pin(so, test=None):
    env = dict(os.environ, ONCO_ORACLE_PLANTED_LIB=so)
    target = PINS if test is None else f"{PINS}::{test}"
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", target], cwd=ROOT,
                          env=env, capture_output=True, text=True, timeout=300)


def test_unplanted_build_passes(planted):
    """The same build recipe with nothing planted passes every step test."""
    r = _run_pin(planted[0])
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_is_caught(planted, plant):
    r = _run_pin(planted[plant], PLANTS[plant])
    assert r.returncode == 1 and "1 failed" in r.stdout, r.stdout[-2000:]
