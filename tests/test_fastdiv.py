"""The integrate kernels form p / 1.6f (the position -> unit-cube map of
position_cell, src/rainbow_fmm.wgsl:144-152) and the two clamped ramps
(count - 16) / 72 (:703-705) and smoothstep(0.30, 0.70, excitation)'s
(excitation - 0.30) / (0.70 - 0.30) (:758) with a
remainder-corrected multiplication instead of a division.  This compiles the
host checker and runs it over all 2^32 inputs of each: the value the kernel
keeps must equal the one from the correctly rounded quotient for every input.  The same checker shows
that the form is NOT exact for other constants (a plausible misuse), so the
test also pins that it detects a difference."""

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "fastdiv"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    exe = tmp_path_factory.mktemp("fastdiv") / "fastdiv_check"
    subprocess.run(
        ["g++", "-O2", "-fopenmp", "-ffp-contract=off", "-o", str(exe), str(HERE / "fastdiv_check.cc")],
        check=True,
    )
    return exe


def test_unit_coord_division_exact_for_all_floats(checker):
    r = subprocess.run([str(checker), "1.6", "unit"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "kept values differing 0" in r.stdout


def test_position_map_exact_for_every_bit_pattern(checker):
    # The kernel's full form, dividend clamped to [-2, 2] first, over all
    # 2^32 bit patterns: +inf must land where the division puts it (the last
    # cell), -inf and NaN in the first.
    r = subprocess.run([str(checker), "1.6", "position"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "kept values differing 0" in r.stdout


def test_integrate_ramps_exact_for_all_inputs(checker):
    # The excitation ramp over every float excitation, the density ramp over
    # every uint32 body count (divisors as the kernel forms them).
    for d, mode in (("0.39999998", "excitation"), ("72", "density")):
        r = subprocess.run([str(checker), d, mode], capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "kept values differing 0" in r.stdout


def test_checker_detects_inexact_constant(checker):
    # x / 72 clamped to [0, 1] is not reproduced for every float by the
    # corrected product: the checker must report it.
    r = subprocess.run([str(checker), "72", "clamp01"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 1, r.stdout + r.stderr
