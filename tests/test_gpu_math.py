"""The kernels' correctly rounded sqrt and 1/sqrt forms (Newton steps from the
hardware estimates, csrc/rainbow_math.h) checked on the device against the
compiler's sqrtf and 1.0f / sqrtf over all 2^32 float inputs, and the
one-instruction saturating f32 -> i32 conversion against its reference form
(truncation, i32 range clamp, NaN to 0) over the same inputs
(bin/math_check)."""

import subprocess
from pathlib import Path

import pytest

EXE = Path(__file__).resolve().parent.parent / "fmm-rainbow_amd" / "bin" / "math_check"


@pytest.mark.gpu
def test_root_forms_exhaustive():
    assert EXE.exists(), "bin/math_check is not built (run __graft_entry__.build())"
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "inputs 4294967296" in r.stdout
    assert "i32 conversion, all inputs: 0 mismatches" in r.stdout
