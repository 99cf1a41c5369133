"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/sanitize_driver.cc runs every oracle entry point (plain and
compensated solves and steps, the f64 twin, accelerations, leaf choice, the
FP64 direct sum and conserved quantities) on one particle file.  It is built
twice, with and without the sanitizers; on edge-case inputs (one and two
bodies, every body at one point, coordinates on cell faces and beyond the
world edge, half-way fixed-point ties, NaN and infinite coordinates, dense
clusters) the sanitized build must run clean and print the same digests as
the plain build.
"""

import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as orc
import rainbow_fmm as rf

DRIVER = Path(orc.ORACLE_DIR) / "sanitize_driver.cc"
SOURCE = Path(orc.ORACLE_DIR) / "rainbow_oracle.cc"
BASE = ["g++", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-g"]
SANITIZE = ["-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    out = tmp_path_factory.mktemp("sanitize")
    plain, sanitized = out / "driver_plain", out / "driver_asan"
    subprocess.run([*BASE, "-O2", "-o", str(plain), str(SOURCE), str(DRIVER)], check=True, capture_output=True)
    subprocess.run([*BASE, *SANITIZE, "-o", str(sanitized), str(SOURCE), str(DRIVER)], check=True,
                   capture_output=True)
    return out, plain, sanitized


def edge_input(kind, n, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 8), dtype=np.float32)
    if kind == "plummer":
        return rf.ic_plummer(n, seed, 0.1, 1.5, 1.0)
    if kind == "ball":
        return rf.ic_uniform_ball(n, seed, 1.0)
    if kind == "point":
        p[:, :3] = np.float32(0.25)
    elif kind == "faces":
        grid = np.array([-1.7, -1.6, -0.8, -0.4, 0.0, 0.2, 0.4, 0.8, 1.2, 1.6, 1.7, 2.5], dtype=np.float32)
        p[:, :3] = rng.choice(grid, size=(n, 3))
        p[:, 4:7] = rng.normal(0, 0.3, size=(n, 3)).astype(np.float32)
    elif kind == "ties":
        scale = orc.plan(n)[2]
        k = rng.integers(-1000, 1000, size=(n, 3))
        p[:, :3] = ((k + 0.5) / scale).astype(np.float32)
    elif kind == "nonfinite":
        p[:, :3] = rng.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32)
        p[0, 0], p[1, 1], p[2, 2], p[3, :3] = np.nan, np.inf, -np.inf, np.nan
        p[4, 4] = np.inf
    elif kind == "clustered":
        centres = rng.uniform(-1.2, 1.2, size=(5, 3))
        idx = rng.integers(0, 5, size=n)
        p[:, :3] = (centres[idx] + rng.normal(0, 0.02, size=(n, 3))).astype(np.float32)
        p[:, 4:7] = rng.normal(0, 0.5, size=(n, 3)).astype(np.float32)
    p[:, 3] = rng.uniform(0, 1, size=n).astype(np.float32)
    p[:, 7] = rng.uniform(0, 1, size=n).astype(np.float32)
    return p


CASES = [("ball", 1), ("ball", 2), ("point", 17), ("clustered", 100), ("faces", 777), ("nonfinite", 999),
         ("ties", 4096), ("plummer", 4096), ("clustered", 16385), ("faces", 32769), ("plummer", 70000)]


@pytest.mark.parametrize("kind,n", CASES, ids=[f"{k}-{n}" for k, n in CASES])
def test_sanitized_oracle_runs_clean_and_agrees(kind, n, drivers):
    out, plain, sanitized = drivers
    path = out / f"{kind}-{n}.bin"
    edge_input(kind, n, 4321 + n).tofile(path)
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    a = subprocess.run([str(sanitized), str(path), "2"], env=env, capture_output=True, text=True, timeout=600)
    assert a.returncode == 0, a.stderr[-4000:]
    assert "runtime error" not in a.stderr and "Sanitizer" not in a.stderr, a.stderr[-4000:]
    b = subprocess.run([str(plain), str(path), "2"], env=env, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-2000:]
    assert a.stdout == b.stdout
    # the digests of the plain build are the oracle library's own
    parts = np.fromfile(path, dtype=np.float32).reshape(n, 8)
    slots = orc.build_plans(n)
    m, f = orc.solve_fields(slots, parts)
    lines = dict(line.split() for line in b.stdout.splitlines())
    assert int(lines["moments"], 16) == orc.fnv1a(m)
    assert int(lines["fields"], 16) == orc.fnv1a(f)
