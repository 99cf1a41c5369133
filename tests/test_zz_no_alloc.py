"""The allocation-free FP64 direct sum (rainbow_hip_direct_sum_f64_noalloc)
against the allocating one (rainbow_hip_direct_sum_f64), through the C ABI:
both run the same kernels on the same scratch layout, so the results must be
equal bit for bit.  Two calls in a row, the second on a smaller system with
fewer targets (the array keeps the first call's contents); targets listed
twice; more targets than the array holds are refused."""

import ctypes

import numpy as np
import pytest

import rainbow_fmm as rf

pytestmark = pytest.mark.gpu


def direct(fn, solver, targets):
    t = np.ascontiguousarray(targets, dtype=np.uint32)
    out = np.full((t.shape[0], 4), np.nan)
    rc = fn(solver._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), t.shape[0],
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


def test_noalloc_direct_sum_equals_the_allocating_one():
    lib = rf.lib()
    prm = rf.physics(dt=5e-4, gravity=1.0, softening=1e-3)
    for n, nt, seed in ((200000, 4096, 1), (30000, 37, 2)):
        print(f"[case] no-alloc direct sum, n={n}, {nt} targets", flush=True)
        s = rf.Solver(n, prm)
        s.upload(rf.ic_plummer(n, 900 + seed, 0.1, 1.5, 1.0))
        s.step(3)  # bodies are stored in Morton order: targets are caller indices
        rng = np.random.default_rng(seed)
        targets = rng.choice(n, nt, replace=False).astype(np.uint32)
        targets[-1] = targets[0]  # a target listed twice
        rc, got = direct(lib.rainbow_hip_direct_sum_f64_noalloc, s, targets)
        assert rc == 0, lib.rainbow_last_error().decode()
        rc, want = direct(lib.rainbow_hip_direct_sum_f64, s, targets)
        assert rc == 0, lib.rainbow_last_error().decode()
        assert np.isfinite(got).all()
        assert got.tobytes() == want.tobytes()
        assert got[-1].tobytes() == got[0].tobytes()
        rc, _ = direct(lib.rainbow_hip_direct_sum_f64_noalloc, s, np.arange(4097))
        assert rc == rf.E_INVALID and "at most 4096" in lib.rainbow_last_error().decode()
        s.close()
