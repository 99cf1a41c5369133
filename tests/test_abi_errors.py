"""Error paths of the C ABI that are decided before any device work, run on
the CPU: every call returns a negative status with a message from
rainbow_last_error() and nothing aborts (the reference surfaces device
errors to its host, app.js:289-296; here each entry point reports its own).
"""

import ctypes

import numpy as np

import rainbow_fmm as rf


def last_error():
    return rf.lib().rainbow_last_error().decode()


def test_solver_create_rejects_bad_arguments():
    lib = rf.lib()
    h = ctypes.c_void_p()
    assert lib.rainbow_hip_create_ex(0, None, 0, ctypes.byref(h)) == rf.E_INVALID
    assert "particles" in last_error() and h.value is None
    assert lib.rainbow_hip_create_ex(100, None, 0x10, ctypes.byref(h)) == rf.E_INVALID
    assert "option" in last_error()
    assert lib.rainbow_hip_create_ex(100, None, 0, None) == rf.E_INVALID
    assert "null" in last_error()


def test_null_handles_are_rejected():
    lib = rf.lib()
    buf = np.zeros(64, np.float32)
    for rc in (lib.rainbow_hip_step(None, 1), lib.rainbow_hip_solve_fields(None),
               lib.rainbow_hip_upload_particles(None, buf.ctypes.data, buf.nbytes),
               lib.rainbow_hip_download(None, 0, buf.ctypes.data, buf.nbytes),
               lib.rainbow_hip_set_sort_interval(None, 4), lib.rainbow_batch_step(None, 1),
               lib.rainbow_server_poll(None, None)):
        assert rc == rf.E_INVALID
        assert "null" in last_error()


def test_batch_create_rejects_bad_arguments():
    lib = rf.lib()
    h = ctypes.c_void_p()
    counts = (ctypes.c_uint32 * 3)(100, 0, 200)
    assert lib.rainbow_batch_create(3, counts, None, 0, ctypes.byref(h)) == rf.E_INVALID
    assert "system 1 has no bodies" in last_error()
    assert lib.rainbow_batch_create(0, counts, None, 0, ctypes.byref(h)) == rf.E_INVALID
    assert lib.rainbow_batch_create(1, counts, None, 0x8, ctypes.byref(h)) == rf.E_INVALID


def test_server_create_and_submit_reject_bad_arguments():
    lib = rf.lib()
    h = ctypes.c_void_p()
    assert lib.rainbow_server_create(10**9, 0, None, 0, ctypes.byref(h)) == rf.E_INVALID
    assert "max_slots" in last_error()
    assert lib.rainbow_server_create(10**9, 4, None, 0x2, ctypes.byref(h)) == rf.E_INVALID
    r = rf.Request(1, 100, 1, None)
    assert lib.rainbow_server_submit(None, ctypes.byref(r)) == rf.E_INVALID
    assert rf.Server.request_bytes(0) == 0
    # request bytes grow with the body count and cover at least the bodies
    sizes = [rf.Server.request_bytes(n) for n in (1, 16384, 262144, 2097152, 8388608)]
    assert all(a < b for a, b in zip(sizes, sizes[1:]))
    assert sizes[-1] >= 32 * 8388608
