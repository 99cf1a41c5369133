"""The browser's per-frame time step (app.js:42-43, 711-722).

CPU: rainbow_frame_dt against the values the frame formula fixes -- one
60 Hz frame gives BASE_DT = 0.0009, the interval scales it linearly and is
capped at 100 ms (six frames).  GPU: a frame step equals a plain step at
that dt, and the oracle's step, bit for bit.
"""

import sys

import numpy as np
import pytest

import rainbow_fmm as rf


def test_frame_dt_values():
    f32 = np.float32
    assert rf.frame_dt(1000.0 / 60.0) == f32(0.0009)
    assert rf.frame_dt(2000.0 / 60.0) == f32(0.0018)
    assert rf.frame_dt(50.0) == f32(0.0027)
    # The interval is capped at 100 ms: six 60 Hz frames.
    assert rf.frame_dt(100.0) == f32(0.0054)
    assert rf.frame_dt(250.0) == f32(0.0054)
    assert rf.frame_dt(1e9) == f32(0.0054)
    assert rf.frame_dt(0.0) == 0.0
    xs = np.linspace(0.0, 120.0, 241)
    ys = np.array([rf.frame_dt(x) for x in xs])
    assert np.all(np.diff(ys) >= 0)


@pytest.mark.gpu
def test_gpu_step_frame_matches_plain_step_and_oracle():
    import oracle as orc
    from test_gpu_parity import assert_bits, slots_for

    n = 65536
    print(f"[case] step_frame at 30 Hz, n={n}", flush=True)
    sys.stdout.flush()
    parts = rf.ic_plummer(n, 81, 0.1, 1.5, 1.0)
    g, soft = 1.0, 1e-3
    dt = rf.frame_dt(1000.0 / 30.0)
    s = rf.Solver(n, rf.physics(dt=0.0009, gravity=g, softening=soft))
    s.upload(parts)
    s.step_frame(1000.0 / 30.0)
    s.step_frame(1000.0 / 30.0)
    got = s.particles()
    s.close()
    t = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft))
    t.upload(parts)
    t.step(2)
    assert_bits(got, t.particles(), "frame steps vs plain steps")
    t.close()
    op, _, _ = orc.step(slots_for(n, dt, g, soft), parts, 2)
    assert_bits(got, op, "frame steps vs oracle")
