"""GPU parity at the benchmark configurations that are not the 1M step, and
the failure paths of the C ABI.

* cfg3, the precision form at the budget maximum: 5,242,880 bodies in 64
  Plummer clumps (a = 0.02, cut at 10 a, seed 3), compensated sums: the
  field solve and two steps bit for bit against the oracle, and the FP64
  diagnostics against the oracle's conserved quantities and potential.
* cfg4, the batched form: 256 Plummer systems with N log-uniform in
  [4096, 262144] (seed 4000 + i), three steps in one batch: every system bit
  for bit against its own solo solver, and 16 of them (the smallest, the
  largest and 14 spread over the list) against the oracle.
* cfg5, the serving form: Poisson arrivals of Plummer or uniform-ball
  requests with N log-uniform in [16384, 2097152] (one request at
  2,097,152) and 1..200 steps under a 200 GB budget: every result bit for
  bit against a solo solver run of the same request, the smallest ones
  against the oracle.
* Failure paths: requests over the budget, bad sizes and handles, bodies
  with NaN and infinite coordinates (bit for bit, NaNs compared as NaN), the
  harness CLI's finiteness line, and a fixed-point scale changed between
  steps.
* One system of 150,000,000 bodies (4.8 GB of particles, every body
  buffer past 4 GiB): two steps bit for bit against the oracle.
* The server's request life cycle: cancelled (queued and running) and
  expired requests, a zero-step request, duplicate ids, and a stream of
  arrivals through room for a few systems at a time (arena reuse, systems
  joining and leaving between steps), every result bit for bit.
"""

import re
import subprocess
import sys
import time
from pathlib import Path

import numpy as np
import pytest

import oracle as orc
import rainbow_fmm as rf

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def canon(a):
    """Bit view with every NaN mapped to one pattern (CPU and GPU NaN
    payloads differ; a NaN must still meet a NaN)."""
    a = np.ascontiguousarray(a)
    if a.dtype == np.float32:
        return np.where(np.isnan(a), np.float32(np.nan), a).view(np.uint32)
    return a


def assert_bits(got, want, what):
    g, w = canon(got), canon(want)
    assert g.shape == w.shape, f"{what}: shape {g.shape} != {w.shape}"
    if g.tobytes() == w.tobytes():
        return
    bad = np.argwhere(g != w)
    first = tuple(bad[0])
    raise AssertionError(f"{what}: {len(bad)} of {g.size} words differ; first at {first}: "
                         f"got {np.asarray(got)[first]!r} want {np.asarray(want)[first]!r}")


def slots_for(n, dt, gravity, softening, **kw):
    return orc.build_plans(n, dt=dt, gravity=gravity, softening=softening, **kw)


CFG = dict(dt=5e-4, gravity=1.0, softening=1e-3)


def test_cfg3_clumps_compensated_bit_exact():
    n = 5242880
    say(f"cfg3: {n} bodies in 64 clumps, compensated solve + 2 steps")
    parts = rf.ic_clumps(n, 3, 64, 0.02, 10.0, 1.0)
    slots = slots_for(n, **CFG)
    s = rf.Solver(n, rf.physics(**CFG), compensated=True)
    assert s.depth == 6 and s.scale == 128.0
    s.upload(parts)
    s.solve_fields()
    om, of = orc.solve_fields(slots, parts, compensated=True)
    assert_bits(s.moments(), om, "moments")
    assert_bits(s.fields(), of, "fields (compensated)")
    d = s.diagnostics()
    cons = orc.conserved_f64(parts)
    acc = orc.accelerations(slots, parts, om, of, compensated=True)
    pe = 0.5 * np.sum(acc[:, 3].astype(np.float64)) / n
    assert abs(d[0] - cons[0]) <= 1e-12 * abs(cons[0])
    assert abs(d[1] - pe) <= 1e-11 * abs(pe)
    np.testing.assert_allclose(d[2:], cons[2:], rtol=0, atol=1e-13)
    s.upload(parts)
    s.step(2)
    op, _, of2 = orc.step(slots, parts, 2, compensated=True)
    assert_bits(s.particles(), op, "particles after 2 compensated steps")
    assert_bits(s.fields(), of2, "fields after 2 compensated steps")
    s.close()


def test_system_past_4gib_buffers_bit_exact():
    """150M bodies: byte offsets of every body buffer pass 2^32, the deposit
    has 146,485 ranges and the fixed-point scale is the planner's smallest
    (the sums stay within i32).  Inputs from a vectorized generator (a
    uniform cube of half-width 1.2 with small random velocities), so the
    test's host time stays in seconds."""
    n = 150_000_000
    say(f"{n} bodies: 2 steps")
    rng = np.random.default_rng(150)
    parts = np.empty((n, 8), dtype=np.float32)
    parts[:, 0:3] = rng.uniform(-1.2, 1.2, size=(n, 3)).astype(np.float32)
    parts[:, 3] = rng.uniform(0.0, 1.0, size=n).astype(np.float32)
    parts[:, 4:7] = rng.normal(0.0, 0.05, size=(n, 3)).astype(np.float32)
    parts[:, 7] = 0.0
    assert parts.nbytes > 4 * 2**30
    s = rf.Solver(n, rf.physics(**CFG))
    s.upload(parts)
    s.step(2)
    got = s.particles()
    moments = s.moments()
    s.close()
    op, om, _ = orc.step(slots_for(n, **CFG), parts, 2)
    del parts
    assert_bits(moments, om, "moments")
    assert_bits(got, op, "particles")


def cfg4_counts():
    rng = np.random.default_rng(4)
    return [int(round(np.exp(rng.uniform(np.log(4096), np.log(262144))))) for _ in range(256)]


def test_cfg4_256_systems_batch_equals_solo():
    counts = cfg4_counts()
    say(f"cfg4: 256 systems, {sum(counts)} bodies, 3 batched steps (wave tree items) vs solo and oracle")
    prm = rf.physics(**CFG)
    inputs = [rf.ic_plummer(m, 4000 + i, 0.1, 1.5, 1.0) for i, m in enumerate(counts)]
    b = rf.Batch(counts, prm)
    # The shape the autotuner picks for this batch: wave tree items (the
    # solo runs below take the default block items).
    b.set_shape(512, 1)
    for i, p in enumerate(inputs):
        b.upload(i, p)
    b.step(3)
    got = [b.download(i, rf.PARTICLES) for i in range(len(counts))]
    b.close()
    for i, m in enumerate(counts):
        s = rf.Solver(m, prm)
        s.upload(inputs[i])
        s.step(3)
        assert_bits(got[i], s.particles(), f"system {i} (N={m}) batch vs solo")
        s.close()
    order = np.argsort(counts)
    picks = {int(order[0]), int(order[-1])} | {int(i) for i in np.linspace(0, 255, 14).astype(int)}
    for i in sorted(picks):
        op, _, _ = orc.step(slots_for(counts[i], **CFG), inputs[i], 3)
        assert_bits(got[i], op, f"system {i} (N={counts[i]}) vs oracle")
    assert len(picks) >= 16


def cfg5_requests(nreq):
    rng = np.random.default_rng(5)
    reqs = []
    for j in range(nreq):
        n = int(round(np.exp(rng.uniform(np.log(16384), np.log(2097152)))))
        steps = int(rng.integers(1, 201))
        kind = "plummer" if rng.uniform() < 0.5 else "ball"
        reqs.append([j, n, steps, kind])
    reqs[nreq // 3][1] = 2097152  # the largest size of the mix
    return reqs


def request_input(j, n, kind):
    return rf.ic_plummer(n, 5000 + j, 0.1, 1.5, 1.0) if kind == "plummer" else rf.ic_uniform_ball(n, 5000 + j, 1.0)


def test_cfg5_poisson_serving_results_equal_solo():
    nreq = 24
    reqs = cfg5_requests(nreq)
    say(f"cfg5: {nreq} Poisson requests, N up to {max(r[1] for r in reqs)}, 200 GB budget")
    prm = rf.physics(**CFG)
    data = {j: request_input(j, n, kind) for j, n, _, kind in reqs}
    rng = np.random.default_rng(55)
    arrivals = np.cumsum(rng.exponential(0.004, size=nreq))
    srv = rf.Server(200 * 10**9, 64, prm)
    done = {}
    t0 = time.monotonic()
    nxt = 0
    while len(done) < nreq:
        now = time.monotonic() - t0
        while nxt < nreq and arrivals[nxt] <= now:
            j, n, steps, _ = reqs[nxt]
            srv.submit(j, data[j], steps)
            nxt += 1
        srv.poll()
        for c in srv.completions():
            done[c.id] = c
        assert time.monotonic() - t0 < 60, "serving did not finish in 60 s"
    results = {j: srv.result(j, n) for j, n, _, _ in reqs}
    assert srv.state() == (0, 0, 0)
    srv.close()
    for j, n, steps, _ in reqs:
        s = rf.Solver(n, prm)
        s.upload(data[j])
        s.step(steps)
        assert_bits(results[j], s.particles(), f"request {j} (N={n}, {steps} steps) served vs solo")
        s.close()
        assert done[j].finish_s >= done[j].admit_s >= done[j].submit_s
    small = sorted(reqs, key=lambda r: r[1] * r[2])[:3]
    for j, n, steps, _ in small:
        op, _, _ = orc.step(slots_for(n, **CFG), data[j], steps)
        assert_bits(results[j], op, f"request {j} (N={n}, {steps} steps) vs oracle")


def test_server_rejects_over_budget_and_bad_requests():
    say("server: over-budget and malformed requests leave the state unchanged")
    prm = rf.physics(**CFG)
    small = rf.ic_plummer(20000, 1, 0.1, 1.5, 1.0)
    budget = rf.Server.request_bytes(20000) + 1000
    srv = rf.Server(budget, 4, prm)
    big = np.zeros((300000, 8), np.float32)
    with pytest.raises(rf.RainbowError) as e:
        srv.submit(7, big, 3)
    assert e.value.code == rf.E_BUDGET and "budget" in str(e.value)
    assert srv.state() == (0, 0, 0)
    with pytest.raises(rf.RainbowError) as e:
        srv.submit(8, np.zeros((0, 8), np.float32), 3)
    assert e.value.code == rf.E_INVALID
    with pytest.raises(rf.RainbowError) as e:
        srv.result(99, 20000)
    assert e.value.code == rf.E_STATE
    srv.submit(1, small, 2)
    for _ in range(100):
        srv.poll()
        if srv.completions():
            break
    got = srv.result(1, 20000)
    op, _, _ = orc.step(slots_for(20000, **CFG), small, 2)
    assert_bits(got, op, "the request after the rejected ones")
    assert srv.state() == (0, 0, 0)
    srv.close()


def test_solver_rejects_bad_sizes():
    say("solver: wrong buffer sizes and settings")
    n = 20000
    s = rf.Solver(n, rf.physics(**CFG))
    lib = rf.lib()
    buf = np.zeros((n + 1, 8), np.float32)
    assert lib.rainbow_hip_upload_particles(s._h, buf.ctypes.data, buf.nbytes) == rf.E_INVALID
    assert "upload expects" in lib.rainbow_last_error().decode()
    assert lib.rainbow_hip_download(s._h, rf.PARTICLES, buf.ctypes.data, buf.nbytes) == rf.E_INVALID
    assert lib.rainbow_hip_download(s._h, 7, buf.ctypes.data, 32 * n) == rf.E_INVALID
    assert lib.rainbow_hip_set_sort_interval(s._h, 5000) == rf.E_INVALID
    assert lib.rainbow_hip_set_tree_split(s._h, 1 << 9) == rf.E_INVALID
    # the solver still works afterwards
    p = rf.ic_plummer(n, 3, 0.1, 1.5, 1.0)
    s.upload(p)
    s.step(2)
    op, _, _ = orc.step(slots_for(n, **CFG), p, 2)
    assert_bits(s.particles(), op, "particles after the rejected calls")
    s.close()
    b = rf.Batch([5000, 6000], rf.physics(**CFG))
    assert lib.rainbow_batch_upload(b._h, 2, buf.ctypes.data, 32 * 5000) == rf.E_INVALID
    assert "out of range" in lib.rainbow_last_error().decode()
    b.close()


def nonfinite_input(n, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 8), np.float32)
    p[:, :3] = rng.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32)
    p[:, 4:7] = rng.normal(0, 0.2, size=(n, 3)).astype(np.float32)
    p[0, 0], p[1, 1], p[2, 2], p[3, :3] = np.nan, np.inf, -np.inf, np.nan
    p[4, 4], p[5, 5] = np.inf, -np.nan
    p[6, :3] = (np.float32(3e38), np.float32(-3e38), np.float32(1e-45))
    p[:, 3] = rng.uniform(0, 1, size=n).astype(np.float32)
    return p


@pytest.mark.parametrize("n", [999, 70000])
def test_nonfinite_bodies_bit_exact(n):
    say(f"bodies with NaN and infinite coordinates, n={n}")
    parts = nonfinite_input(n, 77 + n)
    dt, g, soft = 2e-3, 0.7, 2e-3
    slots = slots_for(n, dt, g, soft)
    s = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft))
    s.upload(parts)
    s.solve_fields()
    om, of = orc.solve_fields(slots, parts)
    assert_bits(s.moments(), om, "moments")
    assert_bits(s.fields(), of, "fields")
    s.upload(parts)
    s.step(3)
    op, om3, of3 = orc.step(slots, parts, 3)
    assert_bits(s.particles(), op, "particles")
    assert_bits(s.moments(), om3, "moments after steps")
    assert_bits(s.fields(), of3, "fields after steps")
    s.close()


def test_cli_counts_nonfinite_bodies(tmp_path):
    n = 16384
    steps = 4
    parts = nonfinite_input(n, 5)
    path = tmp_path / "bodies.bin"
    parts.tofile(path)
    exe = ROOT / "fmm-rainbow_amd" / "bin" / "rainbow_native"
    say(f"rainbow_native {n} {steps} file: (NaN and infinite bodies)")
    r = subprocess.run([str(exe), str(n), str(steps), f"file:{path}"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    assert "scene: particles from file" in r.stdout
    fh, mh = (int(v, 16) for v in re.search(r"fieldsHash=([0-9a-f]{8}) momentsHash=([0-9a-f]{8})", r.stdout).groups())
    slots = orc.build_plans(n)
    om, of = orc.solve_fields(slots, parts)
    assert orc.fnv1a(of) == fh and orc.fnv1a(om) == mh
    # The CLI runs 300 warm-up steps, 5 timed per phase and `steps` more.
    op, _, _ = orc.step(slots, parts, 305 + steps)
    probe = op[:min(n, 8192)].reshape(-1)
    finite = np.isfinite(probe)
    want = f"non-finite={int((~finite).sum())} maxAbs={float(np.abs(probe[finite]).max()):.6f}"
    assert want in r.stdout, (want, r.stdout[-400:])
    assert int((~finite).sum()) > 0


def test_scale_change_between_steps():
    say("fixed-point scale changed between steps")
    n = 50000
    parts = rf.ic_plummer(n, 12, 0.1, 1.5, 1.0)
    s = rf.Solver(n, rf.physics(**CFG))
    planner = s.scale
    s.upload(parts)
    s.step(3)
    s.set_params(rf.physics(scale=planner / 4, **CFG))
    s.step(18)
    p1, _, _ = orc.step(slots_for(n, **CFG), parts, 3)
    p2, m2, f2 = orc.step(slots_for(n, scale=planner / 4, **CFG), p1, 18)
    assert_bits(s.particles(), p2, "particles after the scale change")
    assert_bits(s.moments(), m2, "moments at the new scale")
    assert_bits(s.fields(), f2, "fields at the new scale")
    s.close()


def test_server_cancel_timeout_and_zero_steps():
    say("server: cancel queued and running requests, timeouts, a zero-step request")
    prm = rf.physics(**CFG)
    sizes = {1: (60000, 40), 2: (30000, 500), 3: (20000, 5), 4: (40000, 7), 5: (16384, 0), 6: (25000, 3)}
    data = {j: rf.ic_plummer(n, 700 + j, 0.1, 1.5, 1.0) for j, (n, _) in sizes.items()}
    # Room for two systems at a time: request 4 waits in the queue.
    budget = rf.Server.request_bytes(60000) + rf.Server.request_bytes(30000)
    srv = rf.Server(budget, 2, prm)
    for j in (1, 2, 3, 4, 5):
        srv.submit(j, data[j], sizes[j][1])
    with pytest.raises(rf.RainbowError) as e:
        srv.submit(3, data[3], 1)
    assert e.value.code == rf.E_INVALID and "in use" in str(e.value)
    srv.poll()
    srv.cancel(4)  # still queued
    srv.cancel(2)  # running (500 steps)
    srv.submit(6, data[6], 3, timeout_s=1e-9)  # expires before it can run
    srv.drain()
    done = {c.id: c for c in srv.completions()}
    assert sorted(done) == [1, 2, 3, 4, 5, 6]
    assert done[4].status == rf.E_CANCELLED and done[4].steps_run == 0
    assert done[2].status == rf.E_CANCELLED and done[2].steps_run < 500
    assert done[6].status == rf.E_TIMEOUT and done[6].steps_run == 0
    for j in (1, 3, 5):
        n, steps = sizes[j]
        assert done[j].status == rf.OK and done[j].steps_run == steps
        op, _, _ = orc.step(slots_for(n, **CFG), data[j], steps)
        assert_bits(srv.result(j, n), op, f"request {j} beside cancelled ones")
    for j in (2, 4, 6):
        with pytest.raises(rf.RainbowError):
            srv.result(j, sizes[j][0])
    with pytest.raises(rf.RainbowError) as e:
        srv.cancel(1)
    assert e.value.code == rf.E_STATE
    assert srv.state() == (0, 0, 0)
    st = srv.stats()
    assert st["steps"] > 0 and st["gpu_step_s"] > 0.0
    srv.close()


def test_server_steady_arrivals_reuse_the_arena():
    # Many more requests than fit at once: every admission after the first
    # few reuses arena memory freed by a finished request, and systems join
    # and leave between steps while others run.
    say("server: 40 requests through room for 5, results equal the oracle's")
    prm = rf.physics(**CFG)
    rng = np.random.default_rng(31)
    reqs = [(j, int(rng.integers(1000, 40000)), int(rng.integers(1, 12))) for j in range(40)]
    data = {}
    for j, n, _ in reqs:
        # Odd requests are borrowed from pinned host arrays (no copy at submit).
        data[j] = rf.host_array(n) if j % 2 else np.empty((n, 8), np.float32)
        data[j][:] = rf.ic_uniform_ball(n, 300 + j, 1.0)
    budget = 5 * rf.Server.request_bytes(40000)
    srv = rf.Server(budget, 64, prm)
    for j, n, steps in reqs[:10]:
        srv.submit(j, data[j], steps, borrow=bool(j % 2))
    nxt = 10
    while nxt < len(reqs):
        srv.poll()
        j, n, steps = reqs[nxt]
        srv.submit(j, data[j], steps, borrow=bool(j % 2))
        nxt += 1
    srv.drain()
    done = {c.id: c for c in srv.completions()}
    assert len(done) == len(reqs)
    for j, n, steps in reqs:
        assert done[j].status == rf.OK and done[j].steps_run == steps
        op, _, _ = orc.step(slots_for(n, **CFG), data[j], steps)
        if j % 3 == 0:
            assert_bits(srv.result_view(j), op, f"request {j} (result view)")
            srv.release(j)
        else:
            assert_bits(srv.result(j, n), op, f"request {j}")
    assert srv.state() == (0, 0, 0)
    srv.close()
