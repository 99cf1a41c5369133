"""GPU parity: the HIP path against the oracle, through the C ABI.

The bar is bit-exactness for everything: the integer moments (order-free
fixed-point sums), and the f32 fields and particles, because the kernels
evaluate the reference expressions operation for operation with correctly
rounded division/sqrt and no contraction, and keep the reference's
summation order.  The only exception is the optional control field, which
calls exp (device and host libm differ in the last bits): it is checked to
a relative tolerance of 1e-5.
"""

import sys
import time

import numpy as np
import pytest

import oracle as orc
import rainbow_fmm as rf
import workloads as wl

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def assert_bits(got, want, what):
    got = np.ascontiguousarray(got)
    want = np.ascontiguousarray(want)
    assert got.shape == want.shape, f"{what}: shape {got.shape} != {want.shape}"
    if got.tobytes() == want.tobytes():
        return
    g = got.view(np.uint32) if got.dtype.itemsize == 4 else got
    w = want.view(np.uint32) if want.dtype.itemsize == 4 else want
    bad = np.argwhere(g != w)
    first = tuple(bad[0])
    raise AssertionError(
        f"{what}: {len(bad)} of {g.size} words differ; first at {first}: got {got[first]!r} want {want[first]!r}"
    )


def slots_for(n, dt, gravity, softening, control=None):
    return orc.build_plans(n, dt=dt, gravity=gravity, softening=softening, control=control)


def run_solve(parts, dt, gravity, softening, compensated=False):
    n = parts.shape[0]
    s = rf.Solver(n, rf.physics(dt=dt, gravity=gravity, softening=softening), compensated=compensated)
    s.upload(parts)
    s.solve_fields()
    return s, s.moments(), s.fields()


def test_cfg1_field_solve_bit_exact(golden):
    say("cfg1 solve_fields, 65536 uniform ball")
    parts = wl.cfg1_particles()
    assert orc.fnv1a(parts) == golden["cfg1"]["input_digest"]
    c = wl.CFG1
    s, m, f = run_solve(parts, c["dt"], c["gravity"], c["softening"])
    om, of = orc.solve_fields(slots_for(c["n"], c["dt"], c["gravity"], c["softening"]), parts)
    assert_bits(m, om, "moments")
    assert_bits(f, of, "fields")
    assert s.digest(rf.MOMENTS) == golden["cfg1"]["moments_digest"]
    assert s.digest(rf.FIELDS) == golden["cfg1"]["fields_digest"]
    s.close()


def test_cfg1_ten_steps_bit_exact(golden):
    say("cfg1 10 steps")
    parts = wl.cfg1_particles()
    c = wl.CFG1
    s = rf.Solver(c["n"], wl.params_for(c))
    s.upload(parts)
    s.step(10)
    p = s.particles()
    op, om, of = orc.step(slots_for(c["n"], c["dt"], c["gravity"], c["softening"]), parts, 10)
    assert_bits(p, op, "particles")
    assert_bits(s.moments(), om, "moments")
    assert_bits(s.fields(), of, "fields")
    assert orc.fnv1a(p) == golden["cfg1"]["step_particles_digest"]
    pinned = rf.host_array(c["n"])
    assert s.particles(pinned) is pinned
    assert_bits(pinned, op, "particles into a pinned buffer")
    with pytest.raises(ValueError):
        s.particles(np.empty((c["n"], 4), np.float32))
    s.close()


def test_cfg2_plummer_million_bit_exact(golden):
    say("cfg2 1M Plummer solve + 3 steps")
    parts = wl.cfg2_particles()
    c = wl.CFG2
    assert orc.fnv1a(parts) == golden["cfg2"]["input_digest"]
    slots = slots_for(c["n"], c["dt"], c["gravity"], c["softening"])
    s, m, f = run_solve(parts, c["dt"], c["gravity"], c["softening"])
    om, of = orc.solve_fields(slots, parts)
    assert_bits(m, om, "moments")
    assert_bits(f, of, "fields")
    s.upload(parts)
    s.step(3)
    op, om3, of3 = orc.step(slots, parts, 3)
    assert_bits(s.particles(), op, "particles after 3 steps")
    assert_bits(s.fields(), of3, "fields after 3 steps")
    assert orc.fnv1a(op) == golden["cfg2"]["step_particles_digest"]
    s.close()


def _edge_input(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "plummer":
        return rf.ic_plummer(n, seed, 0.1, 1.5, 1.0)
    if kind == "ball":
        return rf.ic_uniform_ball(n, seed, 1.0)
    p = np.zeros((n, 8), dtype=np.float32)
    if kind == "point":
        p[:, :3] = np.float32(0.25)
    elif kind == "faces":
        # Coordinates exactly on cell faces of several levels, and beyond
        # the world edge on both sides.
        grid = np.array([-1.7, -1.6, -0.8, -0.4, 0.0, 0.2, 0.4, 0.8, 1.2, 1.6, 1.7, 2.5], dtype=np.float32)
        p[:, :3] = rng.choice(grid, size=(n, 3))
        p[:, 4:7] = rng.normal(0, 0.3, size=(n, 3)).astype(np.float32)
    elif kind == "ties":
        # x * scale lands exactly on .5: round-half-to-even cases.
        scale = rf.plan_scale(n)
        k = rng.integers(-1000, 1000, size=(n, 3))
        p[:, :3] = ((k + 0.5) / scale).astype(np.float32)
    elif kind == "clustered":
        centres = rng.uniform(-1.2, 1.2, size=(5, 3))
        idx = rng.integers(0, 5, size=n)
        p[:, :3] = (centres[idx] + rng.normal(0, 0.02, size=(n, 3))).astype(np.float32)
        p[:, 4:7] = rng.normal(0, 0.5, size=(n, 3)).astype(np.float32)
    p[:, 3] = rng.uniform(0, 1, size=n).astype(np.float32)
    p[:, 7] = rng.uniform(0, 1, size=n).astype(np.float32)
    return p


EDGE_CASES = [
    ("ball", 1), ("ball", 2), ("point", 17), ("clustered", 100), ("faces", 777), ("ties", 4096),
    ("plummer", 4096), ("clustered", 16385), ("faces", 32769), ("plummer", 262145), ("point", 300000),
]


@pytest.mark.parametrize("kind,n", EDGE_CASES, ids=[f"{k}-{n}" for k, n in EDGE_CASES])
def test_edge_cases_bit_exact(kind, n):
    say(f"edge {kind} n={n}")
    parts = _edge_input(kind, n, 1234 + n)
    dt, g, soft = 2e-3, 0.7, 2e-3
    slots = slots_for(n, dt, g, soft)
    s, m, f = run_solve(parts, dt, g, soft)
    om, of = orc.solve_fields(slots, parts)
    assert_bits(m, om, "moments")
    assert_bits(f, of, "fields")
    s.upload(parts)
    s.step(3)
    op, om3, of3 = orc.step(slots, parts, 3)
    assert_bits(s.particles(), op, "particles")
    assert_bits(s.moments(), om3, "moments after steps")
    assert_bits(s.fields(), of3, "fields after steps")
    s.close()


def test_compensated_form_bit_exact():
    say("compensated 200k Plummer")
    n = 200000
    parts = rf.ic_plummer(n, 77, 0.1, 1.5, 1.0)
    dt, g, soft = 5e-4, 1.0, 1e-3
    slots = slots_for(n, dt, g, soft)
    s, m, f = run_solve(parts, dt, g, soft, compensated=True)
    om, of = orc.solve_fields(slots, parts, compensated=True)
    assert_bits(m, om, "moments")
    assert_bits(f, of, "fields (compensated)")
    acc = s.accelerations()
    oacc = orc.accelerations(slots, parts, om, of, compensated=True)
    assert_bits(acc[:, :3], oacc[:, :3], "accelerations (compensated)")
    s.upload(parts)
    s.step(2)
    op, _, _ = orc.step(slots, parts, 2, compensated=True)
    assert_bits(s.particles(), op, "particles (compensated)")
    s.close()


def test_accelerations_match_oracle():
    say("accelerations 65536")
    parts = wl.cfg1_particles()
    c = wl.CFG1
    slots = slots_for(c["n"], c["dt"], c["gravity"], c["softening"])
    s, m, f = run_solve(parts, c["dt"], c["gravity"], c["softening"])
    acc = s.accelerations()
    oacc = orc.accelerations(slots, parts, m, f)
    assert_bits(acc, oacc, "accelerations and potential")
    s.close()


def test_control_field_within_tolerance():
    say("control field on")
    n = 50000
    parts = rf.ic_plummer(n, 5, 0.1, 1.5, 1.0)
    control = ((0.1, 0.05, -0.02, 1.0), (0.0, 1.0, 0.0, 3.0), 0.7, 0.22)
    dt, g, soft = 1e-3, 1.0, 1e-3
    slots = slots_for(n, dt, g, soft, control=control)
    s = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft, control=control))
    s.upload(parts)
    s.step(1)
    op, _, _ = orc.step(slots, parts, 1)
    got = s.particles()
    # Relative tolerance 1e-5 on position and velocity (exp differs by ulps
    # between the device and host libraries).
    for sl in (slice(0, 3), slice(4, 7)):
        err = np.linalg.norm(got[:, sl] - op[:, sl], axis=1)
        ref = np.linalg.norm(op[:, sl], axis=1) + 1e-6
        assert np.max(err / ref) < 1e-5
    s.close()


def test_batch_equals_solo():
    say("batch of 6 uneven systems")
    counts = [4096, 70000, 1, 300000, 20000, 262144]
    dt, g, soft = 5e-4, 1.0, 1e-3
    inputs = [rf.ic_plummer(n, 4000 + i, 0.1, 1.5, 1.0) for i, n in enumerate(counts)]
    b = rf.Batch(counts, rf.physics(dt=dt, gravity=g, softening=soft))
    for i, p in enumerate(inputs):
        b.upload(i, p)
    b.step(5)
    for i, n in enumerate(counts):
        op, om, of = orc.step(slots_for(n, dt, g, soft), inputs[i], 5)
        assert_bits(b.download(i, rf.PARTICLES), op, f"system {i} particles")
        assert_bits(b.download(i, rf.MOMENTS), om, f"system {i} moments")
        assert_bits(b.download(i, rf.FIELDS), of, f"system {i} fields")
    b.close()


@pytest.mark.parametrize("seed", range(4))
def test_random_batches_bit_exact(seed):
    """Seeded random batches: 3..10 systems of 1..300000 bodies (Plummer,
    ball or clumps) under shared random physics, steps, summation form,
    shape, sort interval and tree split; every system against the oracle."""
    rng = np.random.default_rng(700 + seed)
    k = int(rng.integers(3, 11))
    counts = [int(round(np.exp(rng.uniform(0.0, np.log(300000))))) for _ in range(k)]
    inputs = []
    for i, n in enumerate(counts):
        kind = int(rng.integers(0, 3))
        if kind == 0:
            inputs.append(rf.ic_plummer(n, 7000 + 16 * seed + i, 0.1, 1.5, 1.0))
        elif kind == 1:
            inputs.append(rf.ic_uniform_ball(n, 7000 + 16 * seed + i, float(rng.uniform(0.3, 1.5))))
        else:
            inputs.append(rf.ic_clumps(n, 7000 + 16 * seed + i, min(n, int(rng.integers(2, 33))), 0.02, 10.0, 1.0))
    phys = dict(dt=float(rng.uniform(1e-4, 3e-3)), gravity=float(rng.uniform(0.2, 2.0)),
                softening=float(np.exp(rng.uniform(np.log(5e-4), np.log(2e-2)))))
    steps = int(rng.integers(1, 13))
    comp = bool(rng.uniform() < 0.3)
    shape = SHAPES[int(rng.integers(0, len(SHAPES)))]
    interval = int(rng.choice([0, 1, 3, 8, 20]))
    split = [None, 1, 0x7C, 4 | (int(rng.integers(0, 32)) << 2)][int(rng.integers(0, 4))]
    say(f"random batch {seed}: counts {counts} {phys} steps={steps} compensated={comp} shape={shape} "
        f"sort interval={interval} split={split}")
    b = rf.Batch(counts, rf.physics(**phys), compensated=comp)
    b.set_shape(*shape)
    b.set_sort_interval(interval)
    if split is not None:
        b.set_tree_split(split)
    for i, p in enumerate(inputs):
        b.upload(i, p)
    b.step(steps)
    for i, n in enumerate(counts):
        op, om, of = orc.step(orc.build_plans(n, **phys), inputs[i], steps, compensated=comp)
        assert_bits(b.download(i, rf.PARTICLES), op, f"system {i} (N={n}) particles")
        assert_bits(b.download(i, rf.MOMENTS), om, f"system {i} (N={n}) moments")
        assert_bits(b.download(i, rf.FIELDS), of, f"system {i} (N={n}) fields")
    b.close()


def test_graph_replay_consistency():
    say("graph replays: 16-step graph vs single steps vs eager oracle")
    n = 100000
    parts = rf.ic_plummer(n, 9, 0.1, 1.5, 1.0)
    dt, g, soft = 5e-4, 1.0, 1e-3
    a = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft))
    a.upload(parts)
    a.step(37)  # two 16-step graph replays + five single-step replays
    b = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft))
    b.upload(parts)
    for _ in range(37):
        b.step(1)
        b.sync()
    op, _, _ = orc.step(slots_for(n, dt, g, soft), parts, 37)
    assert_bits(a.particles(), op, "37 steps through graphs")
    assert_bits(b.particles(), op, "37 single-step replays")
    a.close()
    b.close()


def test_server_results_equal_solo():
    say("server: 7 requests, results alone == in the mix")
    dt, g, soft = 5e-4, 1.0, 1e-3
    prm = rf.physics(dt=dt, gravity=g, softening=soft)
    reqs = [(i, n, st) for i, (n, st) in enumerate([(16384, 5), (50000, 2), (1000, 9), (262145, 3), (30000, 1),
                                                       (70000, 7), (20000, 4)])]
    inputs = {i: rf.ic_uniform_ball(n, 5000 + i, 1.0) if i % 2 else rf.ic_plummer(n, 5000 + i, 0.1, 1.5, 1.0)
              for i, n, _ in reqs}
    budget = sum(rf.Server.request_bytes(n) for _, n, _ in reqs[:4])
    srv = rf.Server(budget, 3, prm)
    for i, n, st in reqs:
        srv.submit(i, inputs[i], st)
    done = {}
    t0 = time.monotonic()
    while len(done) < len(reqs) and time.monotonic() - t0 < 60:
        srv.poll()
        for c in srv.completions():
            done[c.id] = c
    assert len(done) == len(reqs)
    for i, n, st in reqs:
        got = srv.result(i, n)
        op, _, _ = orc.step(slots_for(n, dt, g, soft), inputs[i], st)
        assert_bits(got, op, f"request {i}")
        assert done[i].finish_s >= done[i].admit_s >= done[i].submit_s
    q, r, used = srv.state()
    assert (q, r, used) == (0, 0, 0)
    srv.close()


@pytest.mark.parametrize("compensated", [False, True])
def test_million_and_uneven_batch_bit_exact(compensated):
    say(f"compensated={compensated}, 1M Plummer solve + 2 steps, uneven batch")
    parts = wl.cfg2_particles()
    c = wl.CFG2
    slots = slots_for(c["n"], c["dt"], c["gravity"], c["softening"])
    s, m, f = run_solve(parts, c["dt"], c["gravity"], c["softening"], compensated=compensated)
    om, of = orc.solve_fields(slots, parts, compensated=compensated)
    assert_bits(m, om, "moments")
    assert_bits(f, of, "fields")
    s.upload(parts)
    s.step(2)
    op, _, _ = orc.step(slots, parts, 2, compensated=compensated)
    assert_bits(s.particles(), op, "particles")
    s.close()
    # and a small uneven batch through the same variant
    counts = [4096, 1, 70000]
    ins = [rf.ic_plummer(n, 60 + i, 0.1, 1.5, 1.0) for i, n in enumerate(counts)]
    b = rf.Batch(counts, rf.physics(dt=c["dt"], gravity=c["gravity"], softening=c["softening"]),
                 compensated=compensated)
    for i, p in enumerate(ins):
        b.upload(i, p)
    b.step(3)
    for i, n in enumerate(counts):
        op, _, of = orc.step(slots_for(n, c["dt"], c["gravity"], c["softening"]), ins[i], 3, compensated=compensated)
        assert_bits(b.download(i, rf.PARTICLES), op, f"batch system {i}")
        assert_bits(b.download(i, rf.FIELDS), of, f"batch system {i} fields")
    b.close()


def test_determinism_two_runs():
    say("determinism")
    n = 1 << 20
    parts = wl.cfg2_particles()
    out = []
    for _ in range(2):
        s = rf.Solver(n, wl.params_for(wl.CFG2))
        s.upload(parts)
        s.step(2)
        out.append((s.particles(), s.moments(), s.fields()))
        s.close()
    for k in range(3):
        assert_bits(out[0][k], out[1][k], f"run-to-run buffer {k}")


def test_diagnostics_fp64():
    say("FP64 diagnostics")
    n = 65536
    parts = rf.ic_plummer(n, 11, 0.1, 1.5, 1.0)
    c = dict(dt=5e-4, gravity=1.0, softening=1e-3)
    slots = slots_for(n, **c)
    s, m, f = run_solve(parts, **c)
    d = s.diagnostics()
    cons = orc.conserved_f64(parts)
    acc = orc.accelerations(slots, parts, m, f)
    pe = 0.5 * np.sum(acc[:, 3].astype(np.float64)) / n
    assert abs(d[0] - cons[0]) <= 1e-12 * abs(cons[0]) + 1e-15
    assert abs(d[1] - pe) <= 1e-12 * abs(pe)
    np.testing.assert_allclose(d[2:], cons[2:], rtol=0, atol=1e-14)
    s.close()


@pytest.mark.parametrize("interval", [0, 1, 3, 20, 32])
def test_sort_interval_does_not_change_results(interval):
    say(f"sort interval {interval}")
    n = 150000
    parts = rf.ic_plummer(n, 91, 0.1, 1.5, 1.0)
    dt, g, soft = 5e-4, 1.0, 1e-3
    s = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft))
    s.set_sort_interval(interval)
    s.upload(parts)
    s.step(21)
    op, om, of = orc.step(slots_for(n, dt, g, soft), parts, 21)
    assert_bits(s.particles(), op, "particles")
    assert_bits(s.moments(), om, "moments")
    assert_bits(s.fields(), of, "fields")
    s.close()


def test_autotune_restores_state():
    say("autotune")
    n = 200000
    parts = rf.ic_plummer(n, 92, 0.1, 1.5, 1.0)
    s = rf.Solver(n, rf.physics(dt=5e-4, gravity=1.0, softening=1e-3))
    s.upload(parts)
    s.step(5)
    before = s.particles()
    chosen = s.autotune(8)
    assert chosen in (0, 1, 2, 4, 8, 16, 32)
    # The tuner times both block widths at every pair tile (8, 4, 1 and the
    # per-launch choice 0) and keeps the fastest.
    assert s.shape in {(t, c) for t in (1024, 512) for c in (8, 4, 1, 0)}
    assert_bits(s.particles(), before, "bodies after autotune")
    # The tuned solver still steps bit-exactly.
    s.step(3)
    op, _, _ = orc.step(slots_for(n, 5e-4, 1.0, 1e-3), parts, 8)
    assert_bits(s.particles(), op, "particles after autotune and 3 steps")
    s.close()


SHAPES = [(1024, 8), (512, 8), (1024, 4), (512, 4), (1024, 1), (512, 1), (1024, 0)]


@pytest.mark.parametrize("threads,chunks", SHAPES)
@pytest.mark.parametrize("split", [None, 1, 0x7D])
def test_block_shapes_bit_exact(threads, chunks, split):
    """Every block shape with the measured split, the two-launch form and
    one launch per level: solve + 9 steps (a sort inside) equal the oracle."""
    say(f"shape {threads}x{chunks} split {split}")
    n = 180000
    parts = rf.ic_plummer(n, 95, 0.1, 1.5, 1.0)
    dt, g, soft = 5e-4, 1.0, 1e-3
    s = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft))
    s.set_shape(threads, chunks)
    if split is not None:
        s.set_tree_split(split)
    assert s.shape == (threads, chunks)
    s.upload(parts)
    s.step(9)
    op, om, of = orc.step(slots_for(n, dt, g, soft), parts, 9)
    assert_bits(s.particles(), op, "particles")
    assert_bits(s.moments(), om, "moments")
    assert_bits(s.fields(), of, "fields")
    s.close()


def _random_case(seed):
    """One seeded configuration: bodies (log-uniform in [1000, 400000]),
    initial condition, physics, steps, summation form, block shape, sort
    interval and tree split."""
    rng = np.random.default_rng(seed)
    n = int(round(np.exp(rng.uniform(np.log(1000), np.log(400000)))))
    kind = ["plummer", "ball", "clumps"][int(rng.integers(0, 3))]
    if kind == "plummer":
        parts = rf.ic_plummer(n, 600 + seed, float(rng.uniform(0.05, 0.3)), 1.5, 1.0)
    elif kind == "ball":
        parts = rf.ic_uniform_ball(n, 600 + seed, float(rng.uniform(0.3, 1.5)))
    else:
        parts = rf.ic_clumps(n, 600 + seed, int(rng.integers(2, 65)), 0.02, 10.0, 1.0)
    phys = dict(dt=float(rng.uniform(1e-4, 3e-3)), gravity=float(rng.uniform(0.2, 2.0)),
                softening=float(np.exp(rng.uniform(np.log(5e-4), np.log(2e-2)))),
                speed=float(rng.choice([1.0, 1.5])))
    steps = int(rng.integers(1, 13))
    compensated = bool(rng.uniform() < 0.3)
    shape = SHAPES[int(rng.integers(0, len(SHAPES)))]
    interval = int(rng.choice([0, 1, 3, 8, 20]))
    split = [None, 1, 0x7C, 4 | (int(rng.integers(0, 32)) << 2)][int(rng.integers(0, 4))]
    return n, kind, parts, phys, steps, compensated, shape, interval, split


@pytest.mark.parametrize("seed", range(10))
def test_random_configurations_bit_exact(seed):
    """Seeded random configurations against the oracle: the performance
    settings (shape, sort interval, tree split) must not change any bit, and
    every combination of size, distribution, physics and form must match."""
    n, kind, parts, phys, steps, comp, shape, interval, split = _random_case(seed)
    say(f"random seed {seed}: {kind} n={n} {phys} steps={steps} compensated={comp} shape={shape} "
        f"sort interval={interval} split={split}")
    s = rf.Solver(n, rf.physics(**phys), compensated=comp)
    s.set_shape(*shape)
    s.set_sort_interval(interval)
    if split is not None:
        s.set_tree_split(split)
    s.upload(parts)
    s.step(steps)
    slots = orc.build_plans(n, **phys)
    op, om, of = orc.step(slots, parts, steps, compensated=comp)
    assert_bits(s.particles(), op, "particles")
    assert_bits(s.moments(), om, "moments")
    assert_bits(s.fields(), of, "fields")
    s.close()


@pytest.mark.parametrize("threads,chunks", [(512, 4), (1024, 4), (1024, 1)])
def test_block_shapes_compensated_and_fused(threads, chunks):
    """The compensated form and a fused run of levels (2..5 with ancestor
    recomputation, 6 alone) under the smaller tiles, and under wave items
    (the fused run then takes block items, level 6 wave items)."""
    say(f"shape {threads}x{chunks} compensated, fused 2-5")
    n = 120000
    parts = rf.ic_plummer(n, 96, 0.1, 1.5, 1.0)
    dt, g, soft = 5e-4, 1.0, 1e-3
    s = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft), compensated=True)
    s.set_shape(threads, chunks)
    s.set_tree_fusion(5)
    s.upload(parts)
    s.step(4)
    op, om, of = orc.step(slots_for(n, dt, g, soft), parts, 4, compensated=True)
    assert_bits(s.particles(), op, "particles")
    assert_bits(s.fields(), of, "fields")
    s.close()


@pytest.mark.parametrize("threads,chunks", [(512, 4), (1024, 1)])
def test_batch_shapes_equal_solo(threads, chunks):
    """A batch of uneven systems under a smaller tile or wave items equals
    each system stepped alone under the default shape."""
    say(f"batch shape {threads}x{chunks} = solo")
    counts = [4096, 23000, 70000, 131072, 9000]
    prm = rf.physics(dt=5e-4, gravity=1.0, softening=1e-3)
    inputs = [rf.ic_plummer(n, 700 + i, 0.1, 1.5, 1.0) for i, n in enumerate(counts)]
    b = rf.Batch(counts, prm)
    b.set_shape(threads, chunks)
    assert b.shape == (threads, chunks)
    for i, p in enumerate(inputs):
        b.upload(i, p)
    b.step(10)
    for i, n in enumerate(counts):
        s = rf.Solver(n, prm)
        s.upload(inputs[i])
        s.step(10)
        assert_bits(b.download(i, rf.PARTICLES), s.particles(), f"system {i}")
        s.close()
    b.close()


def test_shape_rejects_other_values():
    say("shape validation")
    s = rf.Solver(5000, rf.physics(dt=5e-4, gravity=1.0, softening=1e-3))
    for bad in [(256, 8), (1024, 16), (2048, 4), (512, 3), (1024, 2)]:
        with pytest.raises(rf.RainbowError) as e:
            s.set_shape(*bad)
        assert e.value.code == rf.E_INVALID
        assert "block shape" in str(e.value)
    assert s.shape == (1024, 8)
    s.close()


@pytest.fixture(scope="module")
def cfg2_reference():
    c = wl.CFG2
    parts = wl.cfg2_particles()
    slots = slots_for(c["n"], c["dt"], c["gravity"], c["softening"])
    om, of = orc.solve_fields(slots, parts)
    op, _, _ = orc.step(slots, parts, 2)
    return parts, om, of, op


@pytest.mark.parametrize("hi", [2, 3, 4, 6])
def test_tree_fusion_does_not_change_results(hi, cfg2_reference):
    say(f"tree fusion levels 2..{hi}, 1M Plummer")
    parts, om, of, op = cfg2_reference
    c = wl.CFG2
    s = rf.Solver(c["n"], wl.params_for(c))
    s.set_tree_fusion(hi)
    launches = s.tree_launches()
    assert launches[0] == (2, hi) and launches[-1][1] == s.depth
    assert len(launches) == 1 + (s.depth - hi)
    s.upload(parts)
    s.solve_fields()
    assert_bits(s.moments(), om, "moments")
    assert_bits(s.fields(), of, "fields")
    s.upload(parts)
    s.step(2)
    assert_bits(s.particles(), op, "particles")
    s.close()


@pytest.mark.parametrize("starts", [1 << 5, (1 << 3) | (1 << 5), (1 << 4)])
def test_tree_split_does_not_change_results(starts, cfg2_reference):
    say(f"tree split mask {starts:#x}, 1M Plummer")
    parts, om, of, op = cfg2_reference
    c = wl.CFG2
    s = rf.Solver(c["n"], wl.params_for(c))
    s.set_tree_split(starts)
    launches = s.tree_launches()
    bounds = sorted({2} | {l for l in range(3, 7) if (starts >> l) & 1})
    assert [lo for lo, _ in launches] == bounds and launches[-1][1] == s.depth
    s.upload(parts)
    s.solve_fields()
    assert_bits(s.fields(), of, "fields")
    s.upload(parts)
    s.step(2)
    assert_bits(s.particles(), op, "particles")
    s.close()


def test_measured_split(cfg2_reference):
    # No split set: the first step records the per-level parent counts and
    # the split follows from them (1343 level-6 parents here: levels 2..5
    # would be fused and level 6 alone, so the two-launch form runs
    # instead); results are the same as with any fixed split.
    say("measured tree split, 1M Plummer")
    parts, om, of, op = cfg2_reference
    c = wl.CFG2
    s = rf.Solver(c["n"], wl.params_for(c))
    s.upload(parts)
    s.step(2)
    assert_bits(s.particles(), op, "particles")
    s.step(16)
    assert s.tree_launch_forms() == [("m2l_top", 4, 5), ("finish", 2, 6)]
    t = rf.Solver(c["n"], wl.params_for(c))
    t.set_tree_fusion(6)
    t.upload(parts)
    t.step(18)
    assert s.digest(0) == t.digest(0)
    s.close()
    t.close()


def test_two_launch_tree_matches_oracle(cfg2_reference):
    # Bit 0 of the split mask: the M2L sums of levels 2..5 in one launch,
    # then every level's fields (L2L from those sums, the finest level's own
    # M2L, near) in a second.
    say("two-launch tree, 1M Plummer")
    parts, om, of, op = cfg2_reference
    c = wl.CFG2
    s = rf.Solver(c["n"], wl.params_for(c))
    s.set_tree_split(1)
    assert s.tree_launch_forms() == [("m2l_top", 4, 5), ("finish", 2, 6)]
    s.upload(parts)
    s.solve_fields()
    assert_bits(s.moments(), om, "moments")
    assert_bits(s.fields(), of, "fields")
    s.upload(parts)
    s.step(2)
    assert_bits(s.particles(), op, "particles")
    s.close()


@pytest.mark.parametrize("n,sort_interval", [(4096, 1), (70000, 1), (400000, 2)])
def test_two_launch_tree_with_top_levels_inside(n, sort_interval):
    # A one-system step in the two-launch form runs the top levels inside
    # the M2L launch (block 0), with the items of levels 2..depth-3 taken
    # directly (their level-2 counts and records from the level-3 moments at
    # depth 6); the body sort follows that launch.  Depths 4, 5 and 6.
    say(f"two-launch tree with the top levels inside, n={n}, sort every {sort_interval}")
    parts = rf.ic_plummer(n, 90 + n % 7, 0.1, 1.5, 1.0)
    dt, g, soft = 5e-4, 1.0, 1e-3
    s = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft))
    s.set_tree_split(1)
    s.set_sort_interval(sort_interval)
    s.upload(parts)
    s.solve_fields()
    om, of = orc.solve_fields(slots_for(n, dt, g, soft), parts)
    assert_bits(s.moments(), om, "moments")
    assert_bits(s.fields(), of, "fields")
    s.upload(parts)
    s.step(5)
    op, _, of = orc.step(slots_for(n, dt, g, soft), parts, 5)
    assert_bits(s.particles(), op, "particles")
    assert_bits(s.fields(), of, "fields after 5 steps")
    s.close()


def test_two_launch_tree_top_levels_sparse_octants():
    # Depth 6 with a dense clump in one octant and exactly 16 and 17 bodies
    # in two others: the direct items of levels 2 and 3 meet parents at the
    # 16-body boundary (skipped at 16, taken at 17) and level-2 windows that
    # are mostly empty, their records formed from the level-3 moments.
    say("two-launch tree with the top levels inside, sparse octants at the 16-body boundary")
    rng = np.random.default_rng(123)
    clump = rf.ic_plummer(400000, 77, 0.05, 0.4, 1.0)
    clump[:, 0:3] += np.float32(0.6)
    def few(k, centre):
        p = np.zeros((k, 8), np.float32)
        p[:, 0:3] = (np.asarray(centre, np.float32) + rng.uniform(-0.3, 0.3, (k, 3))).astype(np.float32)
        return p
    parts = np.concatenate([clump, few(16, (-0.8, -0.8, -0.8)), few(17, (-0.8, 0.8, -0.8))])
    n = parts.shape[0]
    dt, g, soft = 5e-4, 1.0, 1e-3
    s = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft))
    assert s.depth == 6
    s.set_tree_split(1)
    assert s.tree_launch_forms()[0][0] == "m2l_top"
    s.upload(parts)
    s.step(3)
    op, _, of = orc.step(slots_for(n, dt, g, soft), parts, 3)
    assert_bits(s.particles(), op, "particles")
    assert_bits(s.fields(), of, "fields")
    s.close()


@pytest.mark.parametrize("compensated", [False, True])
def test_two_launch_tree_ball_and_uneven_batch(compensated):
    say(f"two-launch tree, uniform ball 400k and an uneven batch, compensated={compensated}")
    n = 400000
    parts = rf.ic_uniform_ball(n, 17, 1.0)
    dt, g, soft = 1e-3, 1.0, 1e-3
    s = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft), compensated=compensated)
    s.set_tree_split(1)
    s.upload(parts)
    s.solve_fields()
    om, of = orc.solve_fields(slots_for(n, dt, g, soft), parts, compensated=compensated)
    assert_bits(s.fields(), of, "fields")
    s.close()
    # Systems of depths 4, 4 and 5 in one batch: the finest level differs
    # per system.
    counts = [4096, 1, 70000]
    ins = [rf.ic_plummer(m, 60 + i, 0.1, 1.5, 1.0) for i, m in enumerate(counts)]
    b = rf.Batch(counts, rf.physics(dt=dt, gravity=g, softening=soft), compensated=compensated)
    b.set_tree_split(1)
    for i, p in enumerate(ins):
        b.upload(i, p)
    b.step(3)
    for i, m in enumerate(counts):
        op, _, of = orc.step(slots_for(m, dt, g, soft), ins[i], 3, compensated=compensated)
        assert_bits(b.download(i, rf.PARTICLES), op, f"batch system {i}")
        assert_bits(b.download(i, rf.FIELDS), of, f"batch system {i} fields")
    b.close()


def test_finish_launch_beyond_one_round():
    """A 2,097,152-body uniform ball: about 4,200 finest-level items, more
    than the finish launch's resident blocks, so the launch splits its
    blocks between the long items (several per block) and the groups of
    short items (several per block); solve + 2 steps equal the oracle."""
    say("two-launch tree beyond one round of blocks, uniform ball 2M")
    n = 2097152
    parts = rf.ic_uniform_ball(n, 21, 1.0)
    dt, g, soft = 5e-4, 1.0, 1e-3
    s = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft))
    s.set_tree_split(1)
    s.upload(parts)
    s.step(2)
    op, om, of = orc.step(slots_for(n, dt, g, soft), parts, 2)
    assert_bits(s.particles(), op, "particles")
    assert_bits(s.fields(), of, "fields")
    s.close()


@pytest.mark.parametrize("compensated", [False, True])
def test_tree_fusion_ball_all_levels(compensated):
    say(f"tree fusion 2..6, uniform ball 400k, compensated={compensated}")
    n = 400000
    parts = rf.ic_uniform_ball(n, 17, 1.0)
    dt, g, soft = 1e-3, 1.0, 1e-3
    s = rf.Solver(n, rf.physics(dt=dt, gravity=g, softening=soft), compensated=compensated)
    s.set_tree_fusion(6)
    s.upload(parts)
    s.solve_fields()
    om, of = orc.solve_fields(slots_for(n, dt, g, soft), parts, compensated=compensated)
    assert_bits(s.moments(), om, "moments")
    assert_bits(s.fields(), of, "fields")
    s.close()


def test_direct_sum_f64_matches_oracle():
    say("FP64 direct sum on the device vs the oracle's")
    n = 40000
    parts = rf.ic_plummer(n, 21, 0.1, 1.5, 1.0)
    c = dict(dt=5e-4, gravity=1.0, softening=1e-3)
    s = rf.Solver(n, rf.physics(**c))
    s.upload(parts)
    s.step(3)  # bodies re-ordered internally: targets are caller indices
    now = s.particles()
    targets = np.array([0, 1, 7, 999, n - 1] + list(range(100, 40000, 157)), dtype=np.uint32)
    got = s.direct_sum_f64(targets)
    # The system holds its softening as an f32.
    want = orc.direct_sum_f64(now, c["gravity"], float(np.float32(c["softening"])), targets)
    scale = np.linalg.norm(want[:, :3], axis=1)
    # FP64 sums in a different order: agreement to ~1e-13 of the field size.
    np.testing.assert_allclose(got[:, :3], want[:, :3], rtol=0, atol=float(1e-12 * scale.max()))
    np.testing.assert_allclose(got[:, 3], want[:, 3], rtol=1e-12)
    s.close()


def test_pair_counts_match_oracle_stencils():
    say("pair counts per level vs the oracle's stencils")
    n = 65536
    parts = rf.ic_plummer(n, 31, 0.1, 1.5, 1.0)
    s = rf.Solver(n, rf.physics(dt=5e-4, gravity=1.0, softening=1e-3))
    s.upload(parts)
    s.solve_fields()
    m2l, near = s.pair_counts()
    m = s.moments()
    counts = m[:, 0]
    want_m2l = [0] * 8
    want_near = [0] * 8
    # Brute force from the moments with numpy (same definitions as the pass).
    from itertools import product
    off = 0
    offs = []
    for level in range(s.depth + 1):
        offs.append(off)
        off += 8 ** level
    def cnt(level, x, y, z):
        d = 1 << level
        if min(x, y, z) < 0 or max(x, y, z) >= d:
            return 0
        return int(counts[offs[level] + x + d * (y + d * z)])
    for level in range(1, s.depth + 1):
        d = 1 << level
        for z, y, x in product(range(d), repeat=3):
            c = cnt(level, x, y, z)
            if c == 0 or cnt(level - 1, x // 2, y // 2, z // 2) <= 16:
                continue
            for pz, py, px in product(range(z // 2 - 1, z // 2 + 2), range(y // 2 - 1, y // 2 + 2),
                                      range(x // 2 - 1, x // 2 + 2)):
                for k in range(8):
                    sx, sy, sz = 2 * px + (k & 1), 2 * py + ((k >> 1) & 1), 2 * pz + (k >> 2)
                    if max(abs(sx - x), abs(sy - y), abs(sz - z)) <= 1:
                        continue
                    if cnt(level, sx, sy, sz):
                        want_m2l[level] += 1
            if c <= 16 or level == s.depth:
                for dz, dy, dx in product((-1, 0, 1), repeat=3):
                    if (dx, dy, dz) != (0, 0, 0) and cnt(level, x + dx, y + dy, z + dz):
                        want_near[level] += 1
    assert m2l == want_m2l and near == want_near
    s.close()
