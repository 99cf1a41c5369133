"""Benchmark: particle-steps/s of the p=1 FMM step at 1M bodies on one GPU.

Workload (default): 1,048,576 equal-mass bodies in a seeded Plummer sphere
(scale radius 0.1, cut at 1.5, virial velocities, seed 2), G=1, softening
1e-3, dt 5e-4, planner depth 6; K timed steps after W warm-up steps.  A
step is deposit, upward, top, tree (levels 2..6 in a few launches),
integrate, with the body sort every few steps before the tree pass,
replayed from captured HIP graphs with the bodies resident in HBM.

Prints ONE JSON line.  `roofline` describes the kernel with the largest
mean duration (event pairs recorded by each kernel dispatch, over the same
number of steps run un-graphed after the timed region): algorithmic bytes
per launch / mean duration against the 8 TB/s HBM peak, or for a tree
launch algorithmic FP32 operations against the 157.3 TFLOP/s FP32 peak.
`cpu_baseline` times the oracle (the C++ CPU restatement of the reference
step, OpenMP) on the same input for a bounded number of steps.
"""

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "fmm-rainbow_amd"), str(ROOT / "tests")]

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


# Algorithmic FP32 operations per non-empty source pair: the reference's
# pair_field plus add_field (src/rainbow_fmm.wgsl:100-118, 163-203) as
# written, 60 with the square root and each division counted once (the
# source's centre of mass and mass share, which the reference forms per
# pair, included; the kernels form them once per cell).
FLOPS_PER_PAIR = 60


def algorithmic_bytes(n, depth, stats, finest_nonempty):
    """Compulsory HBM bytes per launch of each kernel (each unique address once).

    deposit    16 N + 16 C_d             read positions, write finest moments
               (standalone; normally inside integrate)
    upward     read the finest cells and their flag bytes; for the cells
               non-empty now or in the last solve (F) zero the other
               buffer's cell and write the source record; write levels
               depth-1..depth-3 and their source records
    top        read level depth-3, write the levels above
    tree(l)    16 T_l + 48 T_l + 48 P_l: read the level's target moments,
               write their fields, read their parents' fields (the 6^3
               source windows are served from cache)
    integrate  64 N + 64 L + 16 C_d      read+write bodies, leaf moments and
               fields, deposit the next step's finest moments
    with C cells, C_d finest cells, T_l targets of level l, P_l their
    parents, L leaves; F is taken as the finest cells non-empty now (they
    change little per step).
    """
    cells = (8 ** (depth + 1) - 1) // 7
    cd = 8 ** depth
    leaves = stats[7]
    top = (8 ** (depth - 2) - 1) // 7  # cells of levels 0..depth-3
    mid = cells - cd - top + 8 ** (depth - 3)  # levels depth-1..depth-3
    out = {"deposit": 16 * n + 16 * cd,
           "upward": 18 * cd + 32 * finest_nonempty + 16 * mid + 16 * mid,
           "top": 16 * (8 ** (depth - 3)) + 16 * (top - 8 ** (depth - 3))}
    for l in range(2, depth + 1):
        t = stats[l]
        # P_l >= ceil(T_l / 8): each parent has at most 8 targets.
        out[f"tree_l{l}"] = 16 * t + 48 * t + 48 * ((t + 7) // 8)
    # A launch over levels lo..hi moves the sum of its levels' bytes.
    for lo in range(2, depth + 1):
        for hi in range(lo + 1, depth + 1):
            out[f"tree_l{lo}-{hi}"] = sum(out[f"tree_l{l}"] for l in range(lo, hi + 1))
    # Two-launch form: the M2L launch reads the target moments and writes
    # the sums (48 B) of levels lo..hi < depth; the finish launch reads
    # those sums back, the target moments and the parents' ancestor sums,
    # and writes every level's fields.
    for lo, hi in tree_spans(depth):
        m2l_bytes = sum(16 * stats[l] + 48 * stats[l] for l in range(lo, hi + 1) if l < depth)
        out[f"tree_m2l_{span_name(lo, hi)}"] = m2l_bytes
        # With the top levels inside: also levels 2..lo-1 and the top's bytes.
        out[f"tree_m2l_top_{span_name(lo, hi)}"] = (
            sum(16 * stats[l] + 48 * stats[l] for l in range(2, hi + 1) if l < depth) + out["top"])
        out[f"tree_finish_{span_name(lo, hi)}"] = sum(
            out[f"tree_l{l}"] + (48 * stats[l] if l < depth else 0) for l in range(lo, hi + 1))
    out["integrate"] = 64 * n + 64 * leaves + 16 * cd
    # sort: read the bodies and their ranks, write the staged bodies, copy back.
    out["sort"] = 36 * n + 4 * n + 36 * n + 72 * n
    return out


def tree_spans(depth):
    """Every level span lo..hi of a tree launch."""
    return [(lo, hi) for lo in range(2, depth + 1) for hi in range(lo, depth + 1)]


def span_name(lo, hi):
    return f"l{lo}" if lo == hi else f"l{lo}-{hi}"


def algorithmic_flops(depth, m2l, near):
    """FP32 operations of the tree launches: non-empty pairs x FLOPS_PER_PAIR.

    Two-launch form: the M2L launch evaluates the M2L pairs above the finest
    level, the finish launch the finest level's M2L pairs and every level's
    near pairs."""
    out = {f"tree_l{l}": FLOPS_PER_PAIR * (m2l[l] + near[l]) for l in range(2, depth + 1)}
    for lo in range(2, depth + 1):
        for hi in range(lo + 1, depth + 1):
            out[f"tree_l{lo}-{hi}"] = sum(out[f"tree_l{l}"] for l in range(lo, hi + 1))
    for lo, hi in tree_spans(depth):
        out[f"tree_m2l_{span_name(lo, hi)}"] = FLOPS_PER_PAIR * sum(m2l[l] for l in range(lo, hi + 1) if l < depth)
        out[f"tree_m2l_top_{span_name(lo, hi)}"] = FLOPS_PER_PAIR * sum(m2l[l] for l in range(2, hi + 1) if l < depth)
        out[f"tree_finish_{span_name(lo, hi)}"] = FLOPS_PER_PAIR * sum(
            near[l] + (m2l[l] if l == depth else 0) for l in range(lo, hi + 1))
    return out


def roofline_of(phase, ms, bytes_per, flops_per):
    """HBM roofline, or for a tree launch the FP32 compute roofline."""
    if phase in flops_per:
        achieved = flops_per[phase] / (ms * 1e-3) / 1e12
        return {"bound": "fp32-valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "flops_per_launch": flops_per[phase]}
    achieved = bytes_per[phase] / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": bytes_per[phase]}


TRAFFIC_TABLE = ROOT / "profiles" / "hbm_traffic_plummer-1m.md"


def measured_traffic(phase):
    """HBM bytes per launch of `phase` from the committed counter summary.

    The summary holds FETCH_SIZE and WRITE_SIZE per dispatch from two
    `rocprofv3 --pmc` runs of this benchmark, each divided by the factor the
    same counter reads for a copy kernel of known size; its table rows are
    `| phase | kernel | fetch KiB | write KiB | bytes |`.  Returns (bytes,
    source) or (None, None) when the phase is not in the table.
    """
    if not TRAFFIC_TABLE.exists():
        return None, None
    for line in TRAFFIC_TABLE.read_text().splitlines():
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) == 5 and cols[0] == phase:
            try:
                return int(cols[4]), str(TRAFFIC_TABLE.relative_to(ROOT))
            except ValueError:
                return None, None
    return None, None


def cpu_baseline(parts, cfg, budget_s):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc

    slots = orc.build_plans(parts.shape[0], dt=cfg["dt"], gravity=cfg["gravity"], softening=cfg["softening"])
    orc.step(slots, parts[:], 1)  # warm-up
    steps = 0
    t0 = time.perf_counter()
    p = parts
    phases = dict.fromkeys(orc.PHASES, 0.0)
    while True:
        p, sec = orc.step_timed(slots, p, 1)
        for k, v in sec.items():
            phases[k] += v
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s and steps >= 3:
            break
    return {
        "value": parts.shape[0] * steps / el,
        "unit": "particle-steps/s",
        "cores": orc.threads(),
        "kind": "port",
        "sample": f"{steps} steps of the same {parts.shape[0]}-body workload on the host "
                  f"({el:.1f} s, OpenMP threads={orc.threads()})",
        "phases_ms": {k: round(v * 1e3 / steps, 3) for k, v in phases.items()},
    }


def cfg4_counts(systems):
    rng = np.random.default_rng(4)
    return [int(round(np.exp(rng.uniform(np.log(4096), np.log(262144))))) for _ in range(systems)]


def batch_leg(args):
    """256 uneven Plummer systems (N log-uniform in [4096, 262144], seed
    4000 + i) advanced together, one launch sequence per step replayed from
    captured graphs, against the same systems stepped one after another on
    their own solvers.  Both legs are timed the same way: every solver is
    first warmed through a full graph chunk and the sort offset (so no graph
    is captured inside a timed region), then its K steps are timed with
    HIP events on its own stream; the sequential GPU time is the sum of the
    solvers' event times.  The host clock around the whole sequential loop
    (each solver synchronised before the next starts) is reported too."""
    import rainbow_fmm as rf

    counts = cfg4_counts(args.systems)
    prm = rf.physics(dt=5e-4, gravity=1.0, softening=1e-3)
    inputs = [rf.ic_plummer(n, 4000 + i, 0.1, 1.5, 1.0) for i, n in enumerate(counts)]
    total = sum(counts)
    steps = args.batch_steps
    b = rf.Batch(counts, prm)
    for i, p in enumerate(inputs):
        b.upload(i, p)
    b.step(args.warmup)
    b.sync()
    tuned = None
    if args.batch_shape:
        # A fixed block shape (integrate threads, tree chunks) instead of the
        # autotuner's pick, e.g. for profiling the tuned step alone.
        b.set_shape(*(int(x) for x in args.batch_shape.split(",")))
    if args.batch_tune_steps > 0:
        # Block shapes, tree split and sort interval picked by the autotuner
        # (the bodies are restored afterwards).
        interval = b.autotune(args.batch_tune_steps)
        tuned = {"integrate_threads": b.shape[0], "tree_chunks": b.shape[1], "sort_interval": interval}
    ms_batch = b.time_steps(steps)
    b.close()
    batch_rate = total * steps / (ms_batch / 1e3)
    if args.batch_no_sequential:
        return {"workload": f"{len(counts)} Plummer systems, {total} bodies; {steps} timed steps (batched leg only)",
                "value": batch_rate, "unit": "particle-steps/s", "ms_per_step": ms_batch / steps, "tuned": tuned}
    solvers = []
    for i, n in enumerate(counts):
        s = rf.Solver(n, prm)
        s.upload(inputs[i])
        s.step(24)  # a 16-step graph chunk after the first sort, both parities captured
        solvers.append(s)
    for s in solvers:
        s.sync()
    seq_gpu_ms = sum(s.time_steps(steps) for s in solvers)
    t0 = time.perf_counter()
    for s in solvers:
        s.step(steps)
        s.sync()
    seq_wall = time.perf_counter() - t0
    for s in solvers:
        s.close()
    return {
        "workload": f"{len(counts)} Plummer systems a=0.1, N log-uniform in [4096, 262144] (seed 4, system i "
                    f"seeded 4000+i), {total} bodies, dt 5e-4, softening 1e-3; {steps} timed steps",
        "value": batch_rate,
        "unit": "particle-steps/s",
        "ms_per_step": ms_batch / steps,
        "sequential_value": total * steps / (seq_gpu_ms / 1e3),
        "sequential_wall_value": total * steps / seq_wall,
        "speedup": batch_rate / (total * steps / (seq_gpu_ms / 1e3)),
        "speedup_wall": batch_rate / (total * steps / seq_wall),
        "tuned": tuned,
    }


def serve_pool(npool, nmax):
    """Distinct request inputs: N log-uniform in [16384, nmax] (one at nmax),
    a seeded Plummer sphere (a = 0.1) or uniform ball, seed 5000 + j."""
    import rainbow_fmm as rf

    rng = np.random.default_rng(5)
    pool = []
    for j in range(npool):
        n = int(round(np.exp(rng.uniform(np.log(16384), np.log(nmax)))))
        if j == npool // 3:
            n = nmax
        plummer = rng.uniform() < 0.5
        data = rf.host_array(n)  # pinned: requests are uploaded by DMA straight from it
        data[:] = rf.ic_plummer(n, 5000 + j, 0.1, 1.5, 1.0) if plummer else rf.ic_uniform_ball(n, 5000 + j, 1.0)
        pool.append((n, "plummer" if plummer else "ball", data))
    return pool


def run_server(srv, pool, reqs, arrivals):
    """Submits reqs (pool index, steps) at their arrival times (seconds from
    the start) while polling; returns (wall s, {id: completion}, t0)."""
    done = {}
    t0 = time.monotonic()
    nxt = 0
    while len(done) < len(reqs):
        now = time.monotonic() - t0
        while nxt < len(reqs) and arrivals[nxt] <= now:
            k, steps = reqs[nxt]
            srv.submit(nxt, pool[k][2], steps, borrow=True)
            nxt += 1
        stepped = srv.poll()
        for c in srv.completions():
            done[c.id] = c
            if c.status == 0:
                srv.result_view(c.id)  # the client takes the result in place
                srv.release(c.id)
        if stepped == 0 and nxt < len(reqs):
            queued, resident, _ = srv.state()
            wait = arrivals[nxt] - (time.monotonic() - t0)
            if queued == 0 and resident == 0 and wait > 0:
                time.sleep(min(wait, 0.0005))
    return time.monotonic() - t0, done, t0


def serve_leg(args):
    """The continuous-batching server on mixed-N requests (N log-uniform in
    [16384, 2097152], 1..200 steps, Plummer or uniform ball, 200 GB budget):

    * saturated: every request submitted at once (four times the resident
      limit by default, so the set stays full while the backlog lasts);
      aggregate particle-steps/s over the wall time (uploads, steps and
      downloads included, ramp and drain too) and the device's busy fraction
      (event time of the steps over the wall);
    * sequential: the same requests one after another on plain solvers
      (upload from and download into pinned buffers, as the server does;
      solvers created and warmed beforehand), on the same host clock; the
      ratio is the served speedup (also reported against the same loop
      downloading into fresh pageable arrays, the plain API's default);
    * Poisson: arrivals at `load` x the saturated rate for `arrival_s`
      seconds; p50/p99 of (completion - arrival), next to the same requests'
      solo latency (upload + steps + download on a plain solver)."""
    import rainbow_fmm as rf

    prm = rf.physics(dt=5e-4, gravity=1.0, softening=1e-3)
    pool = serve_pool(args.serve_pool, args.serve_nmax)
    rng = np.random.default_rng(55)
    budget = int(args.budget_gb * 1e9)

    # Solo latency model per pool entry: upload + download (host clock) and
    # per-step GPU time (graph replays), from a warmed plain solver.  Plain
    # solvers download into pinned buffers, as the server does.
    solvers, results = [], []
    for n, _, data in ([] if args.serve_saturated_only else pool):
        s = rf.Solver(n, prm)
        s.upload(data)
        s.step(24)
        s.sync()
        solvers.append(s)
        results.append(rf.host_array(n))
    updown, per_step = [], []
    for (n, _, data), s, out in zip(pool, solvers, results):
        t0 = time.perf_counter()
        s.upload(data)
        s.particles(out)
        updown.append(time.perf_counter() - t0)
        per_step.append(s.time_steps(32) / 32 / 1e3)

    # Saturated: all at once.
    sat = [(int(rng.integers(0, len(pool))), int(rng.integers(1, 201))) for _ in range(args.requests)]
    work = sum(pool[k][0] * st for k, st in sat)
    srv = rf.Server(budget, args.max_slots, prm)
    wall, done, _ = run_server(srv, pool, sat, np.zeros(len(sat)))
    st_sat = srv.stats()
    srv.close()
    sat_rate = work / wall
    if args.serve_saturated_only:
        return {"workload": f"{len(sat)} requests at once from {len(pool)} seeded inputs (saturated leg only)",
                "value": sat_rate, "unit": "particle-steps/s", "wall_s": wall,
                "gpu_busy": st_sat["gpu_step_s"] / wall, "steps": st_sat["steps"],
                "all_completed": all(c.status == 0 for c in done.values())}

    # Sequential: the same requests one after another on plain solvers,
    # results into pinned buffers (the comparison's baseline), then once more
    # into fresh pageable arrays (the plain API's default download).
    def sequential(pinned):
        t0 = time.perf_counter()
        lat = []
        for k, steps in sat:
            t1 = time.perf_counter()
            s = solvers[k]
            s.upload(pool[k][2])
            s.step(steps)
            s.particles(results[k] if pinned else None)
            lat.append(time.perf_counter() - t1)
        return work / (time.perf_counter() - t0), lat

    seq_rate, solo_lat = sequential(True)
    seq_rate_pageable, _ = sequential(False)

    # Poisson arrivals at load x the saturated rate.
    mean_work = float(np.mean([pool[k][0] * st for k, st in sat]))
    lam = args.load * sat_rate / mean_work
    nreq = max(8, int(lam * args.arrival_s))
    reqs = [(int(rng.integers(0, len(pool))), int(rng.integers(1, 201))) for _ in range(nreq)]
    gaps = rng.exponential(1.0 / lam, size=nreq)
    arrivals = np.cumsum(gaps) - gaps[0]
    srv = rf.Server(budget, args.max_slots, prm)
    wall_p, done_p, t0p = run_server(srv, pool, reqs, arrivals)
    st_p = srv.stats()
    srv.close()
    for s in solvers:
        s.close()
    lat = np.array([done_p[j].finish_s - (t0p + arrivals[j]) for j in range(nreq)])
    solo = np.array([updown[k] + st * per_step[k] for k, st in reqs])
    work_p = sum(pool[k][0] * st for k, st in reqs)
    ok = all(c.status == 0 for c in done_p.values()) and all(c.status == 0 for c in done.values())
    return {
        "workload": f"requests from {len(pool)} seeded inputs (N log-uniform in [16384, {args.serve_nmax}], one "
                    f"at {args.serve_nmax}; Plummer a=0.1 or uniform ball), steps U{{1..200}}, dt 5e-4, "
                    f"softening 1e-3, budget {args.budget_gb:g} GB, at most {args.max_slots} resident",
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "value": work_p / wall_p,
        "unit": "particle-steps/s",
        "poisson": {"requests": nreq, "rate_per_s": lam, "load": args.load, "arrival_s": float(arrivals[-1]),
                    "wall_s": wall_p, "gpu_busy": st_p["gpu_step_s"] / wall_p,
                    "solo_p50_ms": float(np.percentile(solo, 50) * 1e3),
                    "solo_p99_ms": float(np.percentile(solo, 99) * 1e3),
                    "slowdown_p50": float(np.percentile(lat / solo, 50)),
                    "slowdown_p99": float(np.percentile(lat / solo, 99))},
        "saturated": {"requests": len(sat), "value": sat_rate, "wall_s": wall,
                      "gpu_busy": st_sat["gpu_step_s"] / wall, "steps": st_sat["steps"],
                      "host_s": {k: round(v, 4) for k, v in st_sat.items() if k.endswith("_s")}},
        "sequential_value": seq_rate,
        "speedup": sat_rate / seq_rate,
        "sequential_pageable_value": seq_rate_pageable,
        "speedup_vs_pageable": sat_rate / seq_rate_pageable,
        "solo_p50_ms": float(np.percentile(solo_lat, 50) * 1e3),
        "solo_p99_ms": float(np.percentile(solo_lat, 99) * 1e3),
        "all_completed": ok,
    }


def precision_form(args):
    """5,242,880 bodies in 64 Plummer clumps (a = 0.02, cut at 10 a; masses
    log-uniform over two decades; centres uniform in the unit cube), FP32
    with Neumaier-compensated sums, 20 steps.  Force error of the FMM
    acceleration against the FP64 direct sum at 4096 sampled bodies (both
    the compensated and the plain FP32 variant), FP64 energy and momentum
    drift over the steps."""
    import numpy as np

    import rainbow_fmm as rf

    n = args.precision_n
    steps = 20
    parts = rf.ic_clumps(n, 7, 64, 0.02, 10.0, 1.0)
    prm = rf.physics(dt=5e-4, gravity=1.0, softening=1e-3)
    rng = np.random.default_rng(8)
    targets = np.sort(rng.choice(n, size=4096, replace=False)).astype(np.uint32)

    def force_error(compensated):
        s = rf.Solver(n, prm, compensated=compensated)
        s.upload(parts)
        s.solve_fields()
        acc = s.accelerations()[targets]
        ref = s.direct_sum_f64(targets)
        s.close()
        num = np.linalg.norm(acc[:, :3].astype(np.float64) - ref[:, :3], axis=1)
        den = np.linalg.norm(ref[:, :3], axis=1)
        rel = num / den
        return {"median": float(np.median(rel)), "p90": float(np.percentile(rel, 90)),
                "p99": float(np.percentile(rel, 99)), "max": float(rel.max()),
                "rms": float(np.sqrt(np.mean(rel ** 2))),
                "phi_rel_median": float(np.median(np.abs(acc[:, 3] - ref[:, 3]) / np.abs(ref[:, 3])))}

    err_comp = force_error(True)
    err_plain = force_error(False)

    s = rf.Solver(n, prm, compensated=True)
    depth = s.depth
    s.upload(parts)
    s.solve_fields()
    d0 = s.diagnostics()
    ms = s.time_steps(steps)
    s.solve_fields()
    d1 = s.diagnostics()
    s.close()
    e0, e1 = d0[0] + d0[1], d1[0] + d1[1]
    p0, p1 = np.linalg.norm(d0[2:5]), np.linalg.norm(d1[2:5])
    return {
        "metric": "precision form: force error vs FP64 direct sum, energy drift",
        "value": n * steps / (ms / 1e3),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "ms_per_step": ms / steps,
        "higher_is_better": True,
        "dtype": "f32 (compensated sums) + f64 reductions",
        "data": "synthetic",
        "config": {"workload": f"clumps-{n}: 64 Plummer clumps a=0.02 cut at 10a (seed 7), masses 10^U(0,2), "
                               f"centres in the unit cube, G=1, softening 1e-3, dt 5e-4, depth {depth}"},
        "force_error_compensated": err_comp,
        "force_error_plain_f32": err_plain,
        "force_error_targets": 4096,
        "energy": {"E0": e0, "E20": e1, "relative_drift": (e1 - e0) / abs(e0),
                   "kinetic0": d0[0], "potential0": d0[1]},
        "momentum": {"P0": p0, "P20": p1, "angular0": d0[5:8].tolist(), "angular20": d1[5:8].tolist()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--phase-steps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", choices=["plain", "batch", "serve", "precision"], default="plain")
    ap.add_argument("--no-batch", action="store_true", help="leave the batched form out of the plain line")
    ap.add_argument("--no-serve", action="store_true", help="leave the serving form out of the plain line")
    ap.add_argument("--precision-n", type=int, default=5242880)
    ap.add_argument("--systems", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=100, help="timed steps of the batched leg (the configured 100)")
    ap.add_argument("--batch-no-sequential", action="store_true", help="time the batched leg only (profiling)")
    ap.add_argument("--batch-shape", default="", help="fixed batch block shape 'threads,chunks' (with --batch-tune-steps 0)")
    ap.add_argument("--batch-tune-steps", type=int, default=100,
                    help="steps per candidate of the batch's autotuner, the horizon it tunes for (0: default "
                         "shapes and measured split)")
    ap.add_argument("--requests", type=int, default=1024, help="requests of the saturated serving run")
    ap.add_argument("--serve-pool", type=int, default=32, help="distinct seeded request inputs")
    ap.add_argument("--serve-saturated-only", action="store_true", help="time the saturated serving leg only (profiling)")
    ap.add_argument("--arrival-s", type=float, default=10.0, help="seconds of Poisson arrivals")
    ap.add_argument("--serve-nmax", type=int, default=2097152)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-slots", type=int, default=256)
    ap.add_argument("--load", type=float, default=0.8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "decomposition is a design note, not built.", file=sys.stderr)
        sys.exit(2)

    import rainbow_fmm as rf
    import workloads as wl

    if args.form == "batch":
        print(json.dumps({"metric": "batched particle-steps/s, 256 uneven Plummer systems", "n_gpus": 1,
                          "dtype": "f32", "data": "synthetic", **batch_leg(args)}), flush=True)
        return
    if args.form == "serve":
        print(json.dumps({"metric": "serving: p50/p99 request latency and aggregate particle-steps/s",
                          "n_gpus": 1, "dtype": "f32", "data": "synthetic", **serve_leg(args)}), flush=True)
        return
    if args.form == "precision":
        print(json.dumps(precision_form(args)), flush=True)
        return

    cfg = dict(wl.CFG2)
    n = args.n
    parts = wl.cfg2_particles(n)
    solver = rf.Solver(n, wl.params_for(cfg))
    solver.upload(parts)
    solver.step(args.warmup)
    solver.sync()
    wall0 = time.perf_counter()
    ms = solver.time_steps(args.steps)
    wall = time.perf_counter() - wall0
    value = n * args.steps / (ms / 1e3)

    # Per-kernel mean durations, HIP events around each launch (un-graphed).
    solver.solve_fields()
    stats = solver.tree_stats()
    phases = solver.phase_ms(args.phase_steps)
    names = solver.phase_names()
    per = dict(zip(names, phases))
    moments = solver.moments()
    finest_nonempty = int(np.count_nonzero(moments[rf.level_offset(solver.depth):, 0]))
    bytes_per = algorithmic_bytes(n, solver.depth, stats, finest_nonempty)
    m2l, near = solver.pair_counts()
    flops_per = algorithmic_flops(solver.depth, m2l, near)
    # The dominant kernel: the phase with the largest mean duration (the
    # sort is several library kernels and is not a candidate).
    dom = max((k for k in per if k != "sort"), key=per.get)
    traffic, traffic_src = measured_traffic(dom)
    roofline = {"kernel": dom, **roofline_of(dom, per[dom], bytes_per, flops_per),
                "traffic": traffic, "traffic_source": traffic_src, "mean_ms": round(per[dom], 5)}
    # Phases that run every step (the standalone deposit and the sort do not).
    phase_roofline = {k: roofline_of(k, v, bytes_per, flops_per)["frac"]
                      for k, v in per.items() if v > 0 and k in bytes_per and k not in ("deposit", "sort")}
    line = {
        "metric": "particle-steps/s at 1M bodies on 1 GPU; p50/p99 latency for mixed-N serving",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plummer-{n}: seeded Plummer sphere a=0.1 (seed 2), G=1, softening 1e-3, dt 5e-4, "
                        f"depth {solver.depth}",
            "bodies": n,
            "depth": solver.depth,
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "phases_ms": {k: round(v, 5) for k, v in per.items()},
        "phases_roofline_frac": phase_roofline,
        "pairs": {"m2l": m2l[1:solver.depth + 1], "near": near[1:solver.depth + 1]},
        "tree": {"targets_per_level": stats[1:1 + solver.depth], "leaves": stats[7]},
        "wall_s_timed": wall,
    }
    solver.close()
    # The metric's second half and the batched form, on the same device.
    if not args.no_serve:
        print("bench: serving form", file=sys.stderr, flush=True)
        line["serving"] = serve_leg(args)
    if not args.no_batch:
        print("bench: batched form", file=sys.stderr, flush=True)
        line["batch"] = batch_leg(args)
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(parts, cfg, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
