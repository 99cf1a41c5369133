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
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "fmm-rainbow_amd"), str(ROOT / "tests")]

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


FLOPS_PER_PAIR = 50  # pair_field: 40 f32 operations (sqrt, div counted once); 10 additions into the field


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
        return {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
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


def batch_form(args):
    """256 uneven Plummer systems advanced together (one launch sequence per
    step, graph-captured) against the same systems stepped one after another
    with their own solvers."""
    import numpy as np

    import rainbow_fmm as rf

    rng = np.random.default_rng(4)
    counts = [int(round(np.exp(rng.uniform(np.log(4096), np.log(262144))))) for _ in range(args.systems)]
    prm = rf.physics(dt=5e-4, gravity=1.0, softening=1e-3)
    inputs = [rf.ic_plummer(n, 4000 + i, 0.1, 1.5, 1.0) for i, n in enumerate(counts)]
    total = sum(counts)
    b = rf.Batch(counts, prm)
    for i, p in enumerate(inputs):
        b.upload(i, p)
    b.step(args.warmup)
    b.sync()
    ms_batch = b.time_steps(args.steps)
    b.close()
    # Sequential plain launches: each system on its own solver.
    solvers = []
    for i, n in enumerate(counts):
        s = rf.Solver(n, prm)
        s.upload(inputs[i])
        s.step(args.warmup)
        s.sync()
        solvers.append(s)
    # One after another: each solver's steps complete before the next
    # solver starts (separate streams would otherwise overlap them).
    t0 = time.perf_counter()
    for s in solvers:
        s.step(args.steps)
        s.sync()
    seq_s = time.perf_counter() - t0
    for s in solvers:
        s.close()
    batch_rate = total * args.steps / (ms_batch / 1e3)
    seq_rate = total * args.steps / seq_s
    return {
        "metric": "batched particle-steps/s, 256 uneven Plummer systems",
        "value": batch_rate,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_batch / args.steps,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(counts)} Plummer systems, N log-uniform in [4096, 262144] (seed 4), "
                               f"{total} bodies, dt 5e-4, softening 1e-3"},
        "sequential_value": seq_rate,
        "speedup_vs_sequential": batch_rate / seq_rate,
    }


def serve_form(args):
    """Poisson arrivals of N-body requests into the continuous-batching
    server; latency = completion - arrival."""
    import numpy as np

    import rainbow_fmm as rf

    rng = np.random.default_rng(5)
    nreq = args.requests
    reqs = []
    for j in range(nreq):
        n = int(round(np.exp(rng.uniform(np.log(16384), np.log(args.serve_nmax)))))
        steps = int(rng.integers(1, 201))
        kind = "plummer" if rng.uniform() < 0.5 else "ball"
        reqs.append((j, n, steps, kind))
    data = {}
    for j, n, steps, kind in reqs:
        data[j] = (rf.ic_plummer(n, 5000 + j, 0.1, 1.5, 1.0) if kind == "plummer"
                   else rf.ic_uniform_ball(n, 5000 + j, 1.0))
    prm = rf.physics(dt=5e-4, gravity=1.0, softening=1e-3)
    budget = int(args.budget_gb * 1e9)
    # Capacity probe: a mix of the request sizes stepped as one batch.
    probe = [n for _, n, _, _ in reqs[:32]]
    b = rf.Batch(probe, prm)
    for i, n in enumerate(probe):
        b.upload(i, data[i])
    b.step(2)
    b.sync()
    ms = b.time_steps(10)
    b.close()
    capacity = sum(probe) * 10 / (ms / 1e3)
    mean_work = float(np.mean([n * s for _, n, s, _ in reqs]))
    lam = args.load * capacity / mean_work
    gaps = rng.exponential(1.0 / lam, size=nreq)
    arrivals = np.cumsum(gaps) - gaps[0]
    srv = rf.Server(budget, args.max_slots, prm)
    done = {}
    t0 = time.monotonic()
    nxt = 0
    work = 0
    while len(done) < nreq:
        now = time.monotonic() - t0
        while nxt < nreq and arrivals[nxt] <= now:
            j, n, steps, _ = reqs[nxt]
            srv.submit(j, data[j], steps)
            nxt += 1
        stepped = srv.poll()
        for c in srv.completions():
            done[c.id] = c
            work += c.particles * c.steps
        if stepped == 0 and nxt < nreq:
            time.sleep(max(0.0, min(0.001, arrivals[nxt] - (time.monotonic() - t0))))
    wall = time.monotonic() - t0
    lat = np.array([done[j].finish_s - (t0 + arrivals[j]) for j in range(nreq)])
    host = srv.stats()
    srv.close()
    return {
        "metric": "serving: p50/p99 request latency and aggregate particle-steps/s",
        "value": work / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "wall_s": wall,
        "server_host_s": {k: round(v, 4) if isinstance(v, float) else v for k, v in host.items()},
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{nreq} requests, Poisson arrivals at {lam:.1f}/s ({args.load} x probed capacity "
                               f"{capacity:.3g} particle-steps/s), N log-uniform in [16384, {args.serve_nmax}], "
                               f"steps U{{1..200}}, Plummer or uniform ball, budget {args.budget_gb} GB, "
                               f"max {args.max_slots} resident"},
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
    ap.add_argument("--precision-n", type=int, default=5242880)
    ap.add_argument("--systems", type=int, default=256)
    ap.add_argument("--requests", type=int, default=96)
    ap.add_argument("--serve-nmax", type=int, default=2097152)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-slots", type=int, default=256)
    ap.add_argument("--load", type=float, default=0.8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "decomposition is a design note, not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import rainbow_fmm as rf
    import workloads as wl

    if args.form == "batch":
        print(json.dumps(batch_form(args)), flush=True)
        return
    if args.form == "serve":
        print(json.dumps(serve_form(args)), flush=True)
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
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(parts, cfg, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
