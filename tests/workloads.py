"""Seeded inputs of the benchmark workloads, shared by tests and bench.

cfg1: 65,536 bodies uniform in the unit ball, seed 1, G=1, softening 1e-3,
dt 1e-3.  cfg2: 1,048,576 bodies in a Plummer sphere (a=0.1, cut at 1.5,
virial velocities), seed 2, G=1, softening 1e-3, dt 5e-4.
"""

import rainbow_fmm as rf

CFG1 = dict(n=65536, seed=1, dt=1e-3, gravity=1.0, softening=1e-3, steps=10)
CFG2 = dict(n=1 << 20, seed=2, dt=5e-4, gravity=1.0, softening=1e-3, steps=60)


def cfg1_particles():
    return rf.ic_uniform_ball(CFG1["n"], CFG1["seed"], 1.0)


def cfg2_particles(n=None):
    return rf.ic_plummer(n or CFG2["n"], CFG2["seed"], 0.1, 1.5, 1.0)


def params_for(cfg, compensated=False):
    return rf.physics(dt=cfg["dt"], gravity=cfg["gravity"], softening=cfg["softening"])
