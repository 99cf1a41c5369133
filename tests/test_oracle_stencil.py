"""The M2L and near-field stencils of the oracle (m2l_all,
src/rainbow_fmm.wgsl:483-517; near_all, :551-570) against set geometry and
the mass-partition invariant: for every body, the M2L sources of its leaf's
ancestors, the leaf's near neighbours and the leaf itself cover the domain
exactly once, so their body counts sum to N.  A wrong separation test, a
wrong parent neighbourhood or a dropped level breaks the sum."""

import itertools

import numpy as np
import pytest

import oracle as orc
import rainbow_fmm as rf


def level_of(g):
    level, start = 1, 1
    while g >= start + 8 ** level:
        start += 8 ** level
        level += 1
    return level, start


def coords(g):
    level, start = level_of(g)
    dim = 1 << level
    i = g - start
    return level, (i % dim, (i // dim) % dim, i // (dim * dim))


def gid(level, c):
    dim = 1 << level
    return (8 ** level - 1) // 7 + c[0] + dim * (c[1] + dim * c[2])


def expected_m2l(level, tc):
    """Children of the parent's 3x3x3 neighbourhood minus the target's own
    3x3x3 neighbourhood, in pz, py, px, cz, cy, cx order."""
    dim = 1 << level
    pd = dim >> 1
    p = [c // 2 for c in tc]
    out = []
    for pz, py, px in itertools.product((-1, 0, 1), repeat=3):
        sp = (p[0] + px, p[1] + py, p[2] + pz)
        if min(sp) < 0 or max(sp) >= pd:
            continue
        for cz, cy, cx in itertools.product((0, 1), repeat=3):
            sc = (2 * sp[0] + cx, 2 * sp[1] + cy, 2 * sp[2] + cz)
            if all(abs(sc[k] - tc[k]) <= 1 for k in range(3)):
                continue
            out.append(gid(level, sc))
    return out


@pytest.mark.parametrize("level", [1, 2, 3, 4, 5, 6])
def test_m2l_lists(level):
    rng = np.random.default_rng(level)
    dim = 1 << level
    cells = [tuple(rng.integers(0, dim, 3)) for _ in range(40)] + [(0, 0, 0), (dim - 1,) * 3]
    for tc in cells:
        g = gid(level, tc)
        got = orc.m2l_sources(g).tolist()
        assert got == expected_m2l(level, tc)
        assert len(got) == len(set(got)) <= 189
    if level >= 3:
        mid = (dim // 2,) * 3
        assert len(orc.m2l_sources(gid(level, mid))) == 189
    else:
        # All cells of levels 1-2 are within reach of their parent's
        # neighbourhood; level 1 has no well-separated cell at all.
        if level == 1:
            assert len(orc.m2l_sources(gid(1, (0, 0, 0)))) == 0


@pytest.mark.parametrize("level", [1, 3, 6])
def test_near_lists(level):
    dim = 1 << level
    for tc in [(0, 0, 0), (dim // 2, dim // 2, dim // 2), (dim - 1, 0, dim // 2)]:
        got = orc.near_sources(gid(level, tc)).tolist()
        want = []
        for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
            if (dx, dy, dz) == (0, 0, 0):
                continue
            s = (tc[0] + dx, tc[1] + dy, tc[2] + dz)
            if min(s) < 0 or max(s) >= dim:
                continue
            want.append(gid(level, s))
        assert got == want


@pytest.mark.parametrize("n,kind", [(65536, "plummer"), (20000, "ball"), (300000, "clumps")])
def test_mass_partition(n, kind):
    if kind == "plummer":
        parts = rf.ic_plummer(n, 21, 0.1, 1.5, 1.0)
    elif kind == "ball":
        parts = rf.ic_uniform_ball(n, 22, 1.0)
    else:
        parts = rf.ic_clumps(n, 23, 16, 0.02, 10.0, 1.0)
    slots = orc.build_plans(n)
    depth = orc.depth_of(slots)
    m, _ = orc.solve_fields(slots, parts)
    counts = m[:, 0].astype(np.int64)
    leaves = orc.leaf_cells(slots, parts, m)
    for leaf in np.unique(leaves):
        level, tc = coords(int(leaf))
        total = counts[leaf]
        total += sum(counts[s] for s in orc.near_sources(int(leaf)))
        for k in range(1, level + 1):
            anc = tuple(c >> (level - k) for c in tc)
            total += sum(counts[s] for s in orc.m2l_sources(gid(k, anc)))
        assert total == n, (leaf, level)
        # the leaf rule of integrate: first level with count <= 16, or depth
        assert counts[leaf] <= 16 or level == depth


def test_m2l_field_is_the_ordered_sum_of_pair_fields():
    n = 65536
    parts = rf.ic_plummer(n, 31, 0.1, 1.5, 1.0)
    slots = orc.build_plans(n, gravity=1.0, softening=1e-3)
    m, _ = orc.solve_fields(slots, parts)
    fields = np.zeros((m.shape[0], 12), np.float32)
    orc.lib().oracle_m2l_all(orc.slot(slots, 0).ctypes.data, m.ctypes.data, fields.ctypes.data)
    slot0 = orc.slot(slots, 0)
    active = [g for g in range(1, m.shape[0]) if m[g, 0] > 0 and np.abs(fields[g]).sum() > 0]
    rng = np.random.default_rng(0)
    for g in rng.choice(active, 40, replace=False):
        level, tc = coords(int(g))
        dim = 1 << level
        centre = np.array([np.float32((np.float32(c) + np.float32(0.5)) * (np.float32(2.0) / np.float32(dim))
                                      - np.float32(1.0)) * np.float32(1.6) for c in tc], np.float32)
        eps = np.float32(np.float32(0.06) * np.float32(1.6)) / np.float32(dim)
        acc = np.zeros(12, np.float32)
        for s in orc.m2l_sources(int(g)):
            acc = (acc + orc.pair_field(slot0, m[s], centre, float(eps))).astype(np.float32)
        np.testing.assert_array_equal(acc, fields[g])
