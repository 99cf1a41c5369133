"""Integer parts of the oracle -- binning (position_cell,
src/rainbow_fmm.wgsl:144-152), deposit (:408-422) and upward (:424-455) --
against an independent numpy brute force, hand-derived special cases and
invariants.  These are exact: any mismatch is a bug."""

import numpy as np
import pytest

import oracle as orc
import rainbow_fmm as rf


def numpy_bin(p, dim):
    """floor((clamp(p / 1.6, -0.99999, 0.99999) * 0.5 + 0.5) * dim) in f32."""
    p = np.asarray(p, dtype=np.float32)
    n = np.clip(p / np.float32(1.6), np.float32(-0.99999), np.float32(0.99999)).astype(np.float32)
    u = (n * np.float32(0.5) + np.float32(0.5)).astype(np.float32)
    return np.floor(u * np.float32(dim)).astype(np.int64)


def numpy_moments(parts, depth, scale):
    """Per-cell count and fixed-point sums of every level, by bincount."""
    dim = 1 << depth
    c = [numpy_bin(parts[:, k], dim) for k in range(3)]
    key = c[0] + dim * (c[1] + dim * c[2])
    fixed = np.rint(parts[:, :3].astype(np.float64) * scale).astype(np.int64)
    levels = {}
    cnt = np.bincount(key, minlength=dim ** 3)
    sums = [np.bincount(key, weights=fixed[:, k], minlength=dim ** 3).astype(np.int64) for k in range(3)]
    levels[depth] = (cnt, sums)
    for l in range(depth - 1, -1, -1):
        d2 = 1 << (l + 1)
        cnt_c, sums_c = levels[l + 1]
        shape = (1 << l, 2, 1 << l, 2, 1 << l, 2)  # z, dz, y, dy, x, dx
        cnt_p = cnt_c.reshape(d2, d2, d2).reshape(shape).sum(axis=(1, 3, 5)).reshape(-1)
        sums_p = [s.reshape(d2, d2, d2).reshape(shape).sum(axis=(1, 3, 5)).reshape(-1) for s in sums_c]
        levels[l] = (cnt_p, sums_p)
    return levels


def as_i32(x):
    return ((np.asarray(x, dtype=np.int64) + 2**31) % 2**32 - 2**31).astype(np.int64)


def test_position_cell_special_cases():
    # Hand-derived: p = 0 lies on the central face and goes to the upper cell;
    # |p| >= 1.6 clamps into the edge cells; 0.8 / 1.6 = 0.5 exactly.
    for dim in (2, 4, 8, 16, 32, 64):
        assert orc.position_cell(0.0, dim) == dim // 2
        assert orc.position_cell(-1.6, dim) == 0
        assert orc.position_cell(1.6, dim) == dim - 1
        assert orc.position_cell(1e30, dim) == dim - 1
        assert orc.position_cell(-1e30, dim) == 0
        assert orc.position_cell(float("inf"), dim) == dim - 1
    assert orc.position_cell(0.8, 4) == 3  # u = 0.75 -> 3.0
    assert orc.position_cell(-0.8, 4) == 1  # u = 0.25 -> 1.0
    assert orc.position_cell(0.4, 8) == 5  # u = 0.625 -> 5.0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_binning_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.uniform(-2, 2, 20000), rng.normal(0, 0.3, 20000),
                        np.array([0, 0.8, -0.8, 1.6, -1.6, 0.4, -0.4, 1.2, 1e-8, -1e-8])]).astype(np.float32)
    for dim in (2, 16, 64):
        want = numpy_bin(p, dim)
        got = np.array([orc.position_cell(float(x), dim) for x in p[:3000]] +
                       [orc.position_cell(float(x), dim) for x in p[-10:]])
        np.testing.assert_array_equal(got, np.concatenate([want[:3000], want[-10:]]))


@pytest.mark.parametrize("n,kind", [(16384, "ball"), (65536, "plummer"), (100003, "wide")])
def test_deposit_and_upward_match_numpy(n, kind):
    if kind == "ball":
        parts = rf.ic_uniform_ball(n, 1, 1.0)
    elif kind == "plummer":
        parts = rf.ic_plummer(n, 3, 0.1, 1.5, 1.0)
    else:
        rng = np.random.default_rng(5)
        parts = np.zeros((n, 8), dtype=np.float32)
        parts[:, :3] = rng.uniform(-2.0, 2.0, size=(n, 3))
    slots = orc.build_plans(n)
    depth, cells, scale = orc.plan(n)
    m, _ = orc.solve_fields(slots, parts)
    levels = numpy_moments(parts, depth, scale)
    for l in range(depth + 1):
        off = (8 ** l - 1) // 7
        cnt, sums = levels[l]
        got = m[off:off + 8 ** l]
        np.testing.assert_array_equal(got[:, 0].astype(np.int64), cnt)
        for k in range(3):
            np.testing.assert_array_equal(got[:, 1 + k].astype(np.int64), as_i32(sums[k]))
    # Invariants: the root holds every body and the total sums.
    assert m[0, 0] == n
    assert m[0, 1:].astype(np.int64).tolist() == as_i32(
        np.rint(parts[:, :3].astype(np.float64) * scale).astype(np.int64).sum(0)).tolist()


def test_round_half_to_even():
    # x * scale exactly on .5: WGSL round() is half-to-even.
    n = 16384
    scale = orc.plan(n)[2]
    ks = np.array([0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 7.5], dtype=np.float64)
    parts = np.zeros((n, 8), dtype=np.float32)
    parts[:len(ks), 0] = (ks / scale).astype(np.float32)
    m, _ = orc.solve_fields(orc.build_plans(n), parts)
    # every other body sits at 0 and adds 0
    assert m[0, 1] == sum([0, 2, 2, 0, -2, -2, 8])


def test_timed_step_is_the_step():
    # the CPU baseline's per-phase timing runs the same phases in the same
    # order: identical bodies, and a time for each of the seven phases
    n = 20000
    parts = np.zeros((n, 8), np.float32)
    rng = np.random.default_rng(17)
    parts[:, :3] = rng.normal(0, 0.3, (n, 3)).clip(-1.5, 1.5)
    parts[:, 4:7] = rng.normal(0, 0.1, (n, 3))
    slots = orc.build_plans(n, dt=1e-3, gravity=1.0, softening=1e-3)
    a, _, _ = orc.step(slots, parts, 2)
    b, sec = orc.step_timed(slots, parts, 2)
    np.testing.assert_array_equal(a, b)
    assert tuple(sec) == orc.PHASES and all(v >= 0.0 for v in sec.values())
