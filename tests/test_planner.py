"""Planner ABI: the reference's own unit tests (src/lib.rs:163-217), the FFI
clamping of src/lib.rs:113-157, and the closed-form plan table (depth,
cells, scale per body count) against both the library and the oracle's
restatement."""

import numpy as np
import pytest

import oracle as orc
import rainbow_fmm as rf

MAX_DEPTH = 6


def test_octree_layout_is_contiguous():
    # src/lib.rs:163-172
    for level in range(0, MAX_DEPTH + 1):
        assert rf.level_offset(level) + rf.level_cells(level) == rf.level_offset(level + 1)
    assert rf.total_cells(6) == 299_593


def test_million_particle_plan_uses_depth_six():
    # src/lib.rs:174-179
    assert rf.plan_depth(1 << 20) == 6
    assert rf.level_cells(6) == 262_144
    assert rf.clamp_particle_count(0xFFFFFFFF) == 1 << 23


def test_multi_million_plans_stay_at_depth_six():
    # src/lib.rs:181-189
    for n in (1 << 21, 1 << 22, 1 << 23):
        assert rf.plan_depth(n) == 6
    assert rf.total_cells(6) == 299_593


def test_fixed_point_scale_stays_inside_i32():
    # src/lib.rs:191-203
    for n in (1 << 14, 1 << 20, 1 << 21, 1 << 22, 1 << 23):
        scale = float(rf.plan_scale(n))
        worst = n * float(np.float32(1.6)) * scale
        assert worst < 2**31 - 1
        assert np.log2(scale) == np.floor(np.log2(scale))


def test_work_bound_is_linear_for_a_fixed_depth():
    # src/lib.rs:205-210
    a = rf.plan_linear_work(1 << 18, 6)
    b = rf.plan_linear_work(1 << 19, 6)
    assert b - a == 3 * (1 << 18)


def test_ffi_clamping():
    # src/lib.rs:113-157
    assert rf.level_cells(7) == 0 and rf.level_cells(6) == 262144
    assert rf.level_offset(8) == 0 and rf.level_offset(7) == 299593
    assert rf.total_cells(0) == 4681 and rf.total_cells(3) == 4681 and rf.total_cells(9) == 299593
    assert rf.clamp_particle_count(1) == 16384 and rf.clamp_particle_count(100000) == 100000
    assert rf.depth_for_particles(4096) == 4
    assert rf.fixed_point_scale(4096) == 65536.0  # clamped to 2^14 bodies
    assert rf.plan_scale(4096) == 262144.0  # the harness's unclamped planner
    # low 32 bits of 3 n + 199 C
    assert rf.linear_work_lo(1 << 23, 6) == (3 * (1 << 23) + 199 * 299593) & 0xFFFFFFFF
    assert rf.linear_work_lo(5, 2) == 3 * 16384 + 199 * 4681


def test_plan_table_closed_form(golden):
    for row in golden["planner"]:
        n = row["n"]
        assert rf.plan_depth(n) == row["depth"], n
        assert rf.plan_cells(n) == row["cells"], n
        assert rf.plan_scale(n) == row["scale"], n
        assert rf.depth_for_particles(n) == row["ffi_depth"], n
        assert rf.total_cells(rf.depth_for_particles(n)) == row["ffi_cells"], n
        assert rf.fixed_point_scale(n) == row["ffi_scale"], n
        d, c, s = orc.plan(n)
        assert (d, c, s) == (row["depth"], row["cells"], row["scale"]), n


def test_survey_planner_values():
    # Values derived independently of this code (planner table of the
    # reference survey): (n, depth, cells, unclamped scale).
    for n, d, c, s in [(4096, 4, 4681, 262144), (16384, 4, 4681, 65536), (65536, 5, 37449, 16384),
                       (262144, 5, 37449, 4096), (262145, 6, 299593, 4096), (1048576, 6, 299593, 1024),
                       (2097152, 6, 299593, 512), (5242880, 6, 299593, 128), (8388608, 6, 299593, 128)]:
        assert (rf.plan_depth(n), rf.plan_cells(n), rf.plan_scale(n)) == (d, c, s)


@pytest.mark.parametrize("n", [1, 7, 8, 9, 4095, 4097, 32768, 32769, 262144, 262145, 3_000_001, 2**32 - 1])
def test_library_and_oracle_planners_agree(n):
    d, c, s = orc.plan(n)
    assert rf.plan_depth(n) == d and rf.plan_cells(n) == c and rf.plan_scale(n) == s
