"""The oracle's Neumaier-compensated sums (the precision form's FP32
accumulation in M2L, near field and the own-leaf term) against what the
mathematics fixes, without retyping the accumulator:

* known answers where plain FP32 loses every digit and the two branches of
  Neumaier's update differ (a large term added to a small running sum, and
  the reverse), and a run of compensations that must add up;
* random ill-conditioned sums against ``math.fsum`` (the exact sum), within
  Neumaier's bound |s - S| <= u |S| + 2 n u^2 sum|x_i| (u = 2^-24), where
  plain FP32 misses the same bound by orders of magnitude;
* inside the field solve: a level-2 cell whose parent field is zero holds
  exactly the compensated M2L sum, so each of its lanes is the exact sum of
  the f32 pair terms within that bound; the compensated solve and
  accelerations agree with the plain ones and sit no further from the f64
  twin.

Each ``check_*`` raises AssertionError on a mismatch; the mutation tests run
them against oracle builds with planted mistakes.
"""

import math

import numpy as np

import oracle as orc
import rainbow_fmm as rf

U = 2.0 ** -24


def neumaier_bound(terms, exact):
    terms = np.asarray(terms, dtype=np.float64)
    return U * abs(exact) + 2 * len(terms) * U * U * float(np.sum(np.abs(terms)))


def plain_f32(terms, init=0.0):
    s = np.float32(init)
    for t in np.asarray(terms, dtype=np.float32):
        s = np.float32(s + t)
    return s


def check_known_answers():
    # 1 + 1e8 - 1e8: the large term meets a small sum (|t| > |s| branch).
    assert orc.comp_sum_f32([1.0, 1e8, -1e8]) == 1.0
    # 1e8 + 1 - 1e8: the small term meets a large sum (|s| >= |t| branch).
    assert orc.comp_sum_f32([1e8, 1.0, -1e8]) == 1.0
    # Eight lost units must all be carried.
    assert orc.comp_sum_f32([1e8] + [1.0] * 8 + [-1e8]) == 8.0
    # The running value a field already holds is where the sum starts.
    assert orc.comp_sum_f32([1e8, -1e8], init=3.0) == 3.0
    assert orc.comp_sum_f32([], init=-2.5) == -2.5
    # Plain f32 loses all of these (the inputs really are ill-conditioned).
    assert plain_f32([1.0, 1e8, -1e8]) == 0.0
    assert plain_f32([1e8] + [1.0] * 8 + [-1e8]) == 0.0


def check_random_sums_within_bound():
    rng = np.random.default_rng(2024)
    plain_misses = 0
    for trial in range(300):
        n = int(rng.integers(2, 190))
        mags = 10.0 ** rng.uniform(-4, 4, n)
        terms = (rng.choice([-1.0, 1.0], n) * mags).astype(np.float32)
        if trial % 2:
            # cancelling pairs: the exact sum is small against sum |x|
            k = n // 2
            terms[k:2 * k] = -terms[:k] * np.float32(1.0 + 1e-3 * rng.standard_normal())
        exact = math.fsum(float(t) for t in terms)
        got = float(orc.comp_sum_f32(terms))
        bound = neumaier_bound(terms, exact)
        assert abs(got - exact) <= bound, (trial, got, exact, bound)
        if abs(float(plain_f32(terms)) - exact) > 10 * bound:
            plain_misses += 1
    # The sample is hard enough that plain f32 often misses the bound by 10x.
    assert plain_misses >= 100


def _level2_cells(moments):
    """Level-2 cells (global ids 9..72) with more than 16 bodies whose
    level-1 parent also holds more than 16: their field is the M2L sum alone
    (level-1 fields are zero: the M2L list of a level-1 cell is empty and a
    level-1 cell with more than 16 bodies takes no near field)."""
    out = []
    for g in range(9, 73):
        i = g - 9
        x, y, z = i % 4, (i // 4) % 4, i // 16
        parent = 1 + (x // 2) + 2 * ((y // 2) + 2 * (z // 2))
        if moments[g, 0] > 16 and moments[parent, 0] > 16:
            out.append((g, x, y, z))
    return out


def check_solve_lanes_are_exact_sums():
    n = 65536
    parts = rf.ic_plummer(n, 51, 0.3, 1.5, 1.0)
    slots = orc.build_plans(n, gravity=1.0, softening=1e-3)
    m, f = orc.solve_fields(slots, parts, compensated=True)
    slot0 = orc.slot(slots, 0)
    cells = _level2_cells(m)
    assert len(cells) >= 8
    half = np.float32(1.6)
    eps = np.float32(np.float32(np.float32(0.06) * half) / np.float32(4))
    for g, x, y, z in cells:
        centre = np.array([((np.float32(c) + np.float32(0.5)) * np.float32(0.5) - np.float32(1.0)) * half
                           for c in (x, y, z)], dtype=np.float32)
        terms = [orc.pair_field(slot0, m[s], centre, float(eps)) for s in orc.m2l_sources(g)
                 if m[s, 0] != 0]
        terms = np.array(terms, dtype=np.float32)
        for lane in range(10):
            col = terms[:, lane]
            exact = math.fsum(float(t) for t in col)
            assert abs(float(f[g, lane]) - exact) <= neumaier_bound(col, exact) + 1e-45, (g, lane)


def check_compensated_solve_matches_plain_and_f64():
    n = 65536
    parts = rf.ic_plummer(n, 52, 0.1, 1.5, 1.0)
    slots = orc.build_plans(n, gravity=1.0, softening=1e-3)
    mc, fc = orc.solve_fields(slots, parts, compensated=True)
    mp, fp = orc.solve_fields(slots, parts)
    m64, f64 = orc.solve_fields_f64(slots, parts)
    np.testing.assert_array_equal(mc, mp)
    norm = np.linalg.norm(f64[:, :3], axis=1)
    active = norm > 0
    floor = 1e-3 * np.sqrt(np.mean(norm[active] ** 2))
    err_c = np.linalg.norm(fc[:, :3] - f64[:, :3], axis=1)[active] / np.maximum(norm[active], floor)
    err_p = np.linalg.norm(fp[:, :3] - f64[:, :3], axis=1)[active] / np.maximum(norm[active], floor)
    assert err_c.max() < 2e-5 and np.median(err_c) <= np.median(err_p)
    assert np.mean(err_c) <= np.mean(err_p)
    # potentials and tidal lanes too
    for lanes in (slice(3, 4), slice(4, 10)):
        ref = np.abs(f64[active, lanes]).max()
        assert np.abs(fc[active, lanes] - f64[active, lanes]).max() <= 2e-5 * ref
    # per-body accelerations: the compensated own-leaf and expansion sums
    ac = orc.accelerations(slots, parts, mc, fc, compensated=True)[:, :3].astype(np.float64)
    ap = orc.accelerations(slots, parts, mp, fp)[:, :3].astype(np.float64)
    a64 = orc.accelerations_f64(slots, parts, m64, f64)[:, :3]
    nb = np.linalg.norm(a64, axis=1)
    fl = 1e-3 * np.sqrt(np.mean(nb ** 2))
    rc = np.linalg.norm(ac - a64, axis=1) / np.maximum(nb, fl)
    rp = np.linalg.norm(ap - a64, axis=1) / np.maximum(nb, fl)
    assert rc.max() < 1e-4 and np.mean(rc) <= np.mean(rp)


def test_known_answers():
    check_known_answers()


def test_random_sums_within_bound():
    check_random_sums_within_bound()


def test_solve_lanes_are_exact_sums():
    check_solve_lanes_are_exact_sums()


def test_compensated_solve_matches_plain_and_f64():
    check_compensated_solve_matches_plain_and_f64()
