"""The oracle's excitation and hue lanes and the control ("shepherd") branch
of integrate (src/rainbow_fmm.wgsl:659-768), pinned without retyping their
formulas: each case switches on one input of the target excitation at a
time (rest, pure radial motion either way, a 60-degree angle, a leaf of 16,
52 or 100 bodies, a saturated field, a body at the control point, all of
them at once), so the target is a known constant or a known ratio of two
others (the radial shock is the cube of the radial fraction: 1/8 at 60
degrees).  The relaxation toward a fixed target is geometric with the
heating or cooling rate, the hue drift follows the WGSL smoothstep's
textbook Hermite values and the signed radial motion, the control field is
a Gaussian in distance (measured through the excitation, against exp in
f64), attract and repel are exact negatives, the swirl is orthogonal to
both the axis and the offset with the attract magnitude times
sqrt(d^2 + c^2) / d, and the radius clamp, axis scaling, far field, on-axis
and disabled cases are exact identities.

With gravity 0 the tree's fields vanish, so a body's acceleration is the
control field alone.  Planted mistakes in the oracle (squared instead of
cubed shock, swapped heating/cooling rates, unsigned radial motion in the
hue, a linear hue ramp, exp(-d^2/R^2), a core of 0.18 R^2, repel with the
attract sign, the cross product reversed, the radius clamp dropped, the
leaf capacity 8 in the density pulse, a wrong field or control share) each
fail one of these.
"""

import numpy as np

import oracle as orc

BASE = 0.025  # the target excitation of a body at rest, alone, in no field


def body(pos, vel=(0.0, 0.0, 0.0), hue=0.0, exc=0.0):
    return [*pos, hue, *vel, exc]


def run(rows, dt, gravity=0.0, control=None, steps=1):
    parts = np.array(rows, dtype=np.float32)
    slots = orc.build_plans(parts.shape[0], dt=dt, gravity=gravity, softening=1e-3, control=control)
    p, _, _ = orc.step(slots, parts, steps)
    return p


def clump(centre, k, seed, spread=1e-3):
    rng = np.random.default_rng(seed)
    pts = np.asarray(centre) + rng.uniform(-spread, spread, (k, 3))
    return [body(x) for x in pts]


def control(point, mode, strength, radius, axis=(0.0, 0.0, 1.0), enabled=1.0):
    return ((*point, enabled), (*axis, mode), strength, radius)


# dt = 0.1: a heating step's weight dt * 10 rounds to 1 in f32, so the new
# excitation is the target itself (mix(e, target, 1) = target).
DT_FULL = 0.1


def test_target_excitation_terms_one_at_a_time():
    rows = [
        body((-0.5, -0.5, -0.5)),                                   # 0 rest
        body((-0.5, 0.0, 0.0), (-0.25, 0.0, 0.0)),                  # 1 radial out
        body((0.0, -0.5, 0.0), (0.0, 0.25, 0.0)),                   # 2 radial in
        body((0.0, 0.0, -0.5), (0.25 * np.sin(np.pi / 3), 0.0, -0.25 * np.cos(np.pi / 3))),  # 3 at 60 deg
    ]
    p = run(rows, DT_FULL)
    e = p[:, 7].astype(np.float64)
    assert p[0, 7] == np.float32(BASE)
    assert abs(e[1] - (BASE + 0.56)) < 1e-7
    assert p[2, 7] == p[1, 7]  # the shock takes |radial fraction|
    assert abs((e[3] - BASE) / (e[1] - BASE) - 0.125) < 1e-5


def test_density_pulse_of_the_leaf():
    # a clump inside one finest cell at every depth (all cells [0, w) hold
    # 0.025): the leaf is that cell and its count drives the pulse
    for k, pulse in ((16, 0.0), (52, 0.5), (100, 1.0)):
        rows = clump((0.025, 0.025, 0.025), k, seed=k) + [body((-0.6, -0.6, -0.6))]
        p = run(rows, DT_FULL)
        lone = p[k, 7]
        assert lone == np.float32(BASE), k
        got = p[:k, 7].astype(np.float64)
        assert np.abs(got - (BASE + 0.26 * pulse)).max() < 1e-7, k


def test_saturated_field_adds_its_share():
    # a lone body at rest 1.4 from a clump of 100 with gravity 1e4: the
    # field-strength term saturates at 1
    rows = clump((0.4, 0.4, 0.4), 100, seed=5) + [body((-0.4, -0.4, -0.4))]
    p = run(rows, DT_FULL, gravity=1.0e4)
    assert abs(float(p[100, 7]) - (BASE + 0.13)) < 1e-7
    weak = run(rows, DT_FULL, gravity=1.0e-9)
    assert abs(float(weak[100, 7]) - BASE) < 1e-7


def test_all_terms_clamp_at_one():
    # radial motion (0.56) + the control point on the body (0.42) + base
    rows = [body((0.5, 0.0, 0.0), (0.25, 0.0, 0.0)), body((-0.5, -0.5, -0.5))]
    p = run(rows, DT_FULL, control=control((0.5, 0.0, 0.0), 1.0, 1e-6, 0.3))
    assert p[0, 7] == np.float32(1.0)


def test_relaxation_is_geometric_with_the_heating_and_cooling_rates():
    # radial movers on the x axis keep a radial fraction of exactly 1, so
    # the target stays fixed while they drift
    dt, steps = 0.01, 20
    rows = [body((0.3, 0.0, 0.0), (0.01, 0.0, 0.0), exc=0.0),
            body((-0.3, 0.0, 0.0), (-0.01, 0.0, 0.0), exc=1.0)]
    target = run(rows, DT_FULL)[0, 7].astype(np.float64)
    p = run(rows, dt, steps=steps)
    heat = target - (1.0 - dt * 10.0) ** steps * target
    cool = target + (1.0 - dt * 1.05) ** steps * (1.0 - target)
    assert abs(p[0, 7] - heat) < 2e-6
    assert abs(p[1, 7] - cool) < 2e-6


def smoothstep(lo, hi, x):
    """WGSL smoothstep: the cubic Hermite t^2 (3 - 2t) of the clamped ramp."""
    t = np.clip((x - lo) / (hi - lo), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def test_hue_drift_follows_the_signed_radial_motion():
    # equal excitation (|radial fraction| = 1 both ways), opposite sign
    dt = 0.02
    rows = [body((0.5, 0.0, 0.0), (0.25, 0.0, 0.0), hue=0.5), body((0.0, 0.5, 0.0), (0.0, -0.25, 0.0), hue=0.5)]
    p = run(rows, dt)
    e1 = float(p[0, 7])
    assert p[1, 7] == p[0, 7] and e1 < 0.3
    out, inward = float(p[0, 3]) - 0.5, float(p[1, 3]) - 0.5
    assert abs((out - inward) - 2 * 0.030 * dt) < 1e-7
    assert abs((out + inward) / 2 - dt * (0.012 + 0.036 * e1)) < 1e-7


def test_hue_freezes_along_the_smoothstep():
    # bodies at rest: the cooled excitation sets the drift factor
    dt = 0.01
    e0 = np.array([0.05, 0.25, 0.35, 0.42, 0.5, 0.58, 0.65, 0.75, 0.95])
    pts = [(-0.6 + 0.15 * i, 0.37, -0.41) for i in range(e0.size)]
    rows = [body(x, exc=e) for x, e in zip(pts, e0)]
    p = run(rows, dt)
    e1 = p[:, 7].astype(np.float64)
    assert np.all(e1 < e0)  # cooling toward the base
    factor = p[:, 3].astype(np.float64) / (dt * (0.012 + 0.036 * e1))
    np.testing.assert_allclose(factor, 1.0 - smoothstep(0.30, 0.70, e1), atol=2e-5)
    assert factor[0] > 0.99999 and factor[-1] == 0.0


def test_hue_wraps_into_the_unit_interval():
    p = run([body((0.5, 0.0, 0.0), (0.25, 0.0, 0.0), hue=0.9999)], 0.05)
    h = float(p[0, 3])
    assert 0.0 <= h < 0.01
    assert abs((h + 1.0 - 0.9999) - 0.05 * (0.012 + 0.030 + 0.036 * float(p[0, 7]))) < 1e-6


POINT = (0.2, -0.1, 0.3)
R = 0.25
OFFSETS = np.array([(0.05, 0.0, 0.0), (0.0, -0.1, 0.07), (0.12, 0.09, -0.05),
                    (-0.2, 0.1, 0.15), (0.3, -0.25, 0.1)])


def control_rows():
    return [body(np.array(POINT) + o) for o in OFFSETS] + [body(POINT), body((-1.2, -1.2, -1.2))]


def test_control_influence_is_a_gaussian_of_the_distance():
    # strength 1e-7: the control's own field term stays below 1e-7
    p = run(control_rows(), DT_FULL, control=control(POINT, 1.0, 1e-7, R))
    infl = (p[:, 7].astype(np.float64) - BASE) / 0.42
    d2 = np.sum(OFFSETS.astype(np.float32).astype(np.float64) ** 2, axis=1)
    np.testing.assert_allclose(infl[:5], np.exp(-2.0 * d2 / R ** 2), atol=5e-7)
    assert abs(infl[5] - 1.0) < 5e-7          # on the point: exp(0)
    assert p[6, 7] == np.float32(BASE)        # far field: the Gaussian underflows


def kicks(mode, axis=(0.0, 0.0, 1.0), radius=R, enabled=1.0, strength=2.0):
    rows = control_rows()
    p = run(rows, 0.01, control=control(POINT, mode, strength, radius, axis, enabled))
    return p, p[:, 4:7].astype(np.float64) / 0.01


def test_attract_points_at_the_control_point_and_repel_is_its_negative():
    pa, a = kicks(1.0)
    pr, r = kicks(2.0)
    delta = -OFFSETS  # control point minus body
    cos = np.sum(a[:5] * delta, axis=1) / (np.linalg.norm(a[:5], axis=1) * np.linalg.norm(delta, axis=1))
    assert np.all(cos > 1 - 1e-6)
    np.testing.assert_array_equal(pr[:, 4:7], -pa[:, 4:7])
    # on the point and in the far field the body drifts freely
    assert np.all(a[5:] == 0.0) and np.all(pa[5:, :3] == np.array(control_rows(), np.float32)[5:, :3])
    # mode thresholds: below 1.5 attract, below 2.5 repel
    np.testing.assert_array_equal(kicks(1.49)[0], pa)
    np.testing.assert_array_equal(kicks(2.49)[0], pr)


def test_swirl_is_orthogonal_with_the_core_ratio():
    _, a = kicks(1.0)
    ps, s = kicks(3.0, axis=(0.0, 0.0, 2.0))
    delta = -OFFSETS
    assert np.all(s[:5, 2] == 0.0)  # perpendicular to the z axis
    assert np.all(np.abs(np.sum(s[:5] * delta, axis=1)) < 1e-6 * np.linalg.norm(s[:5], axis=1))
    # right-handed: z x delta = (-dy, dx, 0)
    zxd = np.stack([-delta[:, 1], delta[:, 0]], axis=1)
    assert np.all(np.sum(s[:5, :2] * zxd, axis=1) > 0)
    # |swirl| / |attract| = sqrt(d^2 + 0.018 R^2) / d (unit tangent against
    # the cored direction)
    d = np.linalg.norm(OFFSETS, axis=1)
    ratio = np.linalg.norm(s[:5], axis=1) / np.linalg.norm(a[:5], axis=1)
    np.testing.assert_allclose(ratio, np.sqrt(d ** 2 + 0.018 * R ** 2) / d, rtol=2e-6)
    # the axis is normalised: scaling it changes nothing
    np.testing.assert_array_equal(kicks(3.0, axis=(0.0, 0.0, 1.0))[0], ps)


def test_swirl_on_the_axis_gives_no_kick():
    rows = [body((POINT[0], POINT[1], POINT[2] + 0.1)), body((-1.0, -1.0, -1.0))]
    p = run(rows, 0.01, control=control(POINT, 3.0, 2.0, R))
    assert np.all(p[0, 4:7] == 0.0)


def test_control_radius_clamp_and_switch():
    small = kicks(1.0, radius=0.001)[0]
    np.testing.assert_array_equal(small, kicks(1.0, radius=0.02)[0])
    np.testing.assert_array_equal(kicks(1.0, radius=0.0)[0], small)
    off = kicks(1.0, enabled=0.0)[0]
    np.testing.assert_array_equal(off, run(control_rows(), 0.01))
