"""The oracle's field solve and integrator against what the mathematics
fixes: a two-cluster configuration whose p=1 answer is a closed form, the
own-leaf monopole of a two-body system, exact linearity in G, the f64 twin,
and the kick-drift/boundary/special cases of integrate
(src/rainbow_fmm.wgsl:613-771)."""

import numpy as np
import pytest

import oracle as orc
import rainbow_fmm as rf


def monopole(com, target, eps, gm):
    r = np.asarray(com, np.float64) - np.asarray(target, np.float64)
    s2 = r @ r + eps * eps
    a = gm * r / s2 ** 1.5
    J = -gm / s2 ** 1.5 * np.eye(3) + 3 * gm * np.outer(r, r) / s2 ** 2.5
    return a, J


def centre(c, dim):
    return (np.asarray(c, np.float64) + 0.5) * (2.0 / dim) - 1.0


def test_two_clusters_closed_form():
    # 32 bodies at X, one at Y.  X's cells hold 32 > 16 bodies down to the
    # finest level; Y's level-2 cell is in the M2L list of X's level-2 cell
    # and nothing else interacts, so the p=1 answer at X is the first-order
    # expansion of Y's monopole about X's level-2 centre (the L2L shifts
    # telescope).  Y is a level-1 leaf whose only near source is X's level-1
    # cell.  Fixed-point positions are exact and the cluster's centre of mass
    # is X exactly (32 is a power of two), so its own-leaf term vanishes.
    n = 33
    g = 1.0
    X = np.array([-1.25, -1.25, -1.25])
    Y = np.array([1.25, 0.75, 1.0])
    parts = np.zeros((n, 8), np.float32)
    parts[:32, :3] = X
    parts[32, :3] = Y
    slots = orc.build_plans(n, gravity=g, softening=1e-3)
    depth = orc.depth_of(slots)
    m, f = orc.solve_fields(slots, parts)
    acc = orc.accelerations(slots, parts, m, f)
    cX2 = centre([0, 0, 0], 4) * np.float64(np.float32(1.6))
    eps2 = float(np.float32(np.float32(0.06) * np.float32(1.6)) / np.float32(4))
    a2, J2 = monopole(Y, cX2, eps2, g * 1 / n)
    want_x = a2 + J2 @ (X - cX2)
    np.testing.assert_allclose(acc[0, :3], want_x, rtol=2e-6)
    np.testing.assert_array_equal(acc[:32, :3], np.repeat(acc[:1, :3], 32, axis=0))
    cY1 = centre([1, 1, 1], 2) * np.float64(np.float32(1.6))
    eps_near = max(1e-3, float(np.float32(np.float32(0.72) * np.float32(1.6)) / np.float32(2)))
    a1, J1 = monopole(X, cY1, eps_near, g * 32 / n)
    want_y = a1 + J1 @ (Y - cY1)
    np.testing.assert_allclose(acc[32, :3], want_y, rtol=2e-6)
    assert depth == 4


def test_two_bodies_own_leaf_monopole():
    # Root holds 2 <= 16 bodies: no M2L or near field; both bodies share a
    # level-1 leaf and feel its softened monopole about the centre of mass.
    n = 2
    g = 0.7
    A = np.array([0.25, 0.5, 0.125])
    B = np.array([0.5, 0.25, 0.375])
    parts = np.zeros((n, 8), np.float32)
    parts[0, :3], parts[1, :3] = A, B
    slots = orc.build_plans(n, gravity=g, softening=1e-3)
    m, f = orc.solve_fields(slots, parts)
    assert not f.any()
    acc = orc.accelerations(slots, parts, m, f)
    com = (A + B) / 2
    eps = float(np.float32(np.float32(0.62) * np.float32(1.6)) / np.float32(2))
    for i, P in enumerate((A, B)):
        d = com - P
        want = np.float32(g) * 0.5 * d / (d @ d + eps * eps) ** 1.5
        np.testing.assert_allclose(acc[i, :3], want, rtol=2e-6)


def test_fields_are_exactly_linear_in_g():
    n = 65536
    parts = rf.ic_plummer(n, 41, 0.1, 1.5, 1.0)
    m1, f1 = orc.solve_fields(orc.build_plans(n, gravity=1.0), parts)
    m2, f2 = orc.solve_fields(orc.build_plans(n, gravity=2.0), parts)
    np.testing.assert_array_equal(m1, m2)
    np.testing.assert_array_equal(f2, 2 * f1)


@pytest.mark.parametrize("kind", ["plummer", "ball"])
def test_f32_solve_agrees_with_f64_twin(kind):
    n = 65536
    parts = rf.ic_plummer(n, 42, 0.1, 1.5, 1.0) if kind == "plummer" else rf.ic_uniform_ball(n, 43, 1.0)
    slots = orc.build_plans(n, gravity=1.0, softening=1e-3)
    m32, f32 = orc.solve_fields(slots, parts)
    m64, f64 = orc.solve_fields_f64(slots, parts)
    np.testing.assert_array_equal(m32, m64)
    a32 = f32[:, :3].astype(np.float64)
    a64 = f64[:, :3]
    norm = np.linalg.norm(a64, axis=1)
    active = norm > 0
    floor = 1e-3 * np.sqrt(np.mean(norm[active] ** 2))
    rel = np.linalg.norm(a32 - a64, axis=1)[active] / np.maximum(norm[active], floor)
    assert rel.max() < 2e-5 and np.median(rel) < 1e-6
    # and the per-body accelerations
    acc32 = orc.accelerations(slots, parts, m32, f32)[:, :3].astype(np.float64)
    acc64 = orc.accelerations_f64(slots, parts, m64, f64)[:, :3]
    nb = np.linalg.norm(acc64, axis=1)
    relb = np.linalg.norm(acc32 - acc64, axis=1) / np.maximum(nb, 1e-3 * np.sqrt(np.mean(nb ** 2)))
    assert relb.max() < 1e-4 and np.median(relb) < 1e-6


def _step(parts, dt=1e-3, gravity=1.0, softening=1e-3, speed=None):
    n = parts.shape[0]
    slots = orc.build_plans(n, dt=dt, gravity=gravity, softening=softening, speed=speed)
    p, _, _ = orc.step(slots, parts, 1)
    return p


def test_free_drift_is_exact():
    rng = np.random.default_rng(3)
    n = 5000
    parts = np.zeros((n, 8), np.float32)
    parts[:, :3] = rng.uniform(-1.0, 1.0, (n, 3))
    parts[:, 4:7] = rng.normal(0, 0.5, (n, 3))
    dt = np.float32(1e-3)
    p = _step(parts, dt=float(dt), gravity=0.0)
    np.testing.assert_array_equal(p[:, 4:7], parts[:, 4:7])
    np.testing.assert_array_equal(p[:, :3], (parts[:, :3] + parts[:, 4:7] * dt).astype(np.float32))


def test_absorbing_boundary_and_reflection():
    parts = np.zeros((3, 8), np.float32)
    parts[0, :3] = (1.5, 0.0, 0.0)
    parts[0, 4:7] = (1.0, 0.0, 0.0)   # crosses the edge: damped then reflected
    parts[1, :3] = (0.0, -1.45, 0.0)
    parts[1, 4:7] = (0.0, -0.1, 0.0)  # inside the absorbing band only
    parts[2, :3] = (0.0, 0.0, 0.3)
    dt = np.float32(0.1)
    p = _step(parts, dt=float(dt), gravity=0.0)
    edge = np.float32(1.6) * np.float32(0.985)
    band = np.float32(0.12) * np.float32(1.6)
    # body 0: x = 1.6 -> over = 1.6 - (edge - band); absorb = 1 -> v *= 0.84;
    # then |x| > edge -> x = edge, v *= -0.55
    assert p[0, 0] == edge
    damp = np.float32(np.float32(1.0) - np.float32(1.0) * np.float32(0.16))
    assert p[0, 4] == np.float32(np.float32(1.0) * damp) * np.float32(-0.55)
    x1 = np.float32(np.float32(-1.45) + np.float32(-0.1) * dt)
    over = np.float32(abs(x1) - np.float32(edge - band))
    absorb = min(max(np.float32(over / band), 0), 1)
    assert p[1, 1] == x1
    assert p[1, 5] == np.float32(np.float32(-0.1) * np.float32(np.float32(1.0) - np.float32(absorb) * np.float32(0.16)))
    assert p[2, 2] == np.float32(0.3)


def test_zero_time_step_leaves_the_state_unchanged():
    n = 40000
    parts = rf.ic_plummer(n, 44, 0.1, 1.2, 1.0)
    parts[:, 7] = np.linspace(0, 1, n, endpoint=False, dtype=np.float32)
    p = _step(parts, dt=0.0)
    np.testing.assert_array_equal(p, parts)


def test_hot_bodies_keep_their_hue():
    # excitation >= 0.7 freezes the hue drift (smoothstep(0.3, 0.7, e) = 1)
    n = 20000
    parts = rf.ic_plummer(n, 45, 0.1, 1.2, 1.0)
    parts[:, 7] = 0.95
    p = _step(parts, dt=1e-4)
    hot = p[:, 7] >= 0.7
    assert hot.mean() > 0.99
    np.testing.assert_array_equal(p[hot, 3], parts[hot, 3])
