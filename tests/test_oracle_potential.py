"""The potential lanes of the oracle and its FP64 reference sums, pinned by
closed forms and derivative identities rather than by retyping them:

* the L2L potential update of downward (src/rainbow_fmm.wgsl:605-610): in the
  two-cluster configuration the only expansion is one M2L term at level 2,
  so the potential the shifts carry to level l is
  phi_0 - a_0 . sum(delta) - sum_{i<j} delta_i . J_0 delta_j
  (a shift adds J delta to a and takes a . delta from phi, so the cross
  terms of earlier shifts appear in later ones);
* the per-body potential (local expansion about the leaf centre plus the
  own-leaf softened monopole) in the same configuration, and its gradient:
  with the moments and fields held fixed, moving one body inside its leaf
  changes its potential by -acc . dx, so central differences of the f64
  potential equal the f64 acceleration;
* the FP64 direct sum: two and three bodies in closed form (softened and
  not), Newton's third law over every body, and a = -grad phi by central
  differences;
* the FP64 kinetic energy, linear and angular momentum of a rigid
  translation and a rigid rotation, whose values are known exactly.

Each ``check_*`` raises AssertionError on a mismatch; the mutation tests run
them against oracle builds with planted mistakes.
"""

import numpy as np

import oracle as orc
import rainbow_fmm as rf

F = np.float32


def centre32(c, dim):
    return F(F(F(F(c) + F(0.5)) * F(F(2.0) / F(dim)) - F(1.0)) * F(1.6))


def two_clusters():
    n = 33
    X = np.array([-1.25, -1.25, -1.25])
    Y = np.array([1.25, 0.75, 1.0])
    parts = np.zeros((n, 8), np.float32)
    parts[:32, :3] = X
    parts[32, :3] = Y
    return n, X, Y, parts


def monopole64(src, target, eps, gm):
    r = np.asarray(src, np.float64) - np.asarray(target, np.float64)
    s2 = r @ r + eps * eps
    a = gm * r / s2 ** 1.5
    phi = -gm / np.sqrt(s2)
    J = -gm / s2 ** 1.5 * np.eye(3) + 3 * gm * np.outer(r, r) / s2 ** 2.5
    return a, phi, J


def check_l2l_potential_telescopes():
    n, X, Y, parts = two_clusters()
    g = 1.0
    slots = orc.build_plans(n, gravity=g, softening=1e-3)
    assert orc.depth_of(slots) == 4
    m, f = orc.solve_fields(slots, parts)
    # X's cells: (0,0,0) at levels 2 and 3, (1,1,1) at level 4.
    c2 = np.array([centre32(0, 4)] * 3, np.float64)
    c3 = np.array([centre32(0, 8)] * 3, np.float64)
    c4 = np.array([centre32(1, 16)] * 3, np.float64)
    eps2 = float(F(F(F(0.06) * F(1.6)) / F(4)))
    a0, phi0, J0 = monopole64(Y, c2, eps2, g / n)
    d3, d4 = c3 - c2, c4 - c3
    g2 = 9
    g3 = 73
    g4 = 585 + 1 + 16 * (1 + 16 * 1)
    assert m[g2, 0] == m[g3, 0] == m[g4, 0] == 32
    tol = 3e-6 * abs(phi0)
    assert abs(f[g2, 3] - phi0) <= tol
    assert abs(f[g3, 3] - (phi0 - a0 @ d3)) <= tol
    assert abs(f[g4, 3] - (phi0 - a0 @ (d3 + d4) - d3 @ J0 @ d4)) <= tol
    # the acceleration lanes along the way
    np.testing.assert_allclose(f[g3, :3], a0 + J0 @ d3, rtol=3e-6)
    np.testing.assert_allclose(f[g4, :3], a0 + J0 @ (d3 + d4), rtol=3e-6)


def check_body_potential_closed_form():
    n, X, Y, parts = two_clusters()
    g = 1.0
    soft = 1e-3
    slots = orc.build_plans(n, gravity=g, softening=soft)
    m, f = orc.solve_fields(slots, parts)
    acc = orc.accelerations(slots, parts, m, f)
    # Body at X: leaf is X's level-4 cell; expansion about its centre plus
    # the own-leaf monopole of the 31 others at distance 0.
    c2 = np.array([centre32(0, 4)] * 3, np.float64)
    c3 = np.array([centre32(0, 8)] * 3, np.float64)
    c4 = np.array([centre32(1, 16)] * 3, np.float64)
    eps2 = float(F(F(F(0.06) * F(1.6)) / F(4)))
    a0, phi0, J0 = monopole64(Y, c2, eps2, g / n)
    d3, d4 = c3 - c2, c4 - c3
    phi_leaf = phi0 - a0 @ (d3 + d4) - d3 @ J0 @ d4
    a_leaf = a0 + J0 @ (d3 + d4)
    d = X - c4
    eps_self = max(soft, float(F(F(F(0.62) * F(1.6)) / F(16))))
    want_x = phi_leaf - a_leaf @ d - 0.5 * d @ J0 @ d - g * (31 / n) / eps_self
    assert abs(acc[0, 3] - want_x) <= 3e-6 * abs(want_x)
    # Body at Y: a level-1 leaf of one body (no own-leaf term) whose field
    # is the near monopole of X's level-1 cell.
    cY = np.array([centre32(1, 2)] * 3, np.float64)
    eps_near = max(soft, float(F(F(F(0.72) * F(1.6)) / F(2))))
    a1, phi1, J1 = monopole64(X, cY, eps_near, g * 32 / n)
    dY = Y - cY
    want_y = phi1 - a1 @ dY - 0.5 * dY @ J1 @ dY
    assert abs(acc[32, 3] - want_y) <= 3e-6 * abs(want_y)


def check_body_potential_gradient():
    n = 65536
    parts = rf.ic_plummer(n, 61, 0.1, 1.5, 1.0)
    slots = orc.build_plans(n, gravity=1.0, softening=1e-3)
    m, f64 = orc.solve_fields_f64(slots, parts)
    rng = np.random.default_rng(7)
    idx = rng.choice(n, 64, replace=False)
    # The oracle loops over every body of the plan: perturb the sampled
    # bodies in place in the full array (the moments and fields stay fixed).
    leaf = orc.leaf_cells(slots, parts, m)[idx]
    acc = orc.accelerations_f64(slots, parts, m, f64)[idx]
    # A small step: the own-leaf term varies on the scale of its softening
    # (about 0.016 at depth 6), where a coarse step's truncation error shows.
    h = 2.0 ** -18
    checked = 0
    for k in range(3):
        up, dn = parts.copy(), parts.copy()
        up[idx, k] = (up[idx, k].astype(np.float64) + h).astype(np.float32)
        dn[idx, k] = (dn[idx, k].astype(np.float64) - h).astype(np.float32)
        same = (orc.leaf_cells(slots, up, m)[idx] == leaf) & (orc.leaf_cells(slots, dn, m)[idx] == leaf)
        pu = orc.accelerations_f64(slots, up, m, f64)[idx, 3]
        pd = orc.accelerations_f64(slots, dn, m, f64)[idx, 3]
        step = up[idx, k].astype(np.float64) - dn[idx, k].astype(np.float64)
        grad = (pu - pd) / step
        scale = np.linalg.norm(acc[:, :3], axis=1)
        err = np.abs(grad + acc[:, k])[same]
        assert np.all(err <= 2e-5 * scale[same] + 1e-7), (err / scale[same]).max()
        checked += int(same.sum())
    assert checked >= 150


def check_direct_sum_closed_forms():
    # Two bodies 1.25 apart (a 3-4-5 triangle scaled), G = 2.
    p = np.zeros((2, 8), np.float32)
    p[1, :3] = (0.75, 1.0, 0.0)
    g = 2.0
    for eps in (0.0, 0.5):
        out = orc.direct_sum_f64(p, g, eps, [0, 1])
        s2 = 1.5625 + eps * eps
        a = g * 0.5 * np.array([0.75, 1.0, 0.0]) / s2 ** 1.5
        np.testing.assert_allclose(out[0, :3], a, rtol=1e-15, atol=0)
        np.testing.assert_allclose(out[1, :3], -a, rtol=1e-15, atol=0)
        assert abs(out[0, 3] - (-g * 0.5 / np.sqrt(s2))) <= 1e-15
        assert out[0, 3] == out[1, 3]
    # Equilateral triangle of side s: each body is pulled to the centroid
    # with G m (x_j + x_k - 2 x_i) / (s^2 + eps^2)^(3/2), phi = -2 G m /
    # sqrt(s^2 + eps^2).
    s = 1.0
    tri = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, np.sqrt(3) / 2, 0.0]], np.float32)
    p3 = np.zeros((3, 8), np.float32)
    p3[:, :3] = tri
    x = p3[:, :3].astype(np.float64)
    s2 = np.sum((x[1] - x[0]) ** 2)
    for eps in (0.0, 0.25):
        out = orc.direct_sum_f64(p3, 1.0, eps, [0, 1, 2])
        for i in range(3):
            j, k = [t for t in range(3) if t != i]
            others = [t for t in (j, k)]
            want = np.zeros(3)
            phi = 0.0
            for t in others:
                r = x[t] - x[i]
                q = r @ r + eps * eps
                want += (1 / 3) * r / q ** 1.5
                phi -= (1 / 3) / np.sqrt(q)
            np.testing.assert_allclose(out[i, :3], want, rtol=1e-14, atol=1e-16)
            assert abs(out[i, 3] - phi) <= 1e-15
        # and the centroid direction of the pull
        cen = x.mean(axis=0)
        for i in range(3):
            d = cen - x[i]
            cosang = out[i, :3] @ d / (np.linalg.norm(out[i, :3]) * np.linalg.norm(d))
            assert cosang > 1 - 1e-12
    del s, s2


def check_direct_sum_third_law_and_gradient():
    rng = np.random.default_rng(12)
    n = 60
    p = np.zeros((n, 8), np.float32)
    p[:, :3] = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    eps = 0.05
    out = orc.direct_sum_f64(p, 1.0, eps, np.arange(n))
    total = out[:, :3].sum(axis=0)
    assert np.abs(total).max() <= 1e-13 * np.abs(out[:, :3]).sum()
    h = 2.0 ** -12
    for i in (0, 17, 59):
        for k in range(3):
            up, dn = p.copy(), p.copy()
            up[i, k] = np.float32(np.float64(up[i, k]) + h)
            dn[i, k] = np.float32(np.float64(dn[i, k]) - h)
            pu = orc.direct_sum_f64(up, 1.0, eps, [i])[0, 3]
            pd = orc.direct_sum_f64(dn, 1.0, eps, [i])[0, 3]
            grad = (pu - pd) / (np.float64(up[i, k]) - np.float64(dn[i, k]))
            assert abs(grad + out[i, k]) <= 1e-5 * np.linalg.norm(out[i, :3])


def check_conserved_rigid_motions():
    # Uniform translation of bodies whose centre of mass is (0.5, 0, 0).
    n = 8
    p = np.zeros((n, 8), np.float32)
    corners = np.array([[sx, sy, sz] for sx in (-0.25, 0.25) for sy in (-0.25, 0.25) for sz in (-0.25, 0.25)])
    p[:, :3] = corners + np.array([0.5, 0.0, 0.0])
    V = np.array([0.25, -0.5, 0.125])
    p[:, 4:7] = V
    c = orc.conserved_f64(p)
    assert c[0] == 0.5 * (V @ V)
    assert c[1] == 0.0
    np.testing.assert_allclose(c[2:5], V, rtol=0, atol=1e-16)
    np.testing.assert_allclose(c[5:8], np.cross([0.5, 0.0, 0.0], V), rtol=0, atol=1e-16)
    # Rigid rotation about z at omega = 2 of a ring of radius 0.5.
    n = 16
    ang = np.arange(n) * (2 * np.pi / n)
    q = np.zeros((n, 8), np.float32)
    q[:, 0] = 0.5 * np.cos(ang)
    q[:, 1] = 0.5 * np.sin(ang)
    x = q[:, 0].astype(np.float64)
    y = q[:, 1].astype(np.float64)
    q[:, 4] = -2.0 * q[:, 1]
    q[:, 5] = 2.0 * q[:, 0]
    c = orc.conserved_f64(q)
    r2 = np.mean(x * x + y * y)
    assert abs(c[0] - 0.5 * 4.0 * r2) <= 1e-14
    assert abs(c[7] - 2.0 * r2) <= 1e-14
    assert abs(c[5]) <= 1e-16 and abs(c[6]) <= 1e-16
    np.testing.assert_allclose(c[2:5], 0.0, atol=1e-15)


def test_l2l_potential_telescopes():
    check_l2l_potential_telescopes()


def test_body_potential_closed_form():
    check_body_potential_closed_form()


def test_body_potential_gradient():
    check_body_potential_gradient()


def test_direct_sum_closed_forms():
    check_direct_sum_closed_forms()


def test_direct_sum_third_law_and_gradient():
    check_direct_sum_third_law_and_gradient()


def test_conserved_rigid_motions():
    check_conserved_rigid_motions()
