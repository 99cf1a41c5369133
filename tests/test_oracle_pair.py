"""pair_field (src/rainbow_fmm.wgsl:170-203) of the oracle against the
softened-monopole closed form in f64, its exact symmetries, and the
derivative relations between its lanes (a = -grad phi, J = d a / d target)."""

import numpy as np
import pytest

import oracle as orc


def slot0(n, gravity):
    return orc.slot(orc.build_plans(n, gravity=gravity), 0)


def closed_form(com, target, eps, gm):
    r = np.asarray(com, np.float64) - np.asarray(target, np.float64)
    s2 = r @ r + eps * eps
    s = np.sqrt(s2)
    a = gm * r / s ** 3
    phi = -gm / s
    J = -gm / s ** 3 * np.eye(3) + 3 * gm * np.outer(r, r) / s ** 5
    return a, phi, J


def unpack(f):
    a = f[0:3].astype(np.float64)
    phi = float(f[3])
    J = np.array([[f[4], f[5], f[6]], [f[5], f[7], f[8]], [f[6], f[8], f[9]]], dtype=np.float64)
    return a, phi, J


@pytest.mark.parametrize("seed", range(6))
def test_pair_field_matches_closed_form(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(16384, 1 << 22))
    g = float(rng.uniform(0.1, 3.0))
    slot = slot0(n, g)
    scale = orc.plan(n)[2]
    for _ in range(50):
        count = int(rng.integers(1, 5000))
        com = rng.uniform(-1.5, 1.5, 3)
        sums = np.rint(com * scale * count).astype(np.int64)
        sums = np.clip(sums, -2**31, 2**31 - 1)
        moment = np.array([count, *sums], dtype=np.int64).astype(np.int32)
        target = rng.uniform(-1.5, 1.5, 3).astype(np.float32)
        eps = float(np.float32(rng.choice([1e-3, 0.024, 0.3])))
        f = orc.pair_field(slot, moment, target, eps)
        com64 = sums.astype(np.float64) / (scale * count)
        gm = np.float64(np.float32(g)) * count / n
        a, phi, J = closed_form(com64, target, eps, gm)
        fa, fphi, fJ = unpack(f)
        assert np.linalg.norm(fa - a) <= 5e-6 * np.linalg.norm(a)
        assert abs(fphi - phi) <= 5e-6 * abs(phi)
        assert np.linalg.norm(fJ - J) <= 5e-6 * np.linalg.norm(J)
        assert f[10] == 0.0 and f[11] == 0.0
        # trace(J) = -3 G m eps^2 / (r^2 + eps^2)^(5/2)
        r = com64 - target
        tr = -3 * gm * eps * eps / (r @ r + eps * eps) ** 2.5
        assert abs(np.trace(fJ) - tr) <= 1e-5 * np.linalg.norm(J)


def test_empty_source_contributes_nothing():
    f = orc.pair_field(slot0(65536, 1.0), np.array([0, 123, 4, -5], np.int32), np.zeros(3, np.float32), 0.01)
    assert not f.any()


def test_antisymmetry_for_equal_masses():
    n = 65536
    slot = slot0(n, 1.0)
    scale = orc.plan(n)[2]
    # Single bodies at exactly representable fixed-point positions.
    xa = np.array([3000, -1200, 777], np.int64)
    xb = np.array([-9000, 4096, 15], np.int64)
    ma = np.array([1, *xa], np.int32)
    mb = np.array([1, *xb], np.int32)
    ta = (xa / scale).astype(np.float32)
    tb = (xb / scale).astype(np.float32)
    fab = orc.pair_field(slot, ma, tb, 0.01)  # source a at target b
    fba = orc.pair_field(slot, mb, ta, 0.01)
    np.testing.assert_array_equal(fab[0:3], -fba[0:3])
    np.testing.assert_array_equal(fab[3:], fba[3:])


def test_mass_scales_exactly():
    n = 65536
    slot = slot0(n, 1.0)
    m1 = np.array([3, 3000, -1200, 777], np.int32)
    m2 = np.array([6, 6000, -2400, 1554], np.int32)
    t = np.array([0.1, 0.2, -0.3], np.float32)
    f1 = orc.pair_field(slot, m1, t, 0.02)
    f2 = orc.pair_field(slot, m2, t, 0.02)
    np.testing.assert_array_equal(f2, 2 * f1)


@pytest.mark.parametrize("seed", range(3))
def test_field_is_the_gradient_of_the_potential(seed):
    # central differences of the oracle's own phi and a (target shifted by
    # h = 2^-10, exact in f32): a = -grad phi and J = d a / d target, the
    # Jacobian integrate applies to a body's offset from the cell centre
    rng = np.random.default_rng(100 + seed)
    n = 1 << 20
    slot = slot0(n, 1.0)
    scale = orc.plan(n)[2]
    h = 2.0 ** -10
    for _ in range(20):
        count = int(rng.integers(1, 4000))
        com = rng.uniform(-1.0, 1.0, 3)
        moment = np.array([count, *np.rint(com * scale * count)], dtype=np.int64).astype(np.int32)
        while True:
            target = rng.uniform(-1.2, 1.2, 3).astype(np.float32)
            if np.linalg.norm(com - target) > 0.3:
                break
        eps = float(np.float32(rng.choice([1e-3, 0.05])))
        a, _, J = unpack(orc.pair_field(slot, moment, target, eps))
        grad = np.zeros(3)
        jac = np.zeros((3, 3))
        for k in range(3):
            e = np.zeros(3, np.float32)
            e[k] = h
            fp = orc.pair_field(slot, moment, target + e, eps)
            fm = orc.pair_field(slot, moment, target - e, eps)
            grad[k] = (float(fp[3]) - float(fm[3])) / (2 * h)
            jac[:, k] = (unpack(fp)[0] - unpack(fm)[0]) / (2 * h)
        assert np.linalg.norm(a + grad) <= 2e-3 * np.linalg.norm(a)
        assert np.linalg.norm(J - jac) <= 2e-3 * np.linalg.norm(J)
