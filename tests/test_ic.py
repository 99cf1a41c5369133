"""Seeded initial conditions: deterministic, and statistically what they
claim (uniform ball, Plummer sphere in virial equilibrium, clumps)."""

import numpy as np

import oracle as orc
import rainbow_fmm as rf


def test_deterministic_and_seeded():
    a = rf.ic_plummer(10000, 5, 0.1, 1.5, 1.0)
    b = rf.ic_plummer(10000, 5, 0.1, 1.5, 1.0)
    c = rf.ic_plummer(10000, 6, 0.1, 1.5, 1.0)
    assert a.tobytes() == b.tobytes()
    assert a.tobytes() != c.tobytes()
    assert a.dtype == np.float32 and a.shape == (10000, 8)


def test_uniform_ball():
    p = rf.ic_uniform_ball(200000, 1, 1.0)
    r = np.linalg.norm(p[:, :3].astype(np.float64), axis=1)
    assert r.max() <= 1.0 + 1e-6
    # uniform in volume: P(r < x) = x^3
    for x in (0.3, 0.5, 0.8):
        assert abs(np.mean(r < x) - x ** 3) < 0.005
    assert not p[:, 4:7].any()
    assert np.all((p[:, 3] >= 0) & (p[:, 3] < 1))


def test_plummer_profile_and_virial():
    n = 100000
    a = 0.1
    p = rf.ic_plummer(n, 2, a, 1.5, 1.0)
    x = p[:, :3].astype(np.float64)
    v = p[:, 4:7].astype(np.float64)
    r = np.linalg.norm(x, axis=1)
    assert r.max() < 1.5 + 1e-3
    # half-mass radius of a Plummer sphere: a / sqrt(2^(2/3) - 1) = 1.305 a
    assert abs(np.median(r) / (a / np.sqrt(2 ** (2 / 3) - 1)) - 1) < 0.03
    assert np.abs(x.mean(0)).max() < 1e-6 and np.abs(v.mean(0)).max() < 1e-6
    # virial ratio 2K / |W| = 1; W from an f64 direct sum on a sample
    m = 1.0 / n
    ke = 0.5 * m * np.sum(v * v)
    rng = np.random.default_rng(0)
    sample = rng.choice(n, 2000, replace=False).astype(np.uint32)
    phi = orc.direct_sum_f64(p, 1.0, 1e-4, sample)[:, 3]
    w = 0.5 * n * m * phi.mean()
    assert abs(2 * ke / abs(w) - 1) < 0.05


def test_clumps():
    n = 300000
    p = rf.ic_clumps(n, 3, 64, 0.02, 10.0, 1.0)
    assert p.shape == (n, 8)
    x = p[:, :3]
    assert np.abs(x).max() < 0.5 + 0.2 + 1e-3
    assert np.isfinite(p).all()


def test_invalid_arguments_are_rejected_with_a_message():
    import pytest

    cases = [
        (lambda: rf.ic_uniform_ball(100, 1, 0.0), "uniform ball"),
        (lambda: rf.ic_plummer(100, 1, -0.1, 1.5, 1.0), "Plummer sphere"),
        (lambda: rf.ic_plummer(100, 1, 0.1, 0.0, 1.0), "Plummer sphere"),
        (lambda: rf.ic_clumps(9, 1, 32, 0.02, 10.0, 1.0), "clumps"),
        (lambda: rf.ic_clumps(100, 1, 0, 0.02, 10.0, 1.0), "clumps"),
    ]
    for call, what in cases:
        with pytest.raises(rf.RainbowError) as e:
            call()
        assert e.value.code == rf.E_INVALID and what in str(e.value)
    # A clump per body is allowed.
    assert rf.ic_clumps(9, 1, 9, 0.02, 10.0, 1.0).shape == (9, 8)
