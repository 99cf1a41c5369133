"""The reference's default scene (initialize, src/rainbow_fmm.wgsl:205-374).

CPU: the oracle's restatement is pinned by properties the reference fixes
without retyping its formulas -- hash32 inverts exactly (xorshift-16 and the
odd multiplier undone in reverse), the enclosed-mass model's closed values
(all mass inside r = 1, the 1e-4 floor at the centre, monotone), the
population shares against the mass model's component weights, which draw
drives which coordinate of a body, and the geometry of each population
(radii, planes, arm phases, circular speeds, orthogonality, excitation
ranges).  Planted mistakes in the restatement (a tilt's sign, the arm
spacing, a speed factor, the lanes' sense, a hash shift, a power law, a
swapped draw or sin/cos) each fail one of these.

GPU: the device seeding against the oracle.  The kernel and the oracle use
different sin/cos/pow/atan2 (device library against host libm; the WGSL
leaves their accuracy to the implementation), so positions and velocities
agree within a stated tolerance, the hash-derived excitations bit for bit,
and a step from the seeded state is bit-exact again; the harness CLI's
digests of the seeded scene equal the oracle's.
"""

import sys

import numpy as np
import pytest

import oracle as orc

TAU = 6.28318530718
G = 0.34  # the harness's gravity (native/main.rs)


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def populations(n):
    """Population of each body from its u5 draw (bulge, halo, disc, polar)."""
    u5 = np.array([orc.hash32(i * 11 + 6) for i in range(n)], dtype=np.float64)
    u5 = (u5.astype(np.float32) * np.float32(1.0 / 4294967296.0)).astype(np.float64)
    return u5 < 0.05, (u5 >= 0.05) & (u5 < 0.75), (u5 >= 0.75) & (u5 < 0.93), u5 >= 0.93


@pytest.fixture(scope="module")
def scene():
    n = 65536
    return n, orc.seed_galaxy(n, G).astype(np.float64), populations(n)


def test_hash32_inverts():
    inv = pow(0x45D9F3B, -1, 1 << 32)

    def xs(x):
        return x ^ (x >> 16)

    def unhash(y):
        return xs((xs((xs(y) * inv) & 0xFFFFFFFF) * inv) & 0xFFFFFFFF)

    rng = np.random.default_rng(0)
    for x in [0, 1, 2, 11, 0xFFFFFFFF, 0x80000000] + [int(v) for v in rng.integers(0, 1 << 32, 2000)]:
        assert unhash(orc.hash32(x)) == x
    assert orc.hash32(0) == 0
    # Bijective: distinct outputs on a run of consecutive seeds.
    outs = {orc.hash32(i * 11 + 3) for i in range(20000)}
    assert len(outs) == 20000


def test_enclosed_mass_closed_values():
    assert orc.enclosed_mass(0.0) == np.float32(0.0001)
    # Every population lies inside r = 1: the fractions sum to one.
    assert abs(orc.enclosed_mass(1.0) - 1.0) < 1e-6
    assert abs(orc.enclosed_mass(5.0) - 1.0) < 1e-6
    rs = np.linspace(0.0, 1.2, 400)
    m = np.array([orc.enclosed_mass(r) for r in rs])
    assert np.all(np.diff(m) >= 0)
    # Only the bulge and halo (uniform balls) below r = 0.055: r^3 laws.
    r = 0.05
    want = 0.05 * (r / 0.11) ** 3 + 0.70 * (r / 0.70) ** 3
    assert abs(orc.enclosed_mass(r) - want) < 1e-6


def draws(n):
    """The six uniform draws u0..u5 of each body (random01 of i * 11 + 1..6)."""
    h = np.array([[orc.hash32(i * 11 + 1 + k) for k in range(6)] for i in range(n)], dtype=np.uint64)
    return h.astype(np.float32) * np.float32(1.0 / 4294967296.0)


def test_draw_mapping(scene):
    # Which draw drives which coordinate, per body: the halo's direction is
    # the uniform-sphere map (y = 2 u0 - 1 the polar axis, azimuth tau u1 from
    # x towards z) and its radius R u2^(1/3); the disc's radius follows u0
    # and its excitation u3; the polar lanes' excitation u4.
    n, p, (bulge, halo, disc, polar) = scene
    u = draws(n)
    ball = bulge | halo
    pos = p[:, :3]
    r = np.linalg.norm(pos, axis=1)
    np.testing.assert_allclose(pos[ball, 1] / r[ball], 2.0 * u[ball, 0] - 1.0, atol=2e-6)
    az = np.arctan2(pos[ball, 2], pos[ball, 0])
    ring = np.sqrt(np.maximum(0.0, 1.0 - (2.0 * u[ball, 0] - 1.0) ** 2))
    ok = ring > 1e-3  # the azimuth is defined off the axis
    d = np.angle(np.exp(1j * (az[ok] - TAU * u[ball, 1][ok])))
    assert np.abs(d).max() < 1e-4
    outer = np.where(bulge[ball], 0.11, 0.70)
    np.testing.assert_allclose(r[ball], outer * u[ball, 2].astype(np.float64) ** (1.0 / 3.0), rtol=2e-6)
    f32 = np.float32
    assert np.array_equal(p[halo, 7].astype(f32), f32(0.025) + u[halo, 4] * f32(0.025))
    assert np.array_equal(p[disc, 7].astype(f32), f32(0.08) + u[disc, 3] * f32(0.05))
    assert np.array_equal(p[polar, 7].astype(f32), f32(0.20) + u[polar, 4] * f32(0.08))
    t = 0.20
    fy = p[disc, 1] * np.cos(t) + p[disc, 2] * np.sin(t)
    radius = np.hypot(p[disc, 0], fy / 0.72)
    np.testing.assert_allclose(radius, 0.055 + 0.60 * u[disc, 0].astype(np.float64) ** 0.72, atol=2e-5)


def test_population_shares(scene):
    n, _, (bulge, halo, disc, polar) = scene
    for mask, share in ((bulge, 0.05), (halo, 0.70), (disc, 0.18), (polar, 0.07)):
        sigma = np.sqrt(share * (1 - share) / n)
        assert abs(mask.mean() - share) < 5 * sigma


def test_halo_and_bulge_geometry(scene):
    n, p, (bulge, halo, disc, polar) = scene
    pos, vel = p[:, :3], p[:, 4:7]
    r = np.linalg.norm(pos, axis=1)
    v = np.linalg.norm(vel, axis=1)
    for mask, outer, exc in ((bulge, 0.11, None), (halo, 0.70, (0.025, 0.05))):
        assert r[mask].max() <= outer * (1 + 1e-6)
        # Uniform ball: P(r < x) = (x / R)^3.
        for q in (0.3, 0.6, 0.9):
            frac = np.mean(r[mask] < q * outer)
            sigma = np.sqrt(q ** 3 * (1 - q ** 3) / mask.sum())
            assert abs(frac - q ** 3) < 5 * sigma
        # Circular speed about an axis through the centre.
        cosang = np.einsum("ij,ij->i", pos[mask], vel[mask]) / (r[mask] * v[mask])
        assert np.abs(cosang).max() < 1e-5
        m = np.array([orc.enclosed_mass(x) for x in r[mask].astype(np.float32)])
        want = 0.985 * np.sqrt(G * m / (r[mask] + 0.018))
        np.testing.assert_allclose(v[mask], want, rtol=2e-5)
        assert np.all(p[mask, 3] == 0.0)  # no hue for the halo
        if exc is None:
            assert np.all(p[mask, 7] == np.float32(0.09))
        else:
            assert p[mask, 7].min() >= exc[0] and p[mask, 7].max() < exc[1]


def test_disc_geometry(scene):
    n, p, (_, _, disc, _) = scene
    idx = np.nonzero(disc)[0]
    pos = p[idx, :3]
    t = 0.20
    # Undo the tilt about x.
    fx = pos[:, 0]
    fy = pos[:, 1] * np.cos(t) + pos[:, 2] * np.sin(t)
    fz = -pos[:, 1] * np.sin(t) + pos[:, 2] * np.cos(t)
    radius = np.hypot(fx, fy / 0.72)
    assert radius.min() >= 0.055 - 1e-6 and radius.max() <= 0.655 + 1e-6
    assert np.all(np.abs(fz) <= 1.5 * (0.0035 + radius * 0.010) + 1e-6)
    # Four arms: angle = arm + 7.8 r + a spread of at most (0.045 + 0.055 r) / 2.
    angle = np.arctan2(fy / 0.72, fx)
    arm = (idx & 3) * (TAU / 4)
    off = np.angle(np.exp(1j * (angle - arm - 7.8 * radius)))
    assert np.all(np.abs(off) <= 0.5 * (0.045 + 0.055 * radius) + 1e-4)
    # Nearly circular prograde orbits in the disc plane: angular momentum
    # along the tilted plane's normal (0, -sin t, cos t).
    L = np.cross(pos, p[idx, 4:7])
    nz = -L[:, 1] * np.sin(t) + L[:, 2] * np.cos(t)
    assert np.all(nz > 0)
    assert p[idx, 7].min() >= 0.08 and p[idx, 7].max() < 0.13
    assert p[idx, 3].min() >= 0.0 and p[idx, 3].max() < 1.0


def test_polar_geometry(scene):
    n, p, (_, _, _, polar) = scene
    idx = np.nonzero(polar)[0]
    pos, vel = p[idx, :3], p[idx, 4:7]
    lineage = idx & 3
    t = -0.54 + lineage * 0.36
    # Undo the tilt about z.
    fx = pos[:, 0] * np.cos(t) + pos[:, 1] * np.sin(t)
    fy = -pos[:, 0] * np.sin(t) + pos[:, 1] * np.cos(t)
    radius = np.hypot(fx, pos[:, 2])
    base = radius - (lineage - 1.5) * 0.007
    assert base.min() >= 0.22 - 1e-5 and base.max() <= 0.53 + 1e-5
    assert np.all(np.abs(fy) <= 1.5 * (0.004 + 0.009 * radius) + 1e-6)
    # Retrograde about the lane's normal: the tilted frame's y axis.
    L = np.cross(pos, vel)
    ny = -L[:, 0] * np.sin(t) + L[:, 1] * np.cos(t)
    assert np.all(ny > 0)
    v = np.linalg.norm(vel, axis=1)
    m = np.array([orc.enclosed_mass(x) for x in radius.astype(np.float32)])
    # The speed uses the sampled radius (the planar one before thickness).
    np.testing.assert_allclose(v, 0.99 * np.sqrt(G * m / (radius + 0.018)), rtol=1e-2)
    assert p[idx, 7].min() >= 0.20 and p[idx, 7].max() < 0.28


# Tolerances of the device seeding against the oracle: sin/cos/pow/atan2
# differ by a few ulp between the device library and the host libm; through
# the radii (<= 0.7) and the arm angle (<= 5.2 rad) that is < 1e-5 in
# position, and through the tangent normalisation (|tangent| >= 0.001) and
# the orbital speed (<= 1.2) < 2e-4 in velocity; the hue wraps at 1, so it
# is compared on the circle.
POS_TOL = 1e-5
VEL_TOL = 2e-4
HUE_TOL = 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("n", [65536, 1 << 20])
def test_gpu_seed_matches_oracle(n):
    import rainbow_fmm as rf

    say(f"galaxy seeding on the device, n={n}")
    s = rf.Solver(n, rf.physics(gravity=G))
    s.seed_galaxy()
    got = s.particles().astype(np.float64)
    s.close()
    want = orc.seed_galaxy(n, G).astype(np.float64)
    dpos = np.abs(got[:, :3] - want[:, :3]).max()
    dvel = np.abs(got[:, 4:7] - want[:, 4:7]).max()
    dh = np.abs(got[:, 3] - want[:, 3])
    dhue = np.minimum(dh, 1.0 - dh).max()
    say(f"max |dpos| {dpos:.3g}, |dvel| {dvel:.3g}, hue {dhue:.3g}")
    assert dpos <= POS_TOL
    assert dvel <= VEL_TOL
    assert dhue <= HUE_TOL
    # The excitations come from the hash draws alone: bit for bit.
    assert got[:, 7].astype(np.float32).tobytes() == want[:, 7].astype(np.float32).tobytes()


@pytest.mark.gpu
def test_gpu_seed_then_step_bit_exact():
    import rainbow_fmm as rf
    from test_gpu_parity import assert_bits, slots_for

    n = 65536
    say(f"galaxy seed then 3 steps, n={n}")
    dt, soft = 1e-3, 1e-3
    s = rf.Solver(n, rf.physics(dt=dt, gravity=G, softening=soft))
    s.seed_galaxy()
    seeded = s.particles()
    s.step(3)
    got = s.particles()
    s.close()
    op, _, _ = orc.step(slots_for(n, dt, G, soft), seeded, 3)
    assert_bits(got, op, "particles after 3 steps from the device seed")


@pytest.mark.gpu
def test_gpu_cli_galaxy_report():
    # The harness CLI seeds the galaxy on the device (its default scene, as
    # the reference harness) and prints the content digests of one field
    # solve: they equal the oracle's digests of that solve from the same
    # seed, downloaded through the binding.
    import ctypes
    import re
    import subprocess
    from pathlib import Path

    import rainbow_fmm as rf

    n = 65536
    exe = Path(__file__).resolve().parent.parent / "fmm-rainbow_amd" / "bin" / "rainbow_native"
    say(f"rainbow_native {n} 3 (galaxy)")
    r = subprocess.run([str(exe), str(n), "3"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    assert "scene: galaxy" in r.stdout and "non-finite=0" in r.stdout
    fh, mh = (int(v, 16) for v in re.search(r"fieldsHash=([0-9a-f]{8}) momentsHash=([0-9a-f]{8})", r.stdout).groups())
    slots = orc.build_plans(n)
    depth = orc.depth_of(slots)
    params = rf.Params.from_buffer_copy(bytes(slots[256 * depth:256 * depth + ctypes.sizeof(rf.Params)]))
    s = rf.Solver(n, params)
    s.seed_galaxy()
    seeded = s.particles()
    s.close()
    om, of = orc.solve_fields(slots, seeded)
    assert orc.fnv1a(of) == fh
    assert orc.fnv1a(om) == mh
