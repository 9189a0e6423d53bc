"""Initial-condition generators of the C++ layer
(include/grasph2/initial_conditions.hpp) against the oracle's restatement
(oracle/ic_oracle.py) and the distributions the reference's sources sample:
InitGenerator (src/initialConditions/InitGenerator.h), the modifiers of
ps::ParticleSource (src/initialConditions/ParticleSource.h), ps::UniformSphere
and ps::PlummerSphere (src/initialConditions/particleSources/).  Host only:
no GPU involved.

Deterministic parts (modifiers, centre-of-mass correction, desired COM, the
file sources) are compared value by value; the random sources are checked
through their distributions (Kolmogorov-Smirnov at the 1e-3 level, fixed
seeds) because the reference draws its points through mpUtils, which is not
available, so a reference run cannot be reproduced draw by draw."""
import math
import os
import subprocess

import numpy as np
import pytest
from scipy import stats
from scipy.integrate import quad
from scipy.spatial.transform import Rotation

from oracle import ic_oracle as ico

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "ic_generate")

f32 = np.float32

CHAIN = [("move", (0.5, -1.0, 2.0)), ("rotate", (1.0, 2.0, 3.0), 0.7), ("angvel", (0.1, -0.2, 0.3)),
         ("linvel", (1.0, 0.0, -1.0)), ("density", 2.5)]


# ------------------------------------------------------------------ oracle pins
def test_rotation_matrix_matches_rotation_vector():
    for axis, angle in [((1, 2, 3), 0.7), ((0, 0, 1), math.pi / 2), ((-1, 0.5, 0.25), -2.0)]:
        u = np.asarray(axis, dtype=np.float64)
        want = Rotation.from_rotvec(u / np.linalg.norm(u) * angle).as_matrix()
        assert np.allclose(ico.rotation_matrix(axis, angle), want, atol=3e-7)
    # quarter turn about z takes x to y (right-handed)
    R = ico.rotation_matrix((0, 0, 1), math.pi / 2)
    assert np.allclose(R @ np.array([1, 0, 0], dtype=f32), [0, 1, 0], atol=1e-7)


def test_desired_com_is_the_modifiers_applied_to_the_origin():
    pos, vel, _ = ico.apply_ops(np.zeros((1, 3)), np.zeros((1, 3)), np.zeros(1), CHAIN)
    com, vcom = ico.desired_com(CHAIN)
    assert np.array_equal(pos[0], com) and np.array_equal(vel[0], vcom)


def test_plummer_mass_fraction_integrates_the_density():
    a = 0.7
    rho = lambda r: 3.0 / (4 * math.pi * a ** 3) * (1 + r * r / (a * a)) ** -2.5  # M = 1
    for r in (0.1, 0.7, 2.0, 10.0):
        m = quad(lambda s: 4 * math.pi * s * s * rho(s), 0, r)[0]
        assert ico.plummer_mass_fraction(r, a) == pytest.approx(m, rel=1e-9)


def test_speed_fraction_cdf_is_a_beta_law():
    # x^2 (1 - x^2)^3.5 dx with t = x^2 is t^(1/2) (1 - t)^(7/2) dt / 2: x^2 ~ Beta(3/2, 9/2)
    x = np.linspace(0, 1, 41)
    assert np.allclose(ico.plummer_speed_fraction_cdf(x), stats.beta.cdf(x * x, 1.5, 4.5), atol=2e-6)


def test_com_correction_hits_the_target():
    rng = np.random.default_rng(3)
    pos = rng.normal(size=(200, 3)).astype(f32)
    vel = rng.normal(size=(200, 3)).astype(f32)
    m = rng.uniform(0.5, 2, size=200).astype(f32)
    p2, v2 = ico.correct_com(pos, m, vel, (1, 2, 3), (-1, 0, 0.5))
    md = m.astype(np.float64)
    assert np.allclose((md[:, None] * p2).sum(0) / md.sum(), [1, 2, 3], atol=2e-6)
    assert np.allclose((md[:, None] * v2).sum(0) / md.sum(), [-1, 0, 0.5], atol=2e-6)
    # only the requested part moves
    p3, v3 = ico.correct_com(pos, m, vel, (1, 2, 3), (0, 0, 0), fix_vel=False)
    assert np.array_equal(v3, vel) and not np.array_equal(p3, pos)


# ------------------------------------------------------------------ the C++ generators
@pytest.fixture(scope="module")
def out(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.skip("tests/cpp/ic_generate not built (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("ic")
    r = subprocess.run([EXE, str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return d


def rows(d, name):
    a = np.fromfile(os.path.join(str(d), name), dtype="<f4").reshape(-1, 8)
    return a[:, 0:3], a[:, 3], a[:, 4:7], a[:, 7]


def ks_ok(sample, cdf):
    return stats.kstest(np.asarray(sample, dtype=np.float64), cdf).pvalue > 1e-3


def isotropic(u):
    """unit vectors: the mean vanishes, cos(polar) and azimuth are uniform"""
    n = len(u)
    ok = np.all(np.abs(u.mean(0)) < 5 * math.sqrt(1 / (3 * n)))
    ok &= ks_ok(u[:, 2], stats.uniform(-1, 2).cdf)
    ok &= ks_ok(np.arctan2(u[:, 1], u[:, 0]), stats.uniform(-math.pi, 2 * math.pi).cdf)
    return ok


def test_uniform_sphere(out):
    pos, m, vel, rho = rows(out, "uniform.bin")
    assert len(m) == 20000
    assert np.all(m == f32(3.0) / f32(20000)) and np.all(rho == 5.0) and np.all(vel == 0)
    r = np.linalg.norm(pos.astype(np.float64), axis=1)
    assert r.max() <= 2.0 * (1 + 1e-6)
    assert ks_ok(r, lambda x: ico.uniform_ball_cdf(x, 2.0))
    assert isotropic(pos / r[:, None])


def test_plummer_sphere_positions_and_velocities(out):
    pos, m, vel, rho = rows(out, "plummer.bin")
    n, a, cut, M, G = 20000, 1.0, 10.0, 1.0, 1.0
    assert np.all(m == f32(M) / f32(n)) and np.all(rho == 1.0)
    p64, v64 = pos.astype(np.float64), vel.astype(np.float64)
    r = np.linalg.norm(p64, axis=1)
    assert r.max() <= cut * (1 + 1e-6)
    assert ks_ok(r, lambda x: ico.plummer_radius_cdf(x, a, cut))
    assert isotropic(p64 / r[:, None])
    # speed as a fraction of the local escape speed of the untruncated sphere
    v = np.linalg.norm(v64, axis=1)
    x = v / np.sqrt(2 * G * M / np.sqrt(r * r + a * a))
    assert x.max() <= 1 + 1e-5
    assert ks_ok(x, ico.plummer_speed_fraction_cdf)
    assert isotropic(v64 / v[:, None])
    # dynamic equilibrium: virial ratio 2K/|W| near 1 (W from 2e6 random pairs)
    K = 0.5 * (m.astype(np.float64) * v * v).sum()
    rng = np.random.default_rng(0)
    i, j = rng.integers(0, n, size=(2, 2_000_000))
    keep = i != j
    inv = 1.0 / np.linalg.norm(p64[i[keep]] - p64[j[keep]], axis=1)
    W = -0.5 * n * (n - 1) * (M / n) ** 2 * inv.mean()
    assert 0.9 < 2 * K / abs(W) < 1.1


def parse_tsv(path):
    vals = []
    with open(path) as f:
        for line in f:
            if line.strip():
                vals.append([f32(float(t)) for t in line.split("\t") if t.strip()])
    return np.array(vals, dtype=f32)


def test_modifier_chain_and_com_correction(out):
    src = parse_tsv(os.path.join(str(out), "chain_in.tsv"))  # x y z m vx vy vz
    pos0, m0, vel0 = src[:, 0:3], src[:, 3], src[:, 4:7]
    pos1, vel1, rho1 = ico.apply_ops(pos0, vel0, np.zeros(len(m0)), CHAIN)
    com, vcom = ico.desired_com(CHAIN)
    want_pos, want_vel = ico.correct_com(pos1, m0, vel1, com, vcom)

    got_desired = np.loadtxt(os.path.join(str(out), "chain_desired.txt"))
    assert np.allclose(got_desired[:3], com, rtol=1e-6, atol=1e-6)
    assert np.allclose(got_desired[3:], vcom, rtol=1e-6, atol=1e-6)

    pos, m, vel, rho = rows(out, "chain.bin")
    assert np.array_equal(m, m0) and np.all(rho == 2.5)
    assert np.allclose(pos, want_pos, rtol=0, atol=2e-6)
    assert np.allclose(vel, want_vel, rtol=0, atol=2e-6)
    md = m.astype(np.float64)
    assert np.allclose((md[:, None] * pos).sum(0) / md.sum(), com, atol=2e-6)
    assert np.allclose((md[:, None] * vel).sum(0) / md.sum(), vcom, atol=2e-6)


def test_random_modifiers_bounded_and_uniform(out):
    pos, m, vel, rho = rows(out, "noise.bin")
    pos0, m0, vel0, rho0 = rows(out, "plain.bin")
    assert np.array_equal(m, m0) and np.array_equal(rho, rho0) and np.all(vel0 == 0)
    dp = (pos.astype(np.float64) - pos0).ravel()
    assert np.abs(dp).max() <= 0.01 * (1 + 1e-4) and np.abs(dp).max() > 0.005
    assert ks_ok(dp, stats.uniform(-0.01, 0.02).cdf)
    dv = vel.astype(np.float64).ravel()
    assert np.abs(dv).max() <= 0.2 and ks_ok(dv, stats.uniform(-0.2, 0.4).cdf)


def test_two_sources_in_order(out):
    pos, m, vel, rho = rows(out, "two.bin")
    assert len(m) == 150
    assert np.all(m[:100] == f32(2.0) / f32(100)) and np.all(rho[:100] == 3.0)
    assert np.all(m[100:] == f32(4.0) / f32(50)) and np.all(rho[100:] == 6.0)
    assert np.linalg.norm(pos[:100], axis=1).max() <= 1.0 + 1e-6
    assert np.linalg.norm(pos[100:] - np.array([3, 0, 0], dtype=f32), axis=1).max() <= 2.0 + 1e-5
