"""Pins of the CPU oracle against closed forms, invariants and physics.

The reference ships no tests, fixtures or golden vectors for this path and
its physics headers cannot be built here (they need mpUtils, absent from the
image), so each oracle function is checked against what the mathematics
fixes, with checks a plausible mistake (dropped term, wrong sign, transposed
index, wrong factor) would fail.  CPU only.
"""
import math

import numpy as np
import pytest

import grasph2_amd as g
from grasph2_amd import ics
from oracle import pyoracle as O


def cubic_lattice(m, d, center=True):
    ax = np.arange(m) * d
    p = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), axis=-1).reshape(-1, 3)
    if center:
        p -= p.mean(0)
    return p


def make(pos, mass, vel=None, fields=ics.SOLID_FIELDS + ("xvel",)):
    n = len(pos)
    h = g.host_particles(n, fields)
    h["posm"][:, :3] = pos
    h["posm"][:, 3] = mass
    if "pos64" in h:
        h["pos64"][:] = np.asarray(pos, dtype=np.float64).T
    if vel is not None:
        h["vel"][:, :3] = vel
    return h


def interior(pos, margin):
    lo, hi = pos.min(0), pos.max(0)
    return np.where(np.all((pos > lo + margin) & (pos < hi - margin), axis=1))[0]


# ---------------------------------------------------------------- kernel
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_kernel_normalised(dim):
    h = 0.37
    r = np.linspace(0.0, h, 20001)
    w = np.array([O.W(x, h, dim) for x in r])
    if dim == 3:
        f = 4 * np.pi * r * r * w
    elif dim == 2:
        f = 2 * np.pi * r * w
    else:
        f = 2 * w
    integral = np.trapezoid(f, r)
    assert abs(integral - 1.0) < 1e-6


def test_kernel_values_and_support():
    h = 0.1
    assert O.W(0.0, h) == pytest.approx(8 / (math.pi * h ** 3), rel=1e-14)
    # continuity where the pieces meet (both give 1/4) and at the support radius h
    for q in (0.5, 1.0):
        lo, hi = O.W(q * h * (1 - 1e-10), h), O.W(q * h * (1 + 1e-10), h)
        assert abs(lo - hi) < 1e-6 * O.W(0, h)
    assert O.W(0.5 * h, h) == pytest.approx(0.25 * 8 / (math.pi * h ** 3), rel=1e-12)
    assert O.W(1.0001 * h, h) == 0.0 and O.W(2 * h, h) == 0.0
    assert O.W(0.999 * h, h) > 0.0


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_kernel_derivative_is_dW_dr(dim):
    h = 0.2
    for q in (0.05, 0.2, 0.45, 0.55, 0.7, 0.95):
        r = q * h
        e = 1e-7 * h
        fd = (O.W(r + e, h, dim) - O.W(r - e, h, dim)) / (2 * e)
        assert O.dW(r, h, dim) == pytest.approx(fd, rel=1e-6)
    assert O.dW(1.5 * h, h, dim) == 0.0


# ---------------------------------------------------------------- EOS
def test_eos():
    assert O.eos_liquid(1.2, 1.0, 4.0) == pytest.approx(0.8)
    assert O.eos_liquid(1.0, 1.0, 4.0) == 0.0
    assert O.eos_murnaghan(1.0, 2.0 / 2.0, 5.0, 3.0) == 0.0
    # bulk modulus K = rho dp/drho at the rest density, for any exponent
    for n in (1.0, 4.0, 7.5):
        K, r0 = 2.0, 1.3
        e = 1e-6
        dp = (O.eos_murnaghan(r0 + e, r0, K, n) - O.eos_murnaghan(r0 - e, r0, K, n)) / (2 * e)
        assert r0 * dp == pytest.approx(K, rel=1e-7)
    assert O.eos_murnaghan(2.0, 1.0, 3.0, 1.0) == pytest.approx(3.0)


# ---------------------------------------------------------------- AV / Balsara / plasticity
def test_artificial_viscosity():
    # receding pair: no viscosity, signal velocity untouched
    II, mv = O.artificial_viscosity(1.0, 1.0, 1.0, [0.5, 0, 0], [1.0, 0, 0], 1.0, 1.0, 1.0, maxvsig=0.25)
    assert II == 0.0 and mv == 0.25
    # head-on approach at 0.5: w=-0.5, vsig = 1+1+1.5 = 3.5, Pi = 0.25*2*1*0.5*3.5/1 = 0.875
    II, mv = O.artificial_viscosity(1.0, 1.0, 1.0, [-0.5, 0, 0], [1.0, 0, 0], 1.0, 1.0, 1.0, maxvsig=0.0)
    assert II == pytest.approx(0.875) and mv == pytest.approx(3.5)
    # Balsara factors and mean density enter linearly / inversely
    II2, _ = O.artificial_viscosity(1.0, 2.0, 2.0, [-0.5, 0, 0], [1.0, 0, 0], 1.0, 1.0, 1.0, 0.5, 0.5)
    assert II2 == pytest.approx(0.875 / 4)


def test_balsara_switch():
    assert O.balsara_switch(3.0, [0, 0, 0], 1.0, 0.1) == pytest.approx(3.0 / (3.0 + 1e-3))
    assert O.balsara_switch(0.0, [1.0, 2.0, 0], 1.0, 0.1) == 0.0
    assert O.balsara_switch(-1.0, [0, 3.0, 4.0], 1.0, 0.1) == pytest.approx(1.0 / (6.0 + 1e-3))


def test_plasticity_von_mises():
    S = np.zeros((3, 3))
    S[0, 1] = S[1, 0] = 0.1
    out = O.plasticity(S, 0.05)  # J2 = 0.01, factor Y^2/(3 J2) = 1/12
    assert out[0, 1] == pytest.approx(0.1 / 12) and out[1, 0] == pytest.approx(0.1 / 12)
    assert np.all(out[np.eye(3, dtype=bool)] == 0)
    assert np.array_equal(O.plasticity(S, 1.0), S)          # inside the yield surface
    assert np.array_equal(O.plasticity(np.zeros((3, 3)), 0.05), np.zeros((3, 3)))  # J2 = 0


def test_dstress_dt_closed_forms():
    mu = 1.7
    # no rotation: Hooke's deviatoric law 2 mu (edot - tr/3 I)
    E = np.array([[0.3, 0.1, 0.0], [0.1, -0.2, 0.05], [0.0, 0.05, 0.4]])
    S = np.array([[0.2, 0.0, 0.1], [0.0, -0.1, 0.0], [0.1, 0.0, -0.1]])
    out = O.dstress_dt(E, np.zeros((3, 3)), S, mu)
    assert np.allclose(out, 2 * mu * (E - np.trace(E) / 3 * np.eye(3)), atol=1e-14)
    # pure rotation: the stress co-rotates, so J2 is conserved and the rate is R S - S R
    R = np.array([[0.0, 0.3, -0.2], [-0.3, 0.0, 0.5], [0.2, -0.5, 0.0]])
    out = O.dstress_dt(np.zeros((3, 3)), R, S, mu)
    assert abs(np.sum(S * out)) < 1e-14
    Q = np.eye(3)
    Sq = S.copy()
    dt = 1e-4
    for _ in range(1000):  # d/dt (Q S Q^T) with Qdot = R Q integrates to exp(Rt) S exp(-Rt)
        Sq = Sq + dt * O.dstress_dt(np.zeros((3, 3)), R, Sq, mu)
    w, V = np.linalg.eig(R * 0.1)
    Qt = (V @ np.diag(np.exp(w)) @ np.linalg.inv(V)).real
    assert np.allclose(Sq, Qt @ S @ Qt.T, atol=2e-5)


# ---------------------------------------------------------------- density pass
def test_density_pair_and_isolated():
    H = 0.1
    prm = ics.liquid_params(H, gravity=False)
    h = make(np.array([[0.0, 0, 0], [0.06, 0, 0], [5.0, 0, 0]]), np.array([1.0, 2.0, 3.0]))
    out = O.calc_density(prm, h)
    Hf = prm.H  # H as the params struct stores it (float32)
    assert out["density"][0] == pytest.approx(2.0 * O.W(0.06, Hf), rel=1e-12)
    assert out["density"][1] == pytest.approx(1.0 * O.W(0.06, Hf), rel=1e-12)
    assert out["density"][2] == 0.0  # self term excluded
    prm.flags |= g.DEAL_WITH_NO_PARTNERS
    out = O.calc_density(prm, h)
    assert out["density"][2] == pytest.approx(0.01 * prm.rho0)


def test_density_lattice_normalisation():
    d = 0.05
    H = 2.5 * d
    pos = cubic_lattice(16, d)
    rho0 = 1.0
    m = rho0 * d ** 3
    prm = ics.liquid_params(H, gravity=False, rho0=rho0)
    h = make(pos, np.full(len(pos), m))
    idx = interior(pos, H * 1.05)
    out = O.calc_density(prm, h, sample=idx)
    rho = out["density"][idx]
    selfterm = m * O.W(0.0, H)
    assert np.all(np.abs(rho + selfterm - rho0) < 0.01 * rho0)
    assert np.all(rho < 0.9 * rho0)  # the missing self term is a ~16% deficit at H = 2.5 spacing


# ---------------------------------------------------------------- strain / rotation pass
def _strain_case(L, S0=None):
    d = 0.05
    H = 2.5 * d
    pos = cubic_lattice(14, d)
    prm = ics.solid_params(H, gravity=False, mu=1.3)
    n = len(pos)
    h = make(pos, np.full(n, d ** 3), vel=pos @ np.asarray(L).T)
    h["density"][:] = 1.0  # m/rho equals the particle volume
    if S0 is not None:
        for c, (a, b) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]):
            h["dstress"][c] = S0[a][b]
    idx = interior(pos, H * 1.05)
    out = O.calc_balsara(prm, h, sample=idx)
    return prm, out, idx


def _full(six, i):
    xx, yy, zz, xy, xz, yz = (six[c][i] for c in range(6))
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


def test_strain_rate_of_linear_field():
    L = np.array([[0.3, 1.0, 0.0], [0.0, -0.1, 0.2], [0.4, 0.0, 0.2]])
    prm, out, idx = _strain_case(L)
    # with S = 0, dS/dt = 2 mu dev(edot) with edot ~ sym(L): asymmetric L pins the transpose
    sym = 0.5 * (L + L.T)
    expect = 2 * prm.shear * (sym - np.trace(sym) / 3 * np.eye(3))
    for i in idx[:: max(1, len(idx) // 20)]:
        got = _full(out["dstress_dt"], i)
        assert np.allclose(got, expect, atol=0.02 * np.abs(expect).max())
    # Balsara: |div v| / (|div v| + |curl v| + eps); curl of L x is (L32-L23, L13-L31, L21-L12)
    div = np.trace(L)
    curl = np.array([L[2, 1] - L[1, 2], L[0, 2] - L[2, 0], L[1, 0] - L[0, 1]])
    b = abs(div) / (abs(div) + np.linalg.norm(curl) + 1e-4 * prm.SOUNDSPEED / prm.H)
    assert np.allclose(out["balsara"][idx], b, rtol=0.03)


def test_stress_corotates_with_rigid_rotation():
    w = np.array([0.3, -0.2, 0.5])
    Om = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])  # Om x = w cross x
    S0 = np.array([[0.2, 0.05, 0.0], [0.05, -0.1, 0.03], [0.0, 0.03, -0.1]])
    prm, out, idx = _strain_case(Om, S0)
    expect = Om @ S0 - S0 @ Om  # d/dt (Q S Q^T) at t = 0
    for i in idx[:: max(1, len(idx) // 20)]:
        got = _full(out["dstress_dt"], i)
        assert np.allclose(got, expect, atol=0.02 * np.abs(expect).max())
    assert np.all(out["balsara"][idx] < 1e-3)  # pure rotation: no compression


# ---------------------------------------------------------------- acceleration pass
def test_two_body_gravity():
    prm = ics.plummer_params(H=0.0)
    h = make(np.array([[0.0, 0, 0], [1.0, 0, 0]]), np.array([1.0, 2.0]))
    out = O.calc_acceleration(prm, h)
    assert np.allclose(out["acc"][0, :3], [2.0, 0, 0], rtol=1e-14)
    assert np.allclose(out["acc"][1, :3], [-1.0, 0, 0], rtol=1e-14)
    prm.H = 0.5  # softening is H: 2 / 1.25^1.5
    out = O.calc_acceleration(prm, h)
    assert out["acc"][0, 0] == pytest.approx(2.0 / 1.25 ** 1.5, rel=1e-7)


def test_gravity_is_minus_grad_potential():
    rng = np.random.default_rng(11)
    pos = rng.normal(size=(40, 3))
    m = rng.uniform(0.5, 1.5, size=40)
    prm = ics.plummer_params(H=0.3)
    h = make(pos, m)
    acc = O.calc_acceleration(prm, h)["acc"][:, :3]
    i, e = 7, 1e-5
    grad = np.zeros(3)
    for d in range(3):
        hp, hm = make(pos.copy(), m), make(pos.copy(), m)
        hp["posm"][i, d] += e
        hm["posm"][i, d] -= e
        hp["pos64"][d, i] += e
        hm["pos64"][d, i] -= e
        grad[d] = (O.potential(prm, hp)[i] - O.potential(prm, hm)[i]) / (2 * e)
    assert np.allclose(acc[i], -grad, rtol=1e-6)
    mf = h["posm"][:, 3].astype(np.float64)  # the masses the oracle saw (float32 storage)
    assert np.allclose((acc * mf[:, None]).sum(0), 0, atol=1e-12)


def _sph_lattice(m=14, d=0.05):
    pos = cubic_lattice(m, d)
    return pos, d, 2.5 * d


def test_pressure_gradient_drives_acceleration():
    pos, d, H = _sph_lattice()
    eps, c = 0.2, 1.5
    rho = 1.0 * (1 + eps * pos[:, 0])
    prm = ics.liquid_params(H, gravity=False, c=c)
    prm.flags &= ~g.ARTIFICIAL_VISCOSITY
    h = make(pos, rho * d ** 3, fields=("posm", "vel", "acc", "density", "maxvsig", "xvel"))
    h["density"][:] = rho
    idx = interior(pos, H * 1.05)
    acc = O.calc_acceleration(prm, h, sample=idx)["acc"][idx, :3]
    expect = -c * c * 1.0 * eps / rho[idx]  # -grad p / rho
    assert np.allclose(acc[:, 0], expect, rtol=0.05)
    assert np.all(np.abs(acc[:, 1:]) < 0.02 * np.abs(expect).max())


def test_shear_stress_divergence():
    pos, d, H = _sph_lattice()
    s = 0.3
    prm = ics.solid_params(H, gravity=False)
    prm.flags &= ~(g.ARTIFICIAL_VISCOSITY | g.BALSARA_SWITCH)
    h = make(pos, np.full(len(pos), d ** 3))
    h["density"][:] = 1.0  # rho = rho0: no pressure
    h["dstress"][3] = s * pos[:, 0]  # sigma_xy = s x  ->  a_y = d sigma_yx / dx = s
    idx = interior(pos, H * 1.05)
    acc = O.calc_acceleration(prm, h, sample=idx)["acc"][idx, :3]
    assert np.allclose(acc[:, 1], s, rtol=0.05)
    assert np.all(np.abs(acc[:, [0, 2]]) < 0.02 * s)


def test_sph_momentum_conservation_and_viscous_dissipation():
    rng = np.random.default_rng(5)
    pos, d, H = _sph_lattice(10)
    pos = pos + rng.uniform(-0.2 * d, 0.2 * d, size=pos.shape)
    n = len(pos)
    m = d ** 3 * rng.uniform(0.9, 1.1, size=n)
    vel = rng.normal(scale=0.3, size=(n, 3))
    prm = ics.solid_params(H, gravity=True)
    h = make(pos, m, vel=vel)
    h["dstress"][:] = rng.normal(scale=0.05, size=(6, n))
    out = O.compute_derivatives(prm, h)
    a = out["acc"][:, :3]
    mf = h["posm"][:, 3].astype(np.float64)
    P = (a * mf[:, None]).sum(0)
    assert np.all(np.abs(P) < 1e-13 * np.abs(a * mf[:, None]).sum())
    # viscosity alone (no pressure, no stress, no gravity) removes kinetic energy
    prm2 = ics.solid_params(H, gravity=False)
    h2 = make(pos, m, vel=vel)
    h2["density"][:] = prm2.rho0
    prm2.flags &= ~(g.BALSARA_SWITCH)
    h2_out = O.calc_acceleration(prm2, h2)
    a2 = h2_out["acc"][:, :3]
    v2 = h2["vel"][:, :3].astype(np.float64)
    assert np.sum(mf[:, None] * v2 * a2) < 0
    assert np.all(h2_out["maxvsig"] == 0)  # the signal velocity is only tracked with the Balsara switch


# ---------------------------------------------------------------- integrator
def test_leapfrog_exact_for_constant_acceleration():
    prm = ics.plummer_params(H=0.0, dt=0.01)
    prm.flags = g.CONSTANT_ACCELERATION
    prm.constant_acceleration[:] = (0.0, -2.0, 0.5)
    x0 = np.array([[0.3, 0.2, -0.1]])
    v0 = np.array([[1.0, 0.5, 0.0]])
    h = make(x0, np.array([1.0]), vel=v0, fields=("posm", "pos64", "vel", "acc"))
    gacc = np.array(prm.constant_acceleration[:])
    for k in range(1, 51):
        h.update({f: v for f, v in O.calc_acceleration(prm, h).items() if f == "acc"})
        o = O.leapfrog(prm, h, not_first_step=(k > 1))
        h["posm"], h["vel"] = o["posm"], o["vel"]
        h["pos64"] = o["posm"][:, :3].T.copy()
        t = k * prm.fixed_timestep
        assert np.allclose(h["posm"][0, :3], x0[0] + v0[0] * t + 0.5 * gacc * t * t, atol=1e-12)
        assert np.allclose(h["vel"][0, :3], v0[0] + gacc * (t - 0.5 * prm.fixed_timestep), atol=1e-12)


def test_leapfrog_circular_orbit():
    prm = ics.plummer_params(H=0.0, dt=2 * math.pi / 2000)
    h = make(np.array([[-0.5, 0, 0], [0.5, 0, 0]]), np.array([0.5, 0.5]), vel=np.array([[0, -0.5, 0], [0, 0.5, 0]]),
             fields=("posm", "pos64", "vel", "acc"))
    e0 = None
    for k in range(2000):
        h["acc"] = O.calc_acceleration(prm, h)["acc"]
        if k == 0:
            inv = O.invariants(prm, h, half_kick=0.0, phi=O.potential(prm, h))
            e0 = inv[g.INV_EKIN] + inv[g.INV_EPOT]
        o = O.leapfrog(prm, h, not_first_step=(k > 0))
        h["posm"], h["vel"] = o["posm"], o["vel"]
        h["pos64"] = o["posm"][:, :3].T.copy()
    # one period: back at the start
    assert np.allclose(h["posm"][:, :3], [[-0.5, 0, 0], [0.5, 0, 0]], atol=2e-4)
    h["acc"] = O.calc_acceleration(prm, h)["acc"]
    inv = O.invariants(prm, h, half_kick=0.5 * prm.fixed_timestep, phi=O.potential(prm, h))
    assert abs(inv[g.INV_EKIN] + inv[g.INV_EPOT] - e0) < 1e-6 * abs(e0)
    assert inv[g.INV_LZ] == pytest.approx(0.25, rel=1e-9)
    assert abs(e0 - (-0.25 + 0.125)) < 1e-12  # E = -m1 m2 / d + 2 * 1/2 m v^2


def test_leapfrog_stress_density_boundary():
    prm = ics.solid_params(0.1, dt=0.01)
    prm.flags |= g.INTEGRATE_DENSITY | g.SIMPLE_BOX_BOUNDARY
    prm.box_min[:] = (-1, -1, -1)
    prm.box_max[:] = (1, 1, 1)
    prm.box_reflectiveness = 0.5
    h = make(np.array([[0.0, 0, 0], [0.999, 0, 0]]), np.array([1.0, 1.0]), vel=np.array([[0, 0, 0], [1.0, 0, 0]]))
    h["density"][:] = [1.0, 0.001]
    h["density_dt"][:] = [2.0, -1.0]
    h["dstress_dt"][3, 0] = 0.1 / prm.fixed_timestep  # S_xy = 0.1 after the drift -> beyond yield
    o = O.leapfrog(prm, h, not_first_step=True)
    assert o["density"][0] == pytest.approx(1.02)
    assert o["density"][1] == 0.0  # clamped at zero
    assert o["dstress"][3, 0] == pytest.approx(0.1 / 12)  # von Mises factor Y^2/(3 J2), Y = 0.05
    assert o["posm"][1, 0] == pytest.approx(1.0)  # clamped to the wall
    assert o["vel"][1, 0] == pytest.approx(-0.5)  # reflected with reflectiveness 0.5


def test_mohr_coulomb_yield_by_hand():
    """PLASTICITY_MC (integration.h:24-27, models.h:207-210): Y = tan(phi) p + c
    with the Murnaghan pressure of the density after the drift; factor
    Y^2 / (3 J2) below yield, unchanged above.  Three particles: compressed
    (p = 1), at rest density (p = 0: only the cohesion), in tension (p = -1:
    the sign of p enters Y, and Y^2 = 0.79^2 rather than 0.81^2)."""
    prm = ics.solid_params(0.1, dt=0.01)  # rho0 = 1, K = 2, n = 1
    prm.flags = (prm.flags | g.INTEGRATE_DENSITY | g.PLASTICITY_MC) & ~g.PLASTICITY_MIESE
    prm.tanfr, prm.cohesion = 0.8, 0.01
    h = make(np.array([[0.0, 0, 0], [3.0, 0, 0], [6.0, 0, 0]]), np.ones(3))
    h["density"][:] = [1.4, 1.0, 0.5]
    h["density_dt"][:] = [10.0, 0.0, 0.0]  # particle 0 reaches rho = 1.5 in the drift: p = 2 (1.5 - 1) = 1
    h["dstress_dt"][3, :] = np.array([1.0, 0.005, 1.0]) / prm.fixed_timestep  # S_xy after the drift
    o = O.leapfrog(prm, h, not_first_step=True)
    assert o["density"][0] == pytest.approx(1.5, rel=1e-7)  # inputs stored in FP32
    J2 = 1.0  # 1/2 (S_xy^2 + S_yx^2) with S_xy = 1
    assert o["dstress"][3, 0] == pytest.approx(0.81 ** 2 / (3 * J2), rel=1e-6)
    assert o["dstress"][3, 1] == pytest.approx(0.005, rel=1e-6)  # Y = c = 0.01: f = 4/3 > 1, unchanged
    assert o["dstress"][3, 2] == pytest.approx(0.79 ** 2 / (3 * J2), rel=1e-6)
    # without INTEGRATE_DENSITY the stored density sets the pressure
    prm.flags &= ~g.INTEGRATE_DENSITY
    o = O.leapfrog(prm, h, not_first_step=True)
    assert o["dstress"][3, 0] == pytest.approx((0.8 * 2 * 0.4 + 0.01) ** 2 / 3, rel=1e-6)


def _liquid_pair(rho_i, rho_j, r, H=0.1, arts=True, matexp=4.0, normalsep=0.03):
    prm = ics.liquid_params(H, gravity=False, c=1.0)
    prm.flags &= ~g.ARTIFICIAL_VISCOSITY
    if arts:
        prm.flags |= g.ARTIFICIAL_STRESS
    prm.mateps, prm.matexp, prm.normalsep = 0.4, matexp, normalsep
    h = make(np.array([[0.0, 0, 0], [r, 0, 0]]), np.array([1.0, 2.0]))
    h["density"][:] = [rho_i, rho_j]
    return prm, h


def test_artificial_pressure_by_hand():
    """Liquid ARTIFICIAL_STRESS (computeDerivatives.h:263-272, models.h:218-221):
    the pair term adds f (R_i + R_j) with f = (W(r) / W(normalsep))^matexp and
    R = -mateps sigma/rho^2 where sigma/rho^2 > 0 (tension), 0 otherwise.  The
    change it makes to a_i is m_j f (R_i + R_j) dW(r)/r r_ij.  Inputs are
    stored in FP32, hence rel 1e-6."""
    r, H = 0.06, 0.1
    Hf = float(np.float32(H))

    def delta(rho_i, rho_j, rr=r, matexp=4.0, ns=0.03):
        on = O.calc_acceleration(*_liquid_pair(rho_i, rho_j, rr, matexp=matexp, normalsep=ns))["acc"][0, :3]
        off = O.calc_acceleration(*_liquid_pair(rho_i, rho_j, rr, arts=False))["acc"][0, :3]
        return on - off

    gradx = O.dW(r, Hf) / r * (-r)  # r_ij = x_i - x_j = -r along x
    # both in tension: sigma = -c^2 (rho - rho0) > 0
    sor_i, sor_j = 0.1 / 0.9 ** 2, 0.2 / 0.8 ** 2
    f = (O.W(r, Hf) / O.W(float(np.float32(0.03)), Hf)) ** 4
    d = delta(0.9, 0.8)
    assert d[0] == pytest.approx(2.0 * f * (-0.4 * sor_i - 0.4 * sor_j) * gradx, rel=1e-6)
    assert d[1] == 0 and d[2] == 0
    # mixed: only the particle in tension contributes (the sign gate)
    d = delta(0.9, 1.1)
    assert d[0] == pytest.approx(2.0 * f * (-0.4 * sor_i) * gradx, rel=1e-6)
    # both compressed: nothing
    assert np.all(delta(1.1, 1.2) == 0)
    # at r = normalsep, f = 1 whatever matexp is
    ns = float(np.float32(r))
    d4, d2 = delta(0.9, 0.8, ns=ns, matexp=4.0), delta(0.9, 0.8, ns=ns, matexp=2.0)
    assert d4[0] == pytest.approx(d2[0], rel=1e-6)  # r in FP64, normalsep in FP32
    assert d4[0] == pytest.approx(2.0 * (-0.4 * sor_i - 0.4 * sor_j) * gradx, rel=1e-6)


def test_xsph_by_hand():
    """XSPH (computeDerivatives.h:291-294): xvel_i = sum_j 2 m_j / (rho_i + rho_j)
    (v_j - v_i) W(r); a pair by hand, zero for a uniform velocity field, and
    sum_i m_i xvel_i = 0 (antisymmetric pair terms); the drift moves with
    v + xsph_factor xvel (integration.h:98-99)."""
    H = 0.1
    Hf = float(np.float32(H))
    prm = ics.liquid_params(H, gravity=False)
    prm.flags = g.SPH | g.XSPH
    h = make(np.array([[0.0, 0, 0], [0.06, 0, 0], [5.0, 0, 0]]), np.array([1.0, 2.0, 1.0]),
             vel=np.array([[1.0, 0, 0], [0.0, 1.0, 0], [0.0, 0, 0]]))
    h["density"][:] = [1.0, 3.0, 1.0]
    o = O.calc_acceleration(prm, h)
    w = O.W(float(np.float32(0.06)), Hf)
    assert np.allclose(o["xvel"][0, :3], np.array([-1.0, 1.0, 0.0]) * (2 * 2.0 / 4.0) * w, rtol=1e-6)
    assert np.allclose(o["xvel"][1, :3], np.array([1.0, -1.0, 0.0]) * (2 * 1.0 / 4.0) * w, rtol=1e-6)
    assert np.all(o["xvel"][2, :3] == 0)
    m = h["posm"][:, 3].astype(np.float64)
    assert np.allclose((o["xvel"][:, :3] * m[:, None]).sum(0), 0, atol=1e-9 * w)
    # uniform flow on a lattice: no correction
    pos = cubic_lattice(6, 0.05)
    hu = make(pos, np.full(len(pos), 0.05 ** 3), vel=np.tile([0.3, -0.2, 0.1], (len(pos), 1)))
    hu["density"][:] = 1.0
    prm.H = 0.125
    assert np.all(O.calc_acceleration(prm, hu)["xvel"][:, :3] == 0)
    # the drift: x += (v + xsph_factor xvel) dt
    prm.H = H
    prm.fixed_timestep = 0.01
    prm.xsph_factor = 0.5
    st = dict(h)
    st["xvel"] = o["xvel"]
    lf = O.leapfrog(prm, st, not_first_step=True)
    dt = float(np.float32(0.01))  # as grasph_params holds it
    expect = h["posm"][0, :3] + (h["vel"][0, :3] + 0.5 * o["xvel"][0, :3]) * dt
    assert np.allclose(lf["posm"][0, :3], expect, rtol=0, atol=1e-12)


def _liquid_rate_case(L, rho=2.0):
    """liquid lattice with v = L x, rho = 2 and m = 2 d^3 (m/rho = volume)"""
    d = 0.05
    H = 2.5 * d
    pos = cubic_lattice(14, d)
    prm = ics.liquid_params(H, gravity=False, rho0=1.0)
    n = len(pos)
    h = make(pos, np.full(n, rho * d ** 3), vel=pos @ np.asarray(L).T)
    h["density"][:] = rho
    idx = interior(pos, H * 1.05)
    return prm, h, idx


def test_density_rate_is_minus_rho_div_v():
    """INTEGRATE_DENSITY (computeDerivatives.h:136-138,152-155):
    rho_dot = -rho div v; v = L x on a lattice gives div v = tr L, so
    rho_dot = -2 tr L at rho = 2 (a dropped rho_i would give -tr L)."""
    L = np.array([[0.3, 1.0, 0.0], [0.0, -0.1, 0.2], [0.4, 0.0, 0.5]])
    for solid in (False, True):
        prm, h, idx = _liquid_rate_case(L)
        prm.flags |= g.INTEGRATE_DENSITY
        if solid:
            prm.flags |= g.SOLIDS
            prm.BULK, prm.dBULKdP, prm.shear = 2.0, 1.0, 1.0
        out = O.calc_balsara(prm, h, sample=idx)
        assert np.allclose(out["density_dt"][idx], -2.0 * np.trace(L), rtol=0.02), solid


def test_liquid_balsara_curl():
    """Liquid BALSARA_SWITCH (computeDerivatives.h:133-135): curl v from
    sum m_j/rho_j v_ij x grad W, div v from the same weights; on v = L x,
    |curl| = |(L32 - L23, L13 - L31, L21 - L12)| and div = tr L, so
    beta = |div| / (|div| + |curl| + 1e-4 c/H); pure rotation gives ~0,
    pure expansion ~1."""
    cases = [
        np.array([[0.3, 1.0, 0.0], [0.0, -0.1, 0.2], [0.4, 0.0, 0.2]]),   # mixed, asymmetric
        np.array([[0.0, -0.5, 0.2], [0.5, 0.0, -0.3], [-0.2, 0.3, 0.0]]),  # rigid rotation
        np.eye(3) * 0.4,                                                   # expansion
    ]
    for k, L in enumerate(cases):
        prm, h, idx = _liquid_rate_case(L)
        prm.flags |= g.BALSARA_SWITCH
        out = O.calc_balsara(prm, h, sample=idx)
        div = np.trace(L)
        curl = np.array([L[2, 1] - L[1, 2], L[0, 2] - L[2, 0], L[1, 0] - L[0, 1]])
        b = abs(div) / (abs(div) + np.linalg.norm(curl) + 1e-4 * prm.SOUNDSPEED / prm.H)
        got = out["balsara"][idx]
        if k == 0:
            assert np.allclose(got, b, rtol=0.03)
        elif k == 1:
            assert np.all(got < 1e-3)
        else:
            assert np.all(got > 0.99)


# ---------------------------------------------------------------- variable step
def _vts_params(H=0.05):
    prm = ics.plummer_params(H=H, dt=0.002)
    prm.flags = g.VARIABLE_TIMESTEP | g.ACCELERATION_CRITERION | g.VELOCITY_CRITERION
    prm.initial_timestep, prm.min_timestep, prm.max_timestep = 0.002, 2e-5, 0.02
    prm.accel_accuracy, prm.velocity_accuracy = 0.01, 0.09
    return prm


def _vts_state(acc, maxvsig):
    n = len(acc)
    h = make(np.zeros((n, 3)), np.ones(n), fields=("posm", "vel", "acc", "maxvsig"))
    h["acc"][:, :3] = acc
    h["maxvsig"][:] = maxvsig
    return h


def test_variable_step_criteria_by_hand():
    """getNewTimestep's criteria (integration.h:134,149-154) and
    useNewTimestep's clamp (:170), each step starting from the next dt an
    acceleration pass leaves (max_timestep; the reset itself is pinned by
    test_acceleration_pass_resets_next_dt).  Parameters are FP32 (as in
    grasph_params), hence the 1e-6 tolerances."""
    prm = _vts_params()
    st = O.TimestepState("f64")
    h = _vts_state([[3.0, 4.0, 0.0], [0.0, 0.0, 0.0]], [0.0, 2.0])
    O.variable_leapfrog(prm, h, st, not_first_step=False)
    assert (st.current, st.next) == pytest.approx((0.002, 0.002), rel=1e-7)  # first step: initial_timestep

    def step(acc, vsig):
        st.ts[1] = np.float64(np.float32(prm.max_timestep))  # as the acceleration pass leaves it
        O.variable_leapfrog(prm, _vts_state(acc, vsig), st)
        return st.current

    # |a| = 5: sqrt(0.01 * 0.05 / 5) = 0.01; 0.09 * 0.05 / 2 = 0.00225 binds
    assert step([[3.0, 4.0, 0.0], [0.0, 0.0, 0.0]], [0.0, 2.0]) == pytest.approx(0.00225, rel=1e-6)
    # |a| = 500: sqrt(1e-6) = 0.001 binds
    assert step([[300.0, 400.0, 0.0], [0.0, 0.0, 0.0]], [0.0, 2.0]) == pytest.approx(0.001, rel=1e-6)
    # 0.09 * 0.05 / 100 = 4.5e-5
    assert step([[0.0, 0.0, 0.0]], [100.0]) == pytest.approx(4.5e-5, rel=1e-6)
    # no criterion binds: max_timestep
    assert step([[0.0, 0.0, 0.0]], [0.0]) == pytest.approx(0.02, rel=1e-6)
    # sqrt(1e-12) = 1e-6 < min_timestep: the next dt keeps it, the step clamps
    assert step([[5e8, 0.0, 0.0]], [0.0]) == pytest.approx(2e-5, rel=1e-6)
    assert st.next == pytest.approx(1e-6, rel=1e-6)
    # without an acceleration pass between two integrate calls the minimum carries over
    O.variable_leapfrog(prm, _vts_state([[0.0, 0.0, 0.0]], [0.0]), st)
    assert st.next == pytest.approx(1e-6, rel=1e-6) and st.current == pytest.approx(2e-5, rel=1e-6)


def _two_body(r, H=0.05):
    h = make(np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]]), np.array([1.0, 1.0]))
    h["density"][:] = 1.0
    prm = _vts_params(H=H)
    prm.flags |= g.SELF_GRAVITY
    prm.flags &= ~g.VELOCITY_CRITERION
    return h, prm


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_acceleration_pass_resets_next_dt(prec):
    """calcAcceleration::do_after sets nextTS = max_timestep
    (computeDerivatives.h:305-308) with the variable step, and only then;
    the other passes leave it."""
    h, prm = _two_body(0.1)
    st = O.TimestepState(prec)
    st.ts[:] = [0.25, 0.125]
    O.calc_density(prm, h, prec)
    O.calc_balsara(prm, h, prec)
    assert list(st.ts) == [0.25, 0.125]
    O.calc_acceleration(prm, h, prec, state=st)
    assert st.current == 0.25 and st.next == pytest.approx(0.02, rel=1e-7)
    st.ts[:] = [0.25, 0.125]
    O.compute_derivatives(prm, h, prec, state=st)
    assert st.next == pytest.approx(0.02, rel=1e-7)
    fixed = prm.copy()
    fixed.flags &= ~g.VARIABLE_TIMESTEP
    st.ts[:] = [0.25, 0.125]
    O.compute_derivatives(fixed, h, prec, state=st)
    assert list(st.ts) == [0.25, 0.125]


def test_main_loop_dt_is_min_of_max_and_criteria():
    """computeDerivatives then integrate(pb, true) from the first step on, as
    main.cu:256-257 (the state zero, as the reference's __device__ symbols
    start): dt = min(max_timestep, sqrt(accel_accuracy H / |a|)) with the
    two-body |a| = m r / (r^2 + H^2)^(3/2) by hand; it shrinks as the pair
    closes and grows back, up to max_timestep, as it separates."""
    H = 0.05
    st = O.TimestepState("f64")
    dts = []
    for r in (0.1, 0.05, 1.0, 0.2):
        h, prm = _two_body(r, H)
        o = O.compute_derivatives(prm, h, "f64", state=st)
        H32 = float(np.float32(H))  # grasph_params holds FP32 values
        amag = r / (r * r + H32 * H32) ** 1.5
        assert np.linalg.norm(o["acc"][1, :3]) == pytest.approx(amag, rel=1e-12)
        O.variable_leapfrog(prm, o, st, not_first_step=True)
        expect = min(float(np.float32(prm.max_timestep)), math.sqrt(float(np.float32(0.01)) * H32 / amag))
        assert st.current == pytest.approx(expect, rel=1e-12), r
        dts.append(st.current)
    assert dts[1] < dts[0] < dts[3] < dts[2]
    assert dts[2] == pytest.approx(0.02, rel=1e-7)  # max_timestep binds at r = 1


def test_variable_step_exact_for_constant_acceleration():
    """Kick-drift-kick with synchronised half kicks integrates constant
    acceleration exactly for any dt sequence: x = x0 + v0 t + a t^2 / 2.
    The acceleration is CONSTANT_ACCELERATION (the only force), each step
    runs the oracle's derivative path (which resets the next dt) and then
    integrate(pb, true); accel_accuracy changes per step, so dt shrinks and
    grows (pins the half kick with the current dt and the drift with the new
    one)."""
    prm = _vts_params(H=0.1)
    prm.flags = g.VARIABLE_TIMESTEP | g.ACCELERATION_CRITERION | g.CONSTANT_ACCELERATION
    a = np.array([0.0, -2.0, 0.5])
    prm.constant_acceleration[:] = a
    x0, v0 = np.array([0.3, 0.2, -0.1]), np.array([1.0, 0.5, 0.0])
    h = make(x0[None], np.array([1.0]), vel=v0[None], fields=("posm", "vel", "acc", "maxvsig"))
    x0, v0, a = (np.asarray(x, dtype=np.float32).astype(np.float64) for x in (x0, v0, a))  # as stored
    st = O.TimestepState("f64")
    t = 0.0
    dts = []
    for k, acc_acc in enumerate([1e-4, 4e-5, 1e-5, 6e-5, 2e-5, 1e-4, 1e-2]):
        prm.accel_accuracy = acc_acc
        h = O.compute_derivatives(prm, h, "f64", state=st)
        o = O.variable_leapfrog(prm, h, st, not_first_step=(k > 0))
        h["posm"], h["vel"] = o["posm"], o["vel"]
        t += st.current
        dts.append(st.current)
        assert np.allclose(h["posm"][0, :3], x0 + v0 * t + 0.5 * a * t * t, rtol=0, atol=1e-13)
        assert np.allclose(h["vel"][0, :3], v0 + a * (t - 0.5 * st.current), rtol=0, atol=1e-13)
    assert dts[0] == pytest.approx(0.002, rel=1e-7)  # integrate(pb, false): initial_timestep
    assert dts[2] < dts[1] < dts[3]  # shrinks, grows back
    assert dts[-1] == pytest.approx(0.02, rel=1e-7)  # max_timestep binds


# ---------------------------------------------------------------- invariants
def test_invariants_two_body():
    prm = ics.plummer_params(H=0.5)
    h = make(np.array([[0.0, 0, 0], [1.0, 0, 0]]), np.array([1.0, 3.0]), vel=np.array([[0, 1.0, 0], [0, -0.5, 2.0]]),
             fields=("posm", "pos64", "vel", "acc"))
    inv = O.invariants(prm, h, phi=O.potential(prm, h))
    assert inv[g.INV_EPOT] == pytest.approx(-3.0 / math.sqrt(1.25), rel=1e-12)
    assert inv[g.INV_EKIN] == pytest.approx(0.5 * 1 + 0.5 * 3 * (0.25 + 4), rel=1e-12)
    assert np.allclose(inv[[g.INV_PX, g.INV_PY, g.INV_PZ]], [0, -0.5, 6.0])
    assert np.allclose(inv[[g.INV_LX, g.INV_LY, g.INV_LZ]], [0, -6.0, -1.5])
    assert inv[g.INV_MASS] == 4.0


def test_fp32_oracle_tracks_fp64():
    h, prm = ics.collision(2000, seed=9)
    o64 = O.compute_derivatives(prm, h, "f64")
    o32 = O.compute_derivatives(prm, h, "f32")
    a64, a32 = o64["acc"][:, :3], o32["acc"][:, :3]
    assert np.linalg.norm(a32 - a64) / np.linalg.norm(a64) < 1e-5
    assert np.allclose(o32["density"], o64["density"], rtol=1e-5)


def test_sph_passes_need_sph_arrays():
    """A state of gravity alone (positions, velocities, accelerations) has no
    SPH arrays: the oracle's SPH passes refuse it instead of writing through
    null pointers; its acceleration pass runs."""
    h, prm = ics.plummer(256, seed=2)
    s = np.arange(16)
    O.calc_acceleration(prm, h, "f32", s)
    for fn in (O.calc_density, O.calc_balsara):
        with pytest.raises(RuntimeError):
            fn(prm, h, "f32", s)
