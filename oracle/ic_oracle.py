"""CPU restatement of GraSPH2's initial-condition recipes, in numpy float32.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

What it restates (reference files, paths relative to the GraSPH2 source tree):
  - the deterministic modifiers of a particle source and the desired
    centre of mass / COM velocity they carry
    (src/initialConditions/ParticleSource.h:114-162, 176-184);
  - the per-source centre-of-mass correction of InitGenerator::generate()
    (src/initialConditions/InitGenerator.h:108-150): mass-weighted sums in
    f1_t in particle order, divided by the total mass, the difference to the
    desired value added to every particle;
  - the distributions the random sources sample (closed forms used by the
    statistical tests): the uniform ball of UniformSphere
    (particleSources/UniformSphere.cu:37-44), the truncated Plummer radius
    (particleSources/PlummerSphere.cu:45-58) and the speed fraction of
    addRandomPlummerVelocity (PlummerSphere.cu:60-88).

Pins (tests/test_initial_conditions.py): the rotation matrix against
scipy's rotation-vector construction; the corrected centre of mass against
the desired one; the desired values against the modifiers applied to a
particle at the origin at rest; the Plummer CDF against numerical
integration of the Plummer density.  The sampling of directions and ball
points is mpUtils' in the reference (mpUtils v0.12.0, absent): the single
draws are parity unpinned, their distributions are pinned by the tests.
"""
import math

import numpy as np

f32 = np.float32


def rotation_matrix(axis, angle):
    """R = cos(a) I + (1 - cos(a)) u u^T + sin(a) [u]x for the normalised axis
    (ParticleSource.h:138-152), in float32."""
    u = np.asarray(axis, dtype=f32)
    u = u / f32(math.sqrt(float(np.dot(u.astype(np.float64), u.astype(np.float64)))))
    c, s = f32(math.cos(angle)), f32(math.sin(angle))
    uu = np.outer(u, u).astype(f32)
    ux = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]], dtype=f32)
    return (c * np.eye(3, dtype=f32) + (f32(1) - c) * uu + s * ux).astype(f32)


def apply_ops(pos, vel, density, ops):
    """The modifiers in order (ParticleSource.h:114-184, 199-203).  ops: list of
    ("move", d) | ("rotate", axis, angle) | ("angvel", omega) | ("linvel", v)
    | ("density", rho).  Returns new float32 arrays."""
    pos = np.array(pos, dtype=f32)
    vel = np.array(vel, dtype=f32)
    density = np.array(density, dtype=f32)
    for op in ops:
        kind = op[0]
        if kind == "move":
            pos = pos + np.asarray(op[1], dtype=f32)
        elif kind == "rotate":
            R = rotation_matrix(op[1], op[2])
            pos = (pos @ R.T).astype(f32)
            vel = (vel @ R.T).astype(f32)
        elif kind == "angvel":
            vel = vel + np.cross(np.asarray(op[1], dtype=f32), pos).astype(f32)
        elif kind == "linvel":
            vel = vel + np.asarray(op[1], dtype=f32)
        elif kind == "density":
            density = np.full_like(density, f32(op[1]))
        else:
            raise ValueError(kind)
    return pos, vel, density


def desired_com(ops):
    """Desired COM and COM velocity after the modifiers, starting from the
    origin at rest (ParticleSource.h:119, 130, 141, 158-159)."""
    com = np.zeros(3, dtype=f32)
    vcom = np.zeros(3, dtype=f32)
    for op in ops:
        kind = op[0]
        if kind == "move":
            com = com + np.asarray(op[1], dtype=f32)
        elif kind == "rotate":
            R = rotation_matrix(op[1], op[2])
            com = (R @ com).astype(f32)
            vcom = (R @ vcom).astype(f32)
        elif kind == "angvel":
            vcom = vcom + np.cross(np.asarray(op[1], dtype=f32), com).astype(f32)
        elif kind == "linvel":
            vcom = vcom + np.asarray(op[1], dtype=f32)
    return com, vcom


def correct_com(pos, mass, vel, com_want, vcom_want, fix_pos=True, fix_vel=True):
    """InitGenerator.h:108-150: float32 sums in particle order."""
    total = f32(0)
    com = np.zeros(3, dtype=f32)
    vcom = np.zeros(3, dtype=f32)
    for i in range(len(mass)):
        total = f32(total + mass[i])
        com = (com + mass[i] * pos[i]).astype(f32)
        vcom = (vcom + mass[i] * vel[i]).astype(f32)
    com = (com / total).astype(f32)
    vcom = (vcom / total).astype(f32)
    dpos = (np.asarray(com_want, dtype=f32) - com) if fix_pos else np.zeros(3, dtype=f32)
    dvel = (np.asarray(vcom_want, dtype=f32) - vcom) if fix_vel else np.zeros(3, dtype=f32)
    return (pos + dpos).astype(f32), (vel + dvel).astype(f32)


def uniform_ball_cdf(r, radius):
    """P(|x| <= r) for a point uniform in the ball: (r/R)^3."""
    return np.clip(np.asarray(r, dtype=np.float64) / radius, 0.0, 1.0) ** 3


def plummer_mass_fraction(r, a):
    """m(r)/M = r^3 / (r^2 + a^2)^(3/2), the inverse of the radius draw at
    PlummerSphere.cu:48-55."""
    r = np.asarray(r, dtype=np.float64)
    return r ** 3 / (r * r + a * a) ** 1.5


def plummer_radius_cdf(r, a, cutoff):
    """CDF of the radius after the redraw above the cut-off."""
    return np.clip(plummer_mass_fraction(r, a) / plummer_mass_fraction(cutoff, a), 0.0, 1.0)


def plummer_speed_fraction_cdf(x):
    """CDF of x = v / v_esc accepted by the rejection g(x) = x^2 (1 - x^2)^3.5
    (PlummerSphere.cu:68-77), by numerical integration of g."""
    grid = np.linspace(0.0, 1.0, 20001)
    g = grid ** 2 * (1.0 - grid ** 2) ** 3.5
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (g[1:] + g[:-1]) * np.diff(grid))])
    cum /= cum[-1]
    return np.interp(np.asarray(x, dtype=np.float64), grid, cum)
