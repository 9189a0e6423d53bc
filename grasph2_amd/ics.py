"""Seeded synthetic initial conditions and parameters of the benchmark workloads.

Every generator is deterministic for a given seed (numpy PCG64) and returns
host arrays in the C-ABI layout (see grasph2_amd.FIELDS) plus the Params.

Workloads (names used by bench.py and the tests):
  plummer    Plummer sphere N-body, N=16384, M=G=a=1, softening H=0.01,
             dt=1e-3 (recipe of src/initialConditions/particleSources/
             PlummerSphere.cu:35-88: radius from the inverted mass profile,
             speed by von Neumann rejection on q^2 (1-q^2)^3.5)
  liquid     SPH liquid blob with self-gravity, N=65536 on a jittered cubic
             lattice in the unit sphere, rho0=1, p=c^2(rho-rho0), c=1,
             H=2.5 spacing, alpha=1, dt=1e-4
  collision  two solid spheres of 65536 each, R=1, relative speed 0.5 at a
             30 deg impact angle, mu=1, K=2, Y=0.05 (von Mises), Balsara,
             alpha=1, self-gravity, dt=1e-3
  contact    the collision with the spheres already overlapping by 2 H
             (the contact phase: up to ~2x the neighbours in the contact zone)
  ensemble   512 independent collisions, N_k ~ U[1024,16384], angle
             U(0,60) deg, speed U(0.1,1.0), the collision material
  large      one rotating solid sphere, N=1048576, +-1% mass noise, spin at
             0.3 of break-up, solid + von Mises + gravity
"""
import math

import numpy as np

from . import FIELDS, host_particles
from ._capi import (ARTIFICIAL_VISCOSITY, BALSARA_SWITCH, PLASTICITY_MIESE, SELF_GRAVITY, SOLIDS, SPH, Params,
                    default_params)


def _base_params():
    p = default_params()
    p.flags = 0
    p.dimension = 3
    return p


def plummer_params(H=0.01, dt=1e-3):
    p = _base_params()
    p.flags = SELF_GRAVITY
    p.H = H
    p.fixed_timestep = dt
    return p


def solid_params(H, dt=1e-3, gravity=True, rho0=1.0, K=2.0, mu=1.0, Y=0.05, alpha=1.0):
    p = _base_params()
    p.flags = SPH | SOLIDS | PLASTICITY_MIESE | ARTIFICIAL_VISCOSITY | BALSARA_SWITCH
    if gravity:
        p.flags |= SELF_GRAVITY
    p.H = H
    p.rho0 = rho0
    p.BULK = K
    p.dBULKdP = 1.0
    p.SOUNDSPEED = math.sqrt(K / rho0)
    p.shear = mu
    p.Y = Y
    p.alpha = alpha
    p.fixed_timestep = dt
    return p


def liquid_params(H, dt=1e-4, gravity=True, rho0=1.0, c=1.0, alpha=1.0):
    p = _base_params()
    p.flags = SPH | ARTIFICIAL_VISCOSITY
    if gravity:
        p.flags |= SELF_GRAVITY
    p.H = H
    p.rho0 = rho0
    p.SOUNDSPEED = c
    p.alpha = alpha
    p.fixed_timestep = dt
    return p


def _fill(h, pos, mass, vel):
    n = pos.shape[0]
    h["posm"][:, :3] = pos
    h["posm"][:, 3] = mass
    if "pos64" in h:
        # FP64 masters hold the same values as the FP32 mirror at t = 0
        h["pos64"][:] = pos.astype(np.float32).astype(np.float64).T
    if "vel" in h:
        h["vel"][:, :3] = vel
    return h


def lattice_sphere(n, radius=1.0, jitter=0.0, rng=None):
    """n points of a cubic lattice filling a sphere: the n lattice sites
    closest to the centre, spacing chosen so the sphere holds n sites
    (no points and spacing None for n = 0)."""
    if n <= 0:
        return np.zeros((0, 3)), None
    d = (4.0 / 3.0 * math.pi * radius ** 3 / n) ** (1.0 / 3.0)
    m = int(math.ceil(radius / d)) + 2
    ax = (np.arange(-m, m + 1) * d)
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), axis=-1).reshape(-1, 3)
    r2 = (g ** 2).sum(1)
    order = np.lexsort((np.arange(len(r2)), r2))[:n]
    pts = g[order]
    if jitter > 0 and rng is not None:
        pts = pts + rng.uniform(-jitter * d, jitter * d, size=pts.shape)
    return pts, d


def plummer(n=16384, seed=1, cutoff=10.0, H=0.01, dt=1e-3, fields=None):
    """Plummer sphere, M=G=a=1, centre of mass and momentum removed."""
    rng = np.random.default_rng(seed)
    a = 1.0
    r = np.empty(n)
    k = 0
    while k < n:
        X = rng.uniform(0.0, 1.0, size=2 * (n - k))
        X = X[X > 0]
        rr = a / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
        rr = rr[rr <= cutoff]
        take = min(len(rr), n - k)
        r[k:k + take] = rr[:take]
        k += take

    def iso(count):
        u = rng.uniform(-1.0, 1.0, size=count)
        phi = rng.uniform(0.0, 2 * math.pi, size=count)
        s = np.sqrt(1 - u * u)
        return np.stack([s * np.cos(phi), s * np.sin(phi), u], axis=1)

    pos = iso(n) * r[:, None]
    q = np.empty(n)
    k = 0
    while k < n:
        x = rng.uniform(0.0, 1.0, size=4 * (n - k))
        y = rng.uniform(0.0, 0.1, size=4 * (n - k))
        acc = x[y < x * x * (1 - x * x) ** 3.5]
        take = min(len(acc), n - k)
        q[k:k + take] = acc[:take]
        k += take
    vesc = np.sqrt(2.0) * (1.0 + r * r) ** (-0.25)
    vel = iso(n) * (q * vesc)[:, None]
    mass = np.full(n, 1.0 / n)
    pos -= (pos * mass[:, None]).sum(0) / mass.sum()
    vel -= (vel * mass[:, None]).sum(0) / mass.sum()
    h = host_particles(n, fields or ("posm", "pos64", "vel", "acc"))
    return _fill(h, pos, mass, vel), plummer_params(H, dt)


def liquid_blob(n=65536, seed=2, dt=1e-4, gravity=True, fields=None):
    rng = np.random.default_rng(seed)
    pos, d = lattice_sphere(n, 1.0, jitter=0.1, rng=rng)
    rho0 = 1.0
    mass = np.full(n, rho0 * d ** 3)
    vel = np.zeros((n, 3))
    f = fields or ("posm", "pos64", "vel", "acc", "density", "maxvsig", "xvel")
    h = host_particles(n, f)
    h = _fill(h, pos, mass, vel)
    if "density" in h:
        h["density"][:] = rho0
    return h, liquid_params(2.5 * d, dt, gravity=gravity, rho0=rho0)


SOLID_FIELDS = ("posm", "pos64", "vel", "acc", "density", "balsara", "maxvsig", "dstress", "dstress_dt",
                "density_dt")


def _collision_arrays(n_total, rng, speed, angle_deg, jitter=0.1, gap=1.0):
    """gap: surface separation of the two spheres along the centre line in
    units of H (1: just out of reach; negative: overlapping by -gap H)"""
    na = n_total // 2
    nb = n_total - na
    pa, da = lattice_sphere(na, 1.0, jitter=jitter, rng=rng)
    pb, db = lattice_sphere(nb, 1.0, jitter=jitter, rng=rng)
    d = da if da is not None else db
    H = 2.5 * d
    sep = 2.0 + gap * H
    pa = pa + np.array([-sep / 2, 0.0, 0.0])
    pb = pb + np.array([sep / 2, 0.0, 0.0])
    th = math.radians(angle_deg)
    u = np.array([math.cos(th), math.sin(th), 0.0])
    ma, mb = float(na), float(nb)
    va = u * speed * mb / (ma + mb)
    vb = -u * speed * ma / (ma + mb)
    pos = np.concatenate([pa, pb])
    vel = np.concatenate([np.tile(va, (na, 1)), np.tile(vb, (nb, 1))])
    return pos, vel, d, H


def collision(n=131072, seed=3, speed=0.5, angle_deg=30.0, dt=1e-3, gravity=True, fields=None, jitter=0.1, gap=1.0):
    rng = np.random.default_rng(seed)
    pos, vel, d, H = _collision_arrays(n, rng, speed, angle_deg, jitter, gap)
    rho0 = 1.0
    mass = np.full(n, rho0 * d ** 3)
    h = host_particles(n, fields or SOLID_FIELDS)
    h = _fill(h, pos, mass, vel)
    if "density" in h:
        h["density"][:] = rho0
    return h, solid_params(H, dt, gravity=gravity, rho0=rho0)


def ensemble(count=512, nmin=1024, nmax=16384, seed=4, dt=1e-3, fields=None):
    """Instances of the collision workload; returns (list of (host, params), offsets)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=count)
    out = []
    for k in range(count):
        ang = rng.uniform(0.0, 60.0)
        spd = rng.uniform(0.1, 1.0)
        h, p = collision(int(sizes[k]), seed=int(rng.integers(1 << 31)), speed=spd, angle_deg=ang, dt=dt,
                         fields=fields)
        out.append((h, p))
    offsets = np.zeros(count + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)
    return out, offsets


def large_sphere(n=1048576, seed=5, dt=1e-3, noise=0.01, spin=0.3, fields=None):
    rng = np.random.default_rng(seed)
    pos, d = lattice_sphere(n, 1.0, jitter=0.0)
    rho0 = 1.0
    mass = rho0 * d ** 3 * (1.0 + rng.uniform(-noise, noise, size=n))
    M = mass.sum()
    omega = spin * math.sqrt(M / 1.0 ** 3)
    vel = np.cross(np.array([0.0, 0.0, omega])[None, :], pos)
    h = host_particles(n, fields or SOLID_FIELDS)
    h = _fill(h, pos, mass, vel)
    if "density" in h:
        h["density"][:] = rho0
    return h, solid_params(2.5 * d, dt, gravity=True, rho0=rho0)


def contact(n=131072, seed=3, **kw):
    """The collision workload already in contact: the spheres overlap by 2 H
    along the centre line, so the contact zone holds up to ~2x the usual
    neighbour count (148 at N = 131072 against ~69 in the bulk)."""
    return collision(n, seed=seed, gap=-2.0, **kw)


def contact_zone(h, count):
    """`count` particles nearest to the midpoint between the two spheres of a
    collision state (the first half of the particles is one sphere): the
    particles that interact across the contact."""
    pos = np.asarray(h["posm"][:, :3], dtype=np.float64)
    n = len(pos)
    mid = 0.5 * (pos[:n // 2].mean(0) + pos[n // 2:].mean(0))
    dist = np.linalg.norm(pos - mid, axis=1)
    return np.sort(np.argsort(dist, kind="stable")[:count])


def workload(name, **kw):
    return {"plummer": plummer, "liquid": liquid_blob, "collision": collision, "contact": contact,
            "large": large_sphere}[name](**kw)
