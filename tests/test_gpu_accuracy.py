"""Accuracy of every workload at its real size against the FP64 oracle, and
the energy drift of its run.

acceleration: the HIP path's acc after one computeDerivatives against the
FP64 oracle evaluated on a sample of particles through the whole chain
(density on the sample's neighbours' neighbours, strain / Balsara on its
neighbours, acceleration on the sample), errors as in tests/_util.py, bound
1e-5 (the accuracy bound).  The FP32 oracle on the same sample is reported beside it:
that is the reference's own FP32 arithmetic.
energy: FP64 invariants (kinetic with the half-kick velocities synchronised,
softened gravitational potential, compressional + elastic) over the
workload's step count with integrate(pb, true) as the reference's main loop;
the drift is reported and bounded loosely (viscosity and plasticity
dissipate, so SPH runs do not conserve it).  Results go to stdout and, when
the environment variable GRASPH_RESULTS_DIR names a folder, to
accuracy.json there.
"""
import json
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree

import grasph2_amd as g
from grasph2_amd import ics
from oracle import pyoracle as O
from tests._util import sampled_oracle, scalar_errors, vec_errors, vec_errors_floored

pytestmark = pytest.mark.gpu

RESULTS = {}


def record(name, **kv):
    RESULTS.setdefault(name, {}).update(kv)
    print("[accuracy]", name, json.dumps(kv), flush=True)
    out = os.environ.get("GRASPH_RESULTS_DIR")
    if out and os.path.isdir(out):
        with open(os.path.join(out, "accuracy.json"), "w") as fh:
            json.dump(RESULTS, fh, indent=1)


def compact_sample(pos, count, seed):
    """`count` particles nearest to a random particle: a compact cluster keeps
    the oracle's dependency rings small at large N."""
    rng = np.random.default_rng(seed)
    c = pos[rng.integers(len(pos))]
    d = np.linalg.norm(pos - c, axis=1)
    return np.sort(np.argsort(d, kind="stable")[:count])


WORKLOADS = [
    # name, builder, sample size, compact, steps for the drift
    ("plummer", lambda: ics.plummer(16384), 2048, False, 100),
    ("liquid", lambda: ics.liquid_blob(65536), 1024, False, 200),
    ("collision", lambda: ics.collision(131072), 1024, False, 100),
    ("large", lambda: ics.large_sphere(1048576), 4096, True, 20),
]


@pytest.mark.parametrize("name,build,ns,compact,steps", WORKLOADS, ids=[w[0] for w in WORKLOADS])
def test_workload_accuracy(gpu_ctx, name, build, ns, compact, steps):
    h, prm = build()
    n = h["posm"].shape[0]
    O.set_threads(min(16, os.cpu_count() or 1))
    d = g.DeviceParticles.from_host(h)
    gpu_ctx.compute_derivatives(d, prm)
    g.hip.synchronize()
    got = d.download()
    pos = h["posm"][:, :3].astype(np.float64)
    sample = compact_sample(pos, ns, 1) if compact else np.sort(
        np.random.default_rng(1).choice(n, size=ns, replace=False))
    ref64 = sampled_oracle(prm, h, sample, "f64")
    ref32 = sampled_oracle(prm, h, sample, "f32")
    mx, nr = vec_errors(got["acc"][sample], ref64["acc"][sample])
    mx32, nr32 = vec_errors(ref32["acc"][sample], ref64["acc"][sample])
    res = {"n": n, "sample": int(ns), "acc_max_rel": mx, "acc_norm_rel": nr, "fp32_oracle_acc_max_rel": mx32,
           "fp32_oracle_acc_norm_rel": nr32}
    if prm.flags & g.SPH:
        dmx, dnr = scalar_errors(got["density"][sample], ref64["density"][sample])
        res.update(density_max_rel=dmx, density_norm_rel=dnr)
    # energy over the run (reference main loop: integrate(pb, true) every step)
    dt = prm.fixed_timestep
    e = []
    for k in range(steps + 1):
        if k > 0:
            gpu_ctx.integrate(d, prm, True)
            gpu_ctx.compute_derivatives(d, prm)
        if k == 0 or k == steps or k % max(1, steps // 4) == 0:
            inv = gpu_ctx.invariants(d, prm, half_kick_dt=0.5 * dt, with_potential=True)
            e.append(inv[g.INV_EKIN] + inv[g.INV_EPOT] + inv[g.INV_EINT])
    d.free()
    drift = abs(e[-1] - e[0]) / abs(e[0])
    res.update(steps=steps, energy_start=e[0], energy_end=e[-1], energy_drift_rel=drift)
    record(name, **res)
    assert nr <= 1e-5, res
    assert mx <= 1e-5, res
    if name == "plummer":
        assert drift < 1e-3, res


def plastic_share(st, prm):
    """share of particles the next drift carries past von Mises yield, 3 J2(S + dS/dt dt) > Y^2"""
    S = st["dstress"].astype(np.float64) + st["dstress_dt"].astype(np.float64) * float(prm.fixed_timestep)
    J2 = 0.5 * (S[0] ** 2 + S[1] ** 2 + S[2] ** 2) + S[3] ** 2 + S[4] ** 2 + S[5] ** 2
    return float(np.mean(3.0 * J2 > float(prm.Y) ** 2))


def test_large_column_batches(gpu_ctx):
    """The large sphere (N = 1048576) with 1 GiB for the far field's partner
    partials (of the 6.4 GB the whole far field takes): seven column batches
    of partner groups with running sums between them.  Two compact samples
    against the FP64 oracle (1e-5 norm-wise; per particle 1e-5 on the first,
    the unbatched far field's error on the second), and the whole
    acceleration against the unbatched far field's to FP32 rounding (same
    order, same lists)."""
    print("large sphere, far field in column batches", flush=True)
    h, prm = ics.large_sphere(1048576)
    O.set_threads(min(16, os.cpu_count() or 1))
    d = g.DeviceParticles.from_host(h)
    # both evaluations sort the same positions into the same slot order, so
    # the neighbour sums agree bit for bit whatever the calls before were
    interval = gpu_ctx.get_option("reorder_interval")
    budget = gpu_ctx.get_option("sym_partial_bytes")
    gpu_ctx.set_option("reorder_interval", 1)
    try:
        gpu_ctx.compute_derivatives(d, prm)
        g.hip.synchronize()
        base = d.download(["acc", "density"])
        gpu_ctx.set_option("sym_partial_bytes", 1 << 30)
        gpu_ctx.compute_derivatives(d, prm)
        g.hip.synchronize()
    finally:
        gpu_ctx.set_option("sym_partial_bytes", budget)
        gpu_ctx.set_option("reorder_interval", interval)
    got = d.download(["acc", "density"])
    d.free()
    assert np.array_equal(got["density"], base["density"])
    mxb, nrb = vec_errors(got["acc"], base["acc"])
    pos = h["posm"][:, :3].astype(np.float64)
    res = {"n": 1048576, "vs_unbatched_max_rel": mxb, "vs_unbatched_norm_rel": nrb}
    # the sample of the unbatched workload check (1e-5 per particle), and a
    # second cluster, where a particle of small |a| has a per-particle error
    # of 1.3e-5 with the unbatched far field too: held there to the norm-wise
    # 1e-5 and to the unbatched per-particle error
    bounds = []
    for key, ns, seed in (("sample1", 4096, 1), ("sample2", 1024, 2)):
        sample = compact_sample(pos, ns, seed)
        ref64 = sampled_oracle(prm, h, sample, "f64")
        mx, nr = vec_errors(got["acc"][sample], ref64["acc"][sample])
        fl = vec_errors_floored(got["acc"][sample], ref64["acc"][sample])
        mxu, _ = vec_errors(base["acc"][sample], ref64["acc"][sample])
        res.update({key + "_size": ns, key + "_acc_max_rel": mx, key + "_acc_norm_rel": nr,
                    key + "_acc_floored": fl, key + "_unbatched_acc_max_rel": mxu})
        bounds.append((nr, mx, 1e-5 if key == "sample1" else max(1e-5, 1.1 * mxu)))
    record("large_column_batches", **res)
    for nr, mx, mx_bound in bounds:
        assert nr <= 1e-5 and mx <= mx_bound, res
    assert nrb <= 1e-6 and mxb <= 1e-5, res


@pytest.mark.parametrize("start", ["overlap", "impact"])
def test_collision_contact_phase(gpu_ctx, start):
    """The collision workload (N = 131072) in contact, where the neighbour
    counts, the stresses and plasticity peak: either started with the
    spheres overlapping by 2 H ("overlap"), or carried through impact from
    the standard start by 600 steps of the reference's main loop
    ("impact").  The contact zone (1024 particles nearest the midpoint of
    the two centres of mass) against the FP64 oracle: at 1e-5 per particle
    for the overlap; after impact at 1e-5 norm-wise and 1e-4 per particle
    (floored at a tenth of the median |a|, against FP64 at the FP32
    positions the kernels read).  The neighbour statistics, the plastic
    share and a 100-step energy drift are recorded."""
    print("contact phase, start=%s" % start, flush=True)
    h, prm = ics.contact(131072) if start == "overlap" else ics.collision(131072)
    O.set_threads(min(16, os.cpu_count() or 1))
    d = g.DeviceParticles.from_host(h)
    if start == "impact":
        for k in range(600):
            gpu_ctx.compute_derivatives(d, prm)
            gpu_ctx.integrate(d, prm, True)
            if (k + 1) % 200 == 0:
                g.hip.synchronize()
                print("  advanced %d steps" % (k + 1), flush=True)
    state = d.download()
    gpu_ctx.compute_derivatives(d, prm)
    g.hip.synchronize()
    got = d.download()
    pairs, mxn, over = gpu_ctx.neighbour_stats()
    sample = ics.contact_zone(state, 1024)
    ref64 = sampled_oracle(prm, state, sample, "f64")
    ref32 = sampled_oracle(prm, state, sample, "f32")
    # FP64 arithmetic at the FP32 mirror the kernels read (the arithmetic
    # error alone; after 600 FP64 drifts the mirror differs from the masters)
    mirror = {k: v for k, v in state.items() if k != "pos64"}
    refm = sampled_oracle(prm, mirror, sample, "f64")
    mx, nr = vec_errors(got["acc"][sample], ref64["acc"][sample])
    mxm, nrm = vec_errors(got["acc"][sample], refm["acc"][sample])
    flm = vec_errors_floored(got["acc"][sample], refm["acc"][sample])
    mx32, nr32 = vec_errors(ref32["acc"][sample], ref64["acc"][sample])
    dmx, dnr = scalar_errors(got["density"][sample], ref64["density"][sample])
    pos = state["posm"][:, :3].astype(np.float64)
    tree = cKDTree(pos)
    na = len(pos) // 2
    cross = [sum(1 for j in tree.query_ball_point(pos[i], float(prm.H)) if (j < na) != (i < na)) for i in sample]
    res = {"n": len(pos), "start": start, "sample": len(sample), "acc_max_rel": mx, "acc_norm_rel": nr,
           "mirror_acc_max_rel": mxm, "mirror_acc_norm_rel": nrm, "mirror_acc_floored": flm,
           "fp32_oracle_acc_max_rel": mx32, "fp32_oracle_acc_norm_rel": nr32, "density_max_rel": dmx,
           "density_norm_rel": dnr, "neighbour_pairs": pairs, "max_neighbours_seen": mxn, "list_overflows": over,
           "sample_mean_cross_neighbours": float(np.mean(cross)), "plastic_share": plastic_share(got, prm),
           "sample_max_density": float(got["density"][sample].max())}
    dt = prm.fixed_timestep
    e = []
    for k in range(101):
        if k > 0:
            gpu_ctx.integrate(d, prm, True)
            gpu_ctx.compute_derivatives(d, prm)
        if k % 25 == 0:
            inv = gpu_ctx.invariants(d, prm, half_kick_dt=0.5 * dt, with_potential=True)
            e.append(inv[g.INV_EKIN] + inv[g.INV_EPOT] + inv[g.INV_EINT])
    d.free()
    res.update(steps=100, energy_start=e[0], energy_end=e[-1], energy_drift_rel=abs(e[-1] - e[0]) / abs(e[0]))
    record("collision_contact_" + start, **res)
    assert np.mean(cross) > 5, res  # the sample does interact across the contact
    assert dnr <= 1e-5 and dmx <= 1e-5, res
    if start == "overlap":
        assert nr <= 1e-5 and mx <= 1e-5, res
    else:
        # 600 steps in, the bodies are crushed: densities up to 4 rho0, up to
        # ~2700 neighbours per particle, nearly every particle past yield.
        # The reference's own FP32 arithmetic is off by ~1% there; this path
        # holds the norm-wise bound, and per particle (floored at a tenth of
        # the median) stays within 1e-4 of FP64 at the same FP32 positions.
        assert nr <= 1e-5 and nrm <= 1e-5 and flm <= 1e-4, res
