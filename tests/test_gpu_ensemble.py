"""The batched ensemble at its real shape: 512 independent solid collisions
with N_k ~ U[1024, 16384] (sum N = 4.55M), autotuned, through the batched
derivative pass and the graph-captured batched step.

- acceleration and density of a spread of instances (the smallest, the
  largest and four at the size quartiles) against the FP64 oracle on 256
  sampled particles each, at the accuracy bound 1e-5;
- two graph-replayed steps equal two eager batched steps bit for bit;
- the 100-step energy drift of those instances (FP64 invariants on each
  instance's sub-view of the batch).
Results are printed and, when the environment variable GRASPH_RESULTS_DIR
names a folder, written to ensemble_accuracy.json there.
"""
import json
import os

import numpy as np
import pytest

import grasph2_amd as g
from grasph2_amd import ics
from oracle import pyoracle as O
from tests._util import sampled_oracle, scalar_errors, sub_view, vec_errors

pytestmark = pytest.mark.gpu

FIELDS = ("posm", "pos64", "vel", "acc", "density", "balsara", "maxvsig", "dstress", "dstress_dt")


def energy(ctx, view, prm, dt):
    inv = ctx.invariants(view, prm, half_kick_dt=0.5 * dt, with_potential=True)
    return inv[g.INV_EKIN] + inv[g.INV_EPOT] + inv[g.INV_EINT]


def test_ensemble_512_at_shape():
    print("ensemble 512: building", flush=True)
    inst, _ = ics.ensemble(512)
    hosts = [h for h, _ in inst]
    params = [p for _, p in inst]
    del inst
    host, offsets = g.pack_batch(hosts)
    sizes = np.diff(offsets)
    order = np.argsort(sizes, kind="stable")
    picks = sorted({int(order[0]), int(order[-1])} |
                   {int(order[int(q * (len(order) - 1))]) for q in (0.25, 0.5, 0.75)} | {int(order[-2])})
    O.set_threads(min(16, os.cpu_count() or 1))
    ctx = g.Context(0)
    b = ctx.batch(offsets, params)
    d = g.DeviceParticles.from_host(host)
    ctx.autotune(d, b)
    print("ensemble 512: tuning", ctx.batch_tuning(), flush=True)
    ctx.batch_compute_derivatives(d, b)
    g.hip.synchronize()
    got = g.unpack_batch(d.download(), offsets)
    res = {"instances": 512, "particles": int(offsets[-1]), "tuning": ctx.batch_tuning(), "checked": {}}
    rng = np.random.default_rng(12)
    worst = 0.0
    for k in picks:
        n = int(sizes[k])
        sample = np.sort(rng.choice(n, size=min(256, n), replace=False))
        ref = sampled_oracle(params[k], hosts[k], sample, "f64")
        mx, nr = vec_errors(got[k]["acc"][sample], ref["acc"][sample])
        dmx, dnr = scalar_errors(got[k]["density"][sample], ref["density"][sample])
        res["checked"][k] = {"n": n, "acc_max_rel": mx, "acc_norm_rel": nr, "density_max_rel": dmx,
                             "density_norm_rel": dnr}
        print("  instance %d (N = %d): acc max %.3g norm %.3g, density max %.3g" % (k, n, mx, nr, dmx), flush=True)
        worst = max(worst, mx, nr, dmx, dnr)
    # graph replays against eager batched steps, bit for bit, from the same state
    dg = g.DeviceParticles.from_host(host)
    de = g.DeviceParticles.from_host(host)
    for _ in range(2):
        ctx.batch_step(dg, b, 1)
        ctx.batch_compute_derivatives(de, b)
        ctx.batch_integrate(de, b, True)
    g.hip.synchronize()
    a, e = dg.download(FIELDS), de.download(FIELDS)
    for f in FIELDS:
        assert np.array_equal(a[f], e[f]), f
    de.free()
    dg.free()
    # energy over 100 graph steps (the first evaluation on the initial state)
    dt = params[0].fixed_timestep
    v = d.view()
    views = {k: sub_view(v, int(offsets[k]), int(sizes[k])) for k in picks}
    e0 = {k: energy(ctx, views[k], params[k], dt) for k in picks}
    ctx.batch_step(d, b, 100)
    g.hip.synchronize()
    ctx.batch_compute_derivatives(d, b)  # accelerations of the final state, for the half-kick synchronisation
    g.hip.synchronize()
    for k in picks:
        e1 = energy(ctx, views[k], params[k], dt)
        res["checked"][k].update(energy_start=e0[k], energy_end=e1, energy_drift_rel=abs(e1 - e0[k]) / abs(e0[k]))
    d.free()
    ctx.close()
    print("[ensemble]", json.dumps(res), flush=True)
    out = os.environ.get("GRASPH_RESULTS_DIR")
    if out and os.path.isdir(out):
        with open(os.path.join(out, "ensemble_accuracy.json"), "w") as fh:
            json.dump(res, fh, indent=1)
    assert worst <= 1e-5, res
