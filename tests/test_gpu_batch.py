"""Batched ensembles: with the one-sided batched sweep for every instance
(option batch_symmetric = 0), the batched launches, the captured-graph
steps, the per-instance launch loop and the autotuned configurations give
the same particles, bit for bit, as the single-instance path on each
instance (same kernels, same summation order per particle); the default
symmetric batched sweep is checked in test_gpu_batch_symmetric.py.  The
single-instance path is
itself checked against the oracle in test_gpu_parity.py, and one instance
is checked against the FP64 oracle here too."""
import numpy as np
import pytest

import grasph2_amd as g
from grasph2_amd import ics
from oracle import pyoracle as O
from tests._util import vec_errors, vec_errors_floored

pytestmark = pytest.mark.gpu

SIZES = (300, 1000, 1024, 2500, 700, 4100, 64, 1)


def make_batch(sizes=SIZES, seed=0):
    rng = np.random.default_rng(seed)
    insts, params = [], []
    for k, n in enumerate(sizes):
        h, prm = ics.collision(n, seed=100 + k, speed=rng.uniform(0.1, 1.0), angle_deg=rng.uniform(0, 60))
        h["vel"][:, :3] += rng.normal(scale=0.05, size=(n, 3))
        h["dstress"][:] = rng.normal(scale=0.02, size=(6, n))
        h["density"][:] *= 1 + rng.uniform(-0.03, 0.03, size=n)
        insts.append(h)
        params.append(prm)
    host, offsets = g.pack_batch(insts)
    return insts, params, host, offsets


def single_results(ctx, insts, params, steps=0):
    outs = []
    for h, prm in zip(insts, params):
        d = g.DeviceParticles.from_host(h)
        if steps == 0:
            ctx.compute_derivatives(d, prm)
        for _ in range(steps):
            ctx.compute_derivatives(d, prm)
            ctx.integrate(d, prm, True)
        g.hip.synchronize()
        outs.append(d.download())
        d.free()
    return outs


FIELDS_D = ("acc", "density", "dstress_dt", "balsara", "maxvsig")
FIELDS_S = ("posm", "pos64", "vel", "acc", "density", "dstress", "dstress_dt", "balsara", "maxvsig")


def assert_same(got_list, ref_list, fields):
    for k, (a, b) in enumerate(zip(got_list, ref_list)):
        for f in fields:
            assert np.array_equal(a[f], b[f]), "instance %d field %s" % (k, f)


@pytest.fixture
def unsplit(gpu_ctx):
    """The single-instance sweep without its j-split and the batch with its
    one-sided persistent sweep for every instance (which never splits an
    instance's j range), so both sum in the same order; the Morton order
    refreshed at every evaluation, so the paths compared refresh it at the
    same times (it changes the order of the compensated sums); the
    launch-loop contexts take the options from the parent context."""
    gpu_ctx.set_option("sweep_slices", 1)
    gpu_ctx.set_option("batch_symmetric", 0)
    gpu_ctx.set_option("reorder_interval", 1)
    yield gpu_ctx
    gpu_ctx.set_option("sweep_slices", 0)  # the defaults
    gpu_ctx.set_option("batch_symmetric", 1)
    gpu_ctx.set_option("reorder_interval", 8)


def test_batched_derivatives_match_single_instances(unsplit):
    gpu_ctx = unsplit
    insts, params, host, offsets = make_batch()
    b = gpu_ctx.batch(offsets, params)
    d = g.DeviceParticles.from_host(host)
    gpu_ctx.batch_compute_derivatives(d, b)
    g.hip.synchronize()
    got = g.unpack_batch(d.download(), offsets)
    d.free()
    ref = single_results(gpu_ctx, insts, params)
    assert_same(got, ref, FIELDS_D)
    o = O.compute_derivatives(params[3], insts[3], "f64")
    mx, nr = vec_errors(got[3]["acc"], o["acc"])
    fl = vec_errors_floored(got[3]["acc"], o["acc"])
    assert nr <= 1e-5 and fl <= 1e-5 and mx <= 1e-4, (mx, fl, nr)


def test_graph_steps_match_loop_and_eager(unsplit):
    gpu_ctx = unsplit
    insts, params, host, offsets = make_batch(seed=1)
    b = gpu_ctx.batch(offsets, params)
    dg = g.DeviceParticles.from_host(host)
    dl = g.DeviceParticles.from_host(host)
    de = g.DeviceParticles.from_host(host)
    # graph replays interleaved with eager work and synchronisation
    for step in range(3):
        gpu_ctx.batch_step(dg, b, 1)
        g.hip.synchronize()
        gpu_ctx.batch_compute_derivatives(de, b)
        gpu_ctx.batch_integrate(de, b, True)
        g.hip.synchronize()
        a, e = dg.download(), de.download()
        for f in FIELDS_S:
            assert np.array_equal(a[f], e[f]), "replay %d field %s" % (step, f)
    gpu_ctx.batch_step_loop(dl, b, 3)
    g.hip.synchronize()
    got = g.unpack_batch(dg.download(), offsets)
    loop = g.unpack_batch(dl.download(), offsets)
    assert_same(got, loop, FIELDS_S)
    ref = single_results(gpu_ctx, insts, params, steps=3)
    assert_same(got, ref, FIELDS_S)
    for x in (dg, dl, de):
        x.free()


def test_autotune_keeps_results(unsplit, tmp_path):
    gpu_ctx = unsplit  # the one-sided batched sweep, whose launch shape the autotuner picks
    insts, params, host, offsets = make_batch(sizes=(900, 2100, 4500, 9000, 1500), seed=2)
    b = gpu_ctx.batch(offsets, params)
    d = g.DeviceParticles.from_host(host)
    gpu_ctx.batch_compute_derivatives(d, b)
    g.hip.synchronize()
    before = d.download()
    cache = str(tmp_path / "tune.txt")
    gpu_ctx.autotune(d, b, cache)
    t = gpu_ctx.batch_tuning()
    assert t["waves_per_block"] in (1, 2, 4) and all(x in (1, 2) for x in t["ni"])
    d.upload(host)
    gpu_ctx.batch_compute_derivatives(d, b)
    g.hip.synchronize()
    after = d.download()
    for f in FIELDS_D:
        assert np.array_equal(before[f], after[f]), f
    # a second tune reads the cache
    gpu_ctx.autotune(d, b, cache)
    assert gpu_ctx.batch_tuning() == t
    assert open(cache).read().count("\n") == 1
    d.free()


def test_batch_rejects_mixed_flags(gpu_ctx):
    insts, params, host, offsets = make_batch(sizes=(100, 200))
    params[1] = params[1].copy()
    params[1].flags &= ~g.BALSARA_SWITCH
    b = gpu_ctx.batch(offsets, params)
    d = g.DeviceParticles.from_host(host)
    with pytest.raises(g.GrasphError) as e:
        gpu_ctx.batch_compute_derivatives(d, b)
    assert e.value.code == g.E_INVALID
    d.free()


def test_split_single_instance_agrees_with_batch(gpu_ctx):
    """With its j-split the single-instance sweep sums gravity in another
    order: equal to FP32 rounding, not bit for bit."""
    insts, params, host, offsets = make_batch(sizes=(2500, 4100, 9000), seed=3)
    b = gpu_ctx.batch(offsets, params)
    d = g.DeviceParticles.from_host(host)
    gpu_ctx.batch_compute_derivatives(d, b)
    g.hip.synchronize()
    got = g.unpack_batch(d.download(), offsets)
    d.free()
    ref = single_results(gpu_ctx, insts, params)
    for a, r in zip(got, ref):
        mx, nr = vec_errors(a["acc"], r["acc"])
        fl = vec_errors_floored(a["acc"], r["acc"])
        assert nr <= 1e-6 and fl <= 1e-5 and mx <= 1e-4, (mx, fl, nr)
        assert np.array_equal(a["density"], r["density"])  # list passes: same neighbours, same order


def _variable(params, initial=0.05):
    out = []
    for p in params:
        q = p.copy()
        q.flags |= g.VARIABLE_TIMESTEP | g.ACCELERATION_CRITERION | g.VELOCITY_CRITERION
        q.initial_timestep, q.min_timestep, q.max_timestep = initial, 1e-4, 0.1
        q.accel_accuracy, q.velocity_accuracy = 0.01, 0.09
        out.append(q)
    return out


def test_batched_variable_step_matches_single_instances(unsplit):
    """Per-instance dt state: the batched integrator (eager and graph steps)
    equals each instance integrated alone on its own context, dt included;
    every batched derivative call resets each instance's next dt to its
    max_timestep as the single-instance acceleration pass does."""
    gpu_ctx = unsplit
    insts, params, host, offsets = make_batch(sizes=(300, 1000, 2500, 64, 1), seed=4)
    params = _variable(params)
    b = gpu_ctx.batch(offsets, params)
    de = g.DeviceParticles.from_host(host)
    dg = g.DeviceParticles.from_host(host)
    gpu_ctx.batch_compute_derivatives(de, b)
    gpu_ctx.batch_integrate(de, b, False)
    for _ in range(3):
        gpu_ctx.batch_compute_derivatives(de, b)
        gpu_ctx.batch_integrate(de, b, True)
    dts = gpu_ctx.batch_current_timesteps(b)
    got = g.unpack_batch(de.download(), offsets)
    # graph steps on a second batch object (fresh dt state): first step eager, then replays
    b2 = gpu_ctx.batch(offsets, [p.copy() for p in params])
    gpu_ctx.batch_compute_derivatives(dg, b2)
    gpu_ctx.batch_integrate(dg, b2, False)
    gpu_ctx.batch_step(dg, b2, 3)
    g.hip.synchronize()
    assert np.array_equal(gpu_ctx.batch_current_timesteps(b2), dts)
    graph = g.unpack_batch(dg.download(), offsets)
    assert_same(graph, got, FIELDS_S)
    de.free()
    dg.free()
    for k, (h, prm) in enumerate(zip(insts, params)):
        c1 = g.Context(0)
        c1.set_option("sweep_slices", 1)
        c1.set_option("reorder_interval", 1)
        d1 = g.DeviceParticles.from_host(h)
        c1.compute_derivatives(d1, prm)
        c1.integrate(d1, prm, False)
        for _ in range(3):
            c1.compute_derivatives(d1, prm)
            c1.integrate(d1, prm, True)
        assert c1.current_timestep(prm) == dts[k], k
        one = d1.download()
        d1.free()
        c1.close()
        for f in FIELDS_S:
            assert np.array_equal(one[f], got[k][f]), "instance %d field %s" % (k, f)
    assert len(set(dts.tolist())) > 1  # the instances chose their own dt
