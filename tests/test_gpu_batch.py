"""Batched ensembles: the batched launches, the captured-graph steps, the
per-instance launch loop and the autotuned configurations give the same
particles, bit for bit, as the single-instance path on each instance (same
kernels, same summation order per particle); the single-instance path is
itself checked against the oracle in test_gpu_parity.py, and one instance
is checked against the FP64 oracle here too."""
import numpy as np
import pytest

import grasph2_amd as g
from grasph2_amd import ics
from oracle import pyoracle as O
from tests._util import vec_errors

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


def test_batched_derivatives_match_single_instances(gpu_ctx):
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
    assert nr <= 1e-5 and mx <= 1e-4, (mx, nr)


def test_graph_steps_match_loop_and_eager(gpu_ctx):
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


def test_autotune_keeps_results(gpu_ctx, tmp_path):
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
