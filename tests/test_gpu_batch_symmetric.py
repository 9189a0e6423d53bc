"""The symmetric sweep over a batch (option batch_symmetric = 1, for the
instances of at least batch_symmetric_min particles): per instance, gravity
between groups of 1024 is evaluated once per pair, the rest through the
candidate sweep; smaller instances stay with the persistent one-sided
sweep.  Checked against the one-sided batched sweep (same neighbour lists, so the list passes agree bit for bit; gravity
to FP32 rounding), against the FP64 oracle for one instance with several
groups, for graph replays against eager steps, and for run-to-run
determinism.  Sizes cover instances without far work (<= 1024 particles),
ragged last super-blocks and several groups."""
import numpy as np
import pytest

import grasph2_amd as g
from oracle import pyoracle as O
from tests._util import vec_errors
from tests.test_gpu_batch import FIELDS_S, make_batch

pytestmark = pytest.mark.gpu

SIZES = (300, 5000, 1024, 2500, 1025, 4100, 64, 1, 3333)


@pytest.fixture
def sym_batch_ctx(gpu_ctx):
    """Instances of 1000 particles or more take the symmetric sweep, the
    smaller ones (300, 64, 1) the one-sided persistent sweep, in one batch."""
    gpu_ctx.set_option("batch_symmetric", 1)
    gpu_ctx.set_option("batch_symmetric_min", 1000)
    gpu_ctx.set_option("reorder_interval", 1)  # graph, eager and repeat runs refresh the order alike
    yield gpu_ctx
    gpu_ctx.set_option("batch_symmetric", 1)  # the defaults
    gpu_ctx.set_option("batch_symmetric_min", 0)
    gpu_ctx.set_option("reorder_interval", 8)


def run_batch(ctx, host, offsets, params):
    b = ctx.batch(offsets, params)
    d = g.DeviceParticles.from_host(host)
    ctx.batch_compute_derivatives(d, b)
    g.hip.synchronize()
    out = g.unpack_batch(d.download(), offsets)
    d.free()
    return out


@pytest.mark.parametrize("gravity_only", [False, True])
def test_matches_one_sided_batch_and_fp64(sym_batch_ctx, gravity_only):
    print("batch symmetric gravity_only=%s" % gravity_only, flush=True)
    ctx = sym_batch_ctx
    insts, params, host, offsets = make_batch(sizes=SIZES, seed=5)
    if gravity_only:
        for p in params:
            p.flags = g.SELF_GRAVITY
    got = run_batch(ctx, host, offsets, params)
    ctx.set_option("batch_symmetric", 0)
    ref = run_batch(ctx, host, offsets, params)
    ctx.set_option("batch_symmetric", 1)
    for k, (a, r) in enumerate(zip(got, ref)):
        if SIZES[k] > 1:
            mx, nr = vec_errors(a["acc"], r["acc"])
            assert nr <= 1e-6 and mx <= 1e-5, (k, mx, nr)
        if not gravity_only:
            assert np.array_equal(a["density"], r["density"]), k  # same lists, same order
            assert np.array_equal(a["dstress_dt"], r["dstress_dt"]), k
    o = O.compute_derivatives(params[1], insts[1], "f64")
    mx, nr = vec_errors(got[1]["acc"], o["acc"])
    assert nr <= 1e-5 and mx <= 1e-5, (mx, nr)


def test_large_instance_flushes_fragments_twice(sym_batch_ctx):
    """An instance of 40000 particles (40 groups): its first rows' pieces
    hold 10 partner groups, 20 far units, more than a wavefront folds
    between flushes (16), so the batch's fragments carry corrections."""
    print("batch symmetric, instance of 40000", flush=True)
    ctx = sym_batch_ctx
    sizes = (40000, 3000, 500)
    insts, params, host, offsets = make_batch(sizes=sizes, seed=9)
    for p in params:
        p.flags = g.SELF_GRAVITY
    got = run_batch(ctx, host, offsets, params)
    ctx.set_option("batch_symmetric", 0)
    ref = run_batch(ctx, host, offsets, params)
    ctx.set_option("batch_symmetric", 1)
    for k, (a, r) in enumerate(zip(got, ref)):
        mx, nr = vec_errors(a["acc"], r["acc"])
        assert nr <= 1e-6 and mx <= 1e-5, (k, mx, nr)
    idx = np.random.default_rng(1).choice(sizes[0], 2000, replace=False)  # FP64 on a sample
    o = O.calc_acceleration(params[0], insts[0], "f64", sample=idx)
    mx, nr = vec_errors(got[0]["acc"][idx], o["acc"][idx])
    assert nr <= 1e-5 and mx <= 1e-5, (mx, nr)


def test_graph_steps_match_eager_and_repeat(sym_batch_ctx):
    print("batch symmetric graph", flush=True)
    ctx = sym_batch_ctx
    insts, params, host, offsets = make_batch(sizes=SIZES, seed=6)
    b = ctx.batch(offsets, params)
    dg = g.DeviceParticles.from_host(host)
    de = g.DeviceParticles.from_host(host)
    for step in range(3):
        ctx.batch_step(dg, b, 1)
        g.hip.synchronize()
        ctx.batch_compute_derivatives(de, b)
        ctx.batch_integrate(de, b, True)
        g.hip.synchronize()
        a, e = dg.download(), de.download()
        for f in FIELDS_S:
            assert np.array_equal(a[f], e[f]), "replay %d field %s" % (step, f)
    dg.free()
    de.free()
    x = run_batch(ctx, host, offsets, params)
    y = run_batch(ctx, host, offsets, params)
    for k, (a, r) in enumerate(zip(x, y)):
        for f in ("acc", "density", "dstress_dt", "maxvsig"):
            assert np.array_equal(a[f], r[f]), (k, f)


def test_autotune_picks_symmetric_split(gpu_ctx, tmp_path):
    """The autotuner also picks from which size bucket on the batch takes the
    symmetric sweep (or none); results stay within FP32 rounding of the
    one-sided batch, the list passes bit for bit, and the cached choice is
    found again."""
    print("batch symmetric autotune", flush=True)
    ctx = gpu_ctx
    insts, params, host, offsets = make_batch(sizes=(1500, 3000, 5000, 9000, 1100, 2200), seed=3)
    ref = run_batch_with(ctx, host, offsets, params, {"batch_symmetric": 0})
    cache = str(tmp_path / "tune.txt")
    try:
        b = ctx.batch(offsets, params)
        d = g.DeviceParticles.from_host(host)
        ctx.autotune(d, b, cache)
        t = ctx.batch_tuning()
        assert t["symmetric_min"] in (None, 0, 2048, 4096, 8192), t
        d.upload(host)
        ctx.batch_compute_derivatives(d, b)
        g.hip.synchronize()
        got = g.unpack_batch(d.download(), offsets)
        for k, (a, r) in enumerate(zip(got, ref)):
            mx, nr = vec_errors(a["acc"], r["acc"])
            assert nr <= 1e-6 and mx <= 1e-5, (k, mx, nr)
            assert np.array_equal(a["density"], r["density"]), k
        # from the cache: same choice
        ctx.set_option("batch_symmetric", 1)
        ctx.set_option("batch_symmetric_min", 0)
        ctx.autotune(d, b, cache)
        assert ctx.batch_tuning() == t
        d.free()
    finally:
        ctx.set_option("batch_symmetric", 1)  # the defaults
        ctx.set_option("batch_symmetric_min", 0)


def run_batch_with(ctx, host, offsets, params, opts):
    old = {k: ctx.get_option(k) for k in opts}
    for k, v in opts.items():
        ctx.set_option(k, v)
    try:
        return run_batch(ctx, host, offsets, params)
    finally:
        for k, v in old.items():
            ctx.set_option(k, v)
