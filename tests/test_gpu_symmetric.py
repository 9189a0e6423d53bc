"""The symmetric single-instance sweep (grasph2_amd/csrc/sym_sweep.hip,
option sweep_symmetric = 1): far pairs computed once for both sides, near
pairs through the candidate sweep.  Checked against the FP64 oracle like the
one-sided sweep, against the one-sided sweep itself (the neighbour lists
must be identical, so the list passes agree bit for bit), and for run-to-run
determinism.  Sizes cover one group only (no far field), several groups,
a ragged last super-block and unsoftened gravity."""
import numpy as np
import pytest

import grasph2_amd as g
from grasph2_amd import ics
from oracle import pyoracle as O
from tests._util import scalar_errors, tensor_errors, vec_errors
from tests.test_gpu_parity import TOL, check_vec, gpu_run, liquid_case, solid_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sym_ctx():
    ctx = g.Context(0)
    ctx.set_option("sweep_symmetric", 1)
    # a fresh Morton order at every evaluation: the slot order, and so the
    # rounding of the sums these tests compare bit for bit, depends on the
    # positions alone, not on the evaluations before
    ctx.set_option("reorder_interval", 1)
    yield ctx
    ctx.close()


# far units of 128 partners (each partner super-block cut into 4) for 5000
# and 25000, of 256 for 20000 and 20037; 5000, 20037 and 25000 end in a
# ragged piece
@pytest.mark.parametrize("n", [1, 255, 1000, 5000, 20000, 20037, 25000])
def test_gravity_matches_fp64(sym_ctx, n):
    print("symmetric gravity n=%d" % n, flush=True)
    h, prm = ics.plummer(n, seed=n)
    out = gpu_run(sym_ctx, h, prm)
    if n == 1:
        assert np.all(out["acc"][:, :3] == 0)
        return
    ref = O.calc_acceleration(prm, h, "f64")
    check_vec(out["acc"], ref["acc"])


# Few far workgroups: a workgroup then stays on a row for more units than it
# folds into its LDS sums between flushes (16), so its fragment is flushed
# several times, compensated, with corrections (3 workgroups: ~500 units
# each, several rows; 40: 19 units, two flushes); 25000 on 40 workgroups
# (15 units each) keeps the plain fragments of one flush.
@pytest.mark.parametrize("n,wgs", [(20037, 3), (20037, 40), (25000, 40)])
def test_gravity_few_far_workgroups(sym_ctx, n, wgs):
    print("symmetric gravity n=%d sym_workgroups=%d" % (n, wgs), flush=True)
    h, prm = ics.plummer(n, seed=n + wgs)
    base = gpu_run(sym_ctx, h, prm)
    sym_ctx.set_option("sym_workgroups", wgs)
    try:
        out = gpu_run(sym_ctx, h, prm)
    finally:
        sym_ctx.set_option("sym_workgroups", 0)
    ref = O.calc_acceleration(prm, h, "f64")
    check_vec(out["acc"], ref["acc"])
    mx, nr = vec_errors(out["acc"], base["acc"])
    assert nr <= 1e-6 and mx <= 1e-5, (mx, nr)


# Column batches: with 1 MiB for the partner-side partials, 20037 particles
# (20 groups) take three batches of partner groups, [1, 13), [13, 18) and
# [18, 20), with running sums between them; with 3 far workgroups each
# batch's fragments are also flushed several times (with corrections).
@pytest.mark.parametrize("wgs", [0, 3])
def test_gravity_column_batches(sym_ctx, wgs):
    print("symmetric gravity in column batches, sym_workgroups=%d" % wgs, flush=True)
    h, prm = ics.plummer(20037, seed=21)
    base = gpu_run(sym_ctx, h, prm)
    sym_ctx.set_option("sym_partial_bytes", 1 << 20)
    sym_ctx.set_option("sym_workgroups", wgs)
    try:
        out = gpu_run(sym_ctx, h, prm)
        again = gpu_run(sym_ctx, h, prm)
    finally:
        sym_ctx.set_option("sym_partial_bytes", 12 << 30)
        sym_ctx.set_option("sym_workgroups", 0)
    assert np.array_equal(out["acc"], again["acc"])  # run to run
    ref = O.calc_acceleration(prm, h, "f64")
    check_vec(out["acc"], ref["acc"])
    mx, nr = vec_errors(out["acc"], base["acc"])
    assert nr <= 1e-6 and mx <= 1e-5, (mx, nr)


def test_partial_budget_option(sym_ctx):
    """sym_partial_bytes: read back as set, at least 1 MiB."""
    assert sym_ctx.get_option("sym_partial_bytes") == 12 << 30
    with pytest.raises(g.GrasphError):
        sym_ctx.set_option("sym_partial_bytes", (1 << 20) - 1)
    sym_ctx.set_option("sym_partial_bytes", 1 << 20)
    assert sym_ctx.get_option("sym_partial_bytes") == 1 << 20
    sym_ctx.set_option("sym_partial_bytes", 12 << 30)


def test_sph_column_batches(sym_ctx):
    """The solid case in column batches: the acceleration pass's SPH terms
    join the gravity in the last batch's combine; the neighbour sums are the
    unbatched run's bit for bit."""
    print("symmetric SPH in column batches", flush=True)
    h, prm = solid_case(20000, seed=23)
    base = gpu_run(sym_ctx, h, prm)
    sym_ctx.set_option("sym_partial_bytes", 1 << 20)
    try:
        out = gpu_run(sym_ctx, h, prm)
    finally:
        sym_ctx.set_option("sym_partial_bytes", 12 << 30)
    for f in ("density", "dstress_dt", "balsara", "maxvsig"):
        assert np.array_equal(out[f], base[f]), f
    mx, nr = vec_errors(out["acc"], base["acc"])
    assert nr <= 1e-6 and mx <= 1e-5, (mx, nr)
    ref = O.compute_derivatives(prm, h, "f64")
    check_vec(out["acc"], ref["acc"])


# Equal masses (the mass factored out of both sides' sums), random masses,
# and equal masses but for the last 50 particles (the equal-mass check
# sees the exception at the end).
@pytest.mark.parametrize("uniform", [True, False, "mixed"])
def test_gravity_equal_and_unequal_masses(sym_ctx, uniform):
    print("symmetric masses uniform=%s" % uniform, flush=True)
    h, prm = ics.plummer(6000, seed=11)
    h["posm"][:, 3] = np.float32(1.0 / 6000)
    if uniform is False:
        h["posm"][:, 3] *= np.random.default_rng(5).uniform(0.5, 1.5, size=6000).astype(np.float32)
    elif uniform == "mixed":
        h["posm"][-50:, 3] *= np.float32(1.5)
    for H in (0.0, 0.01):
        prm.H = H
        out = gpu_run(sym_ctx, h, prm)
        ref = O.calc_acceleration(prm, h, "f64")
        check_vec(out["acc"], ref["acc"])


def test_gravity_unsoftened_and_coincident(sym_ctx):
    print("symmetric unsoftened coincident", flush=True)
    h, prm = ics.plummer(3000, seed=3)
    prm.H = 0.0
    h["posm"][2900] = h["posm"][4]  # coincident particles in different groups: r2 = 0 pairs are skipped
    h["pos64"][:, 2900] = h["pos64"][:, 4]
    out = gpu_run(sym_ctx, h, prm)
    assert np.all(np.isfinite(out["acc"]))
    ref = O.calc_acceleration(prm, h, "f64")
    check_vec(out["acc"], ref["acc"])


@pytest.mark.parametrize("case", ["solid", "solid_nograv", "liquid"])
def test_sph_matches_fp64_and_one_sided_lists(gpu_ctx, sym_ctx, case):
    print("symmetric sph case=%s" % case, flush=True)
    if case == "liquid":
        h, prm = liquid_case(n=6000, seed=9)
    else:
        h, prm = solid_case(n=9000, seed=12, gravity=(case == "solid"))
    out = gpu_run(sym_ctx, h, prm)
    one = gpu_run(gpu_ctx, h, prm)
    # identical neighbour lists: the list passes agree bit for bit
    assert np.array_equal(out["density"], one["density"])
    ref = O.compute_derivatives(prm, h, "f64")
    mx, nr = scalar_errors(out["density"], ref["density"])
    assert nr <= TOL and mx <= TOL, ("density", mx, nr)
    check_vec(out["acc"], ref["acc"])
    mx, nr = vec_errors(out["acc"], one["acc"])
    assert nr <= 1e-6 and mx <= 1e-5, ("symmetric vs one-sided", mx, nr)
    if case != "liquid":
        assert np.array_equal(out["dstress_dt"], one["dstress_dt"])
        mx, nr = tensor_errors(out["dstress_dt"], ref["dstress_dt"])
        assert mx <= 10 * TOL and nr <= TOL, ("dstress_dt", mx, nr)


def test_standalone_density_pass(gpu_ctx, sym_ctx):
    print("symmetric standalone density", flush=True)
    h, prm = solid_case(n=7000, seed=21)
    out = gpu_run(sym_ctx, h, prm, op="density")
    one = gpu_run(gpu_ctx, h, prm, op="density")
    assert np.array_equal(out["density"], one["density"])


def test_list_overflow_scan_is_bitwise_equal(sym_ctx):
    """Lists that overflow fall back to the exact scan of the whole instance,
    which walks the same partners in the same order."""
    print("symmetric list overflow", flush=True)
    h, prm = solid_case(n=9000, seed=14)
    base = gpu_run(sym_ctx, h, prm)
    try:
        sym_ctx.set_option("max_neighbours_grow", 0)  # keep the small capacity
        sym_ctx.set_option("max_neighbours", 8)
        small = gpu_run(sym_ctx, h, prm)
    finally:
        sym_ctx.set_option("max_neighbours", 160)
        sym_ctx.set_option("max_neighbours_grow", 1)
    for f in ("acc", "density", "dstress_dt", "balsara", "maxvsig"):
        assert np.array_equal(base[f], small[f]), f


def test_deterministic(sym_ctx):
    print("symmetric determinism", flush=True)
    h, prm = solid_case(n=9000, seed=13)
    a = gpu_run(sym_ctx, h, prm)
    b = gpu_run(sym_ctx, h, prm)
    for f in ("acc", "density", "dstress_dt", "maxvsig"):
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("partial_bytes", [12 << 30, 1 << 20])
def test_step_captured_in_graph(partial_bytes):
    """computeDerivatives + integrate of the symmetric sweep (the far field on
    the caller's stream, the rest forked onto the context's second stream)
    captured into a HIP graph on the caller's stream: every replay, with
    eager work and synchronisation between replays, equals the same step
    run eagerly, bit for bit.  The Morton order is computed by the warm-up
    step and kept (a long re-sort interval), so both paths sum in the same
    order and the captured step holds no sort.  With 1 MiB for the partner
    partials the far field runs in column batches inside the graph."""
    print("symmetric step in a graph, partial bytes %d" % partial_bytes, flush=True)
    h, prm = solid_case(20000, seed=77)
    ctx = g.Context(0)
    try:
        ctx.set_option("sweep_symmetric", 1)
        ctx.set_option("reorder_interval", 1000)
        ctx.set_option("sym_partial_bytes", partial_bytes)
        s = g.hip.Stream()
        de, dg = g.DeviceParticles.from_host(h), g.DeviceParticles.from_host(h)

        def step(d):
            return lambda st: (ctx.compute_derivatives(d, prm, st), ctx.integrate(d, prm, True, st))

        step(dg)(s)  # warm-up: order, scratch, second stream and events exist before the capture
        s.synchronize()
        graph = g.hip.Graph(s, step(dg))
        ctx2 = g.Context(0)  # the eager run, with the same options and its own order
        ctx2.set_option("sweep_symmetric", 1)
        ctx2.set_option("reorder_interval", 1000)
        ctx2.set_option("sym_partial_bytes", partial_bytes)
        ctx2.compute_derivatives(de, prm, s)  # the warm-up step
        ctx2.integrate(de, prm, True, s)
        for rep in range(3):
            graph.launch(s)
            ctx2.compute_derivatives(de, prm, s)
            ctx2.integrate(de, prm, True, s)
            s.synchronize()
            a, e = dg.download(), de.download()
            for f in ("posm", "pos64", "vel", "acc", "density", "dstress", "dstress_dt", "balsara", "maxvsig"):
                assert np.array_equal(a[f], e[f]), "replay %d field %s" % (rep, f)
        del graph
        de.free()
        dg.free()
        ctx2.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("which", ["spread", "last"])
@pytest.mark.parametrize("symmetric", [0, 1])
def test_sph_gravity_unequal_masses(symmetric, which):
    """With SPH on, the equal-mass check comes from the slot-order gather's
    per-block flags: masses spread by +-20%, or equal but for the last
    particle (in the last, partial block), must take the unequal-mass path
    of the one-sided and of the symmetric sweep (FP64 oracle)."""
    print("unequal masses with SPH symmetric=%d %s" % (symmetric, which), flush=True)
    h, prm = solid_case(n=5000 + 37, seed=31)
    rng = np.random.default_rng(5)
    if which == "spread":
        h["posm"][:, 3] *= rng.uniform(0.8, 1.2, size=h["posm"].shape[0]).astype(np.float32)
    else:
        h["posm"][-1, 3] *= np.float32(1.5)
    ctx = g.Context(0)
    try:
        ctx.set_option("sweep_symmetric", symmetric)
        out = gpu_run(ctx, h, prm)
    finally:
        ctx.close()
    ref = O.compute_derivatives(prm, h, "f64")
    check_vec(out["acc"], ref["acc"])
