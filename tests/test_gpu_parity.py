"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerance: relative 1e-5 against the FP64 oracle for acceleration and
density (the accuracy bound), computed with tests/_util.py's metrics:
  - norm-wise relative error <= 1e-5;
  - every particle's error relative to max(|a|, 0.1 median |a|) <= 1e-5
    (vec_errors_floored: the relative bound wherever |a| is at least a
    tenth of the median, an absolute bound of 1e-6 median below, where the
    random stresses of these inputs make pair terms cancel);
  - per-particle relative error <= 1e-4 over |a| >= 1e-3 median.
Tensor rates are compared relative to their largest entry.  Exactly-defined
quantities (neighbour-order independence, overflow fallback) are compared
bit for bit.
"""
import numpy as np
import pytest

import grasph2_amd as g
from grasph2_amd import ics
from oracle import pyoracle as O
from tests._util import scalar_errors, tensor_errors, vec_errors, vec_errors_floored

pytestmark = pytest.mark.gpu

TOL = 1e-5


def gpu_run(ctx, h, prm, op="derivs", **kw):
    d = g.DeviceParticles.from_host(h)
    if op == "derivs":
        ctx.compute_derivatives(d, prm)
    elif op == "density":
        ctx.calc_density(d, prm)
    elif op == "balsara":
        ctx.calc_balsara_density_dt_dstress_dt(d, prm)
    elif op == "accel":
        ctx.calc_acceleration(d, prm)
    elif op == "integrate":
        ctx.integrate(d, prm, kw.get("not_first_step", True))
    g.hip.synchronize()
    out = d.download()
    d.free()
    return out


def randomise(h, prm, seed, vscale=0.05, sscale=0.02):
    rng = np.random.default_rng(seed)
    n = h["posm"].shape[0]
    if "vel" in h:
        h["vel"][:, :3] += rng.normal(scale=vscale, size=(n, 3))
    if "dstress" in h:
        h["dstress"][:] = rng.normal(scale=sscale, size=(6, n))
    if "density" in h:
        h["density"][:] = prm.rho0 * (1 + rng.uniform(-0.05, 0.05, size=n))
    if "balsara" in h:
        h["balsara"][:] = rng.uniform(0, 1, size=n)
    return h


def check_vec(got, ref, tol=TOL, what="acc"):
    """Norm-wise relative error and floored per-particle error <= tol;
    per-particle relative error over |a_ref| >= 1e-3 median <= 10 tol."""
    mx, nr = vec_errors(got, ref)
    fl = vec_errors_floored(got, ref)
    print("[parity] %s: per-particle rel %.3g (floored %.3g), norm-wise %.3g" % (what, mx, fl, nr), flush=True)
    assert nr <= tol and fl <= tol and mx <= 10 * tol, \
        "%s: per-particle rel %.3g, floored %.3g, norm-wise %.3g" % (what, mx, fl, nr)
    return mx, nr


# ------------------------------------------------------------------ gravity
@pytest.mark.parametrize("n", [1, 2, 255, 1000, 4096])
def test_gravity_matches_fp64(gpu_ctx, n):
    h, prm = ics.plummer(n, seed=n)
    out = gpu_run(gpu_ctx, h, prm)
    ref = O.calc_acceleration(prm, h, "f64")
    if n == 1:
        assert np.all(out["acc"][:, :3] == 0)
        return
    check_vec(out["acc"], ref["acc"])


@pytest.mark.parametrize("uniform", [True, False])
def test_gravity_equal_and_unequal_masses(gpu_ctx, uniform):
    """Equal masses take the sweep that factors the mass out of the sum;
    one particle heavier by one ulp takes the general one.  Both against
    the FP64 oracle, unsoftened and softened."""
    h, prm = ics.plummer(3000, seed=11)
    h["posm"][:, 3] = np.float32(1.0 / 3000)
    if not uniform:
        h["posm"][1234, 3] = np.nextafter(h["posm"][1234, 3], np.float32(1))
    for H in (0.0, 0.01):
        prm.H = H
        out = gpu_run(gpu_ctx, h, prm)
        ref = O.calc_acceleration(prm, h, "f64")
        check_vec(out["acc"], ref["acc"])


def test_gravity_unsoftened_and_coincident(gpu_ctx):
    h, prm = ics.plummer(600, seed=3)
    prm.H = 0.0
    h["posm"][5] = h["posm"][4]  # two coincident particles: r2 = 0 pairs are skipped
    h["pos64"][:, 5] = h["pos64"][:, 4]
    out = gpu_run(gpu_ctx, h, prm)
    assert np.all(np.isfinite(out["acc"]))
    ref = O.calc_acceleration(prm, h, "f64")
    check_vec(out["acc"], ref["acc"])


# ------------------------------------------------------------------ SPH
def liquid_case(n=1500, seed=7):
    h, prm = ics.liquid_blob(n, seed=seed, fields=("posm", "pos64", "vel", "acc", "density", "maxvsig", "xvel",
                                                   "balsara"))
    h = randomise(h, prm, seed, vscale=0.2)
    return h, prm


def solid_case(n=3000, seed=8, gravity=True):
    h, prm = ics.collision(n, seed=seed, gravity=gravity)
    h = randomise(h, prm, seed)
    return h, prm


def test_liquid_derivatives(gpu_ctx):
    h, prm = liquid_case()
    out = gpu_run(gpu_ctx, h, prm)
    ref = O.compute_derivatives(prm, h, "f64")
    mx, nr = scalar_errors(out["density"], ref["density"])
    assert nr <= TOL and mx <= TOL, (mx, nr)
    check_vec(out["acc"], ref["acc"])
    # AV without the Balsara switch does not track the signal velocity
    assert np.all(out["maxvsig"] == 0) and np.all(ref["maxvsig"] == 0)


@pytest.mark.parametrize("gravity", [True, False])
def test_solid_derivatives(gpu_ctx, gravity):
    h, prm = solid_case(gravity=gravity)
    out = gpu_run(gpu_ctx, h, prm)
    ref = O.compute_derivatives(prm, h, "f64")
    mx, nr = scalar_errors(out["density"], ref["density"])
    assert nr <= TOL and mx <= TOL, ("density", mx, nr)
    check_vec(out["acc"], ref["acc"])
    mx, nr = tensor_errors(out["dstress_dt"], ref["dstress_dt"])
    assert mx <= 10 * TOL and nr <= TOL, ("dstress_dt", mx, nr)
    mx, nr = scalar_errors(out["balsara"], ref["balsara"], floor=1e-2)
    assert nr <= 10 * TOL, ("balsara", mx, nr)
    mx, nr = scalar_errors(out["maxvsig"], ref["maxvsig"], floor=1e-2)
    assert nr <= TOL, ("maxvsig", mx, nr)
    assert np.all(out["density_dt"] == 0)  # not integrated: stays zero as in the reference


def test_integrate_density_and_mohr_coulomb(gpu_ctx):
    h, prm = solid_case(2000, seed=12)
    prm.flags = (prm.flags | g.INTEGRATE_DENSITY | g.PLASTICITY_MC) & ~g.PLASTICITY_MIESE
    prm.tanfr, prm.cohesion = 0.8, 0.01
    out = gpu_run(gpu_ctx, h, prm)
    ref = O.compute_derivatives(prm, h, "f64")
    assert np.array_equal(out["density"], h["density"])  # no summation pass with INTEGRATE_DENSITY
    mx, nr = scalar_errors(out["density_dt"], ref["density_dt"], floor=1e-2)
    assert nr <= 10 * TOL, ("density_dt", mx, nr)
    check_vec(out["acc"], ref["acc"])


def test_liquid_extras(gpu_ctx):
    """Balsara, XSPH, artificial pressure, constant acceleration, no-partner density."""
    h, prm = liquid_case(1200, seed=21)
    prm.flags |= g.BALSARA_SWITCH | g.XSPH | g.ARTIFICIAL_STRESS | g.CONSTANT_ACCELERATION | g.DEAL_WITH_NO_PARTNERS
    prm.normalsep = 0.3 * prm.H
    h["posm"][-1, :3] = [50.0, 0, 0]  # an isolated particle
    h["pos64"][:, -1] = [50.0, 0, 0]
    out = gpu_run(gpu_ctx, h, prm)
    ref = O.compute_derivatives(prm, h, "f64")
    assert out["density"][-1] == pytest.approx(0.01 * prm.rho0)
    check_vec(out["acc"], ref["acc"])
    check_vec(out["xvel"], ref["xvel"], what="xvel")


@pytest.mark.parametrize("case", ["solid", "gravity", "none"])
def test_environment_forces(gpu_ctx, case):
    """Constant acceleration on top of SPH + gravity (added by the
    acceleration pass), on gravity alone (added where the sweep stores its
    sums: j-split combine at this N) and with neither (the sweep only
    finishes), against the FP64 oracle."""
    if case == "solid":
        h, prm = solid_case(2500, seed=23)
    else:
        h, prm = ics.plummer(5000, seed=24)
        h["vel"][:, :3] = np.random.default_rng(24).normal(scale=0.3, size=(5000, 3))
        if case == "none":
            prm.flags &= ~g.SELF_GRAVITY
    prm.flags |= g.CONSTANT_ACCELERATION
    prm.constant_acceleration[:] = (0.1, -0.5, 0.25)
    out = gpu_run(gpu_ctx, h, prm)
    ref = O.compute_derivatives(prm, h, "f64")
    check_vec(out["acc"], ref["acc"])
    if "xvel" in out:
        assert np.all(out["xvel"][:, :3] == 0)


@pytest.mark.parametrize("variable", [False, True])
def test_leapfrog_box_boundary(gpu_ctx, variable):
    """SIMPLE_BOX_BOUNDARY (integration.h:33-73): particles drifting out of
    the box are put on the wall and their velocity reflected, for the fixed
    and the variable step, against the FP64 oracle."""
    h, prm = solid_case(2000, seed=53)
    prm.flags |= g.SIMPLE_BOX_BOUNDARY
    lo = np.quantile(h["posm"][:, :3], 0.1, axis=0)  # a box that cuts through the bodies
    hi = np.quantile(h["posm"][:, :3], 0.9, axis=0)
    prm.box_min[:] = lo
    prm.box_max[:] = hi
    prm.box_reflectiveness = 0.6
    rng = np.random.default_rng(5)
    n = h["posm"].shape[0]
    h["vel"][:, :3] = rng.normal(scale=40.0, size=(n, 3))  # many cross a wall in one step
    h["acc"][:, :3] = rng.normal(size=(n, 3))
    if variable:
        prm.flags |= g.VARIABLE_TIMESTEP
        prm.initial_timestep = 2e-3
        ctx = g.Context(0)
        d = g.DeviceParticles.from_host(h)
        ctx.integrate(d, prm, False)
        g.hip.synchronize()
        out = d.download()
        d.free()
        ctx.close()
        st = O.TimestepState("f64")
        ref = O.variable_leapfrog(prm, h, st, not_first_step=False)
    else:
        out = gpu_run(gpu_ctx, h, prm, "integrate", not_first_step=True)
        ref = O.leapfrog(prm, h, not_first_step=True, prec="f64")
    moved = np.any((ref["posm"][:, :3] <= lo) | (ref["posm"][:, :3] >= hi), axis=1)
    assert moved.sum() > 100
    # FP32 velocities (|v| ~ 40) times dt against the oracle's FP64 ones: ~1e-8
    assert np.allclose(out["pos64"].T, ref["posm"][:, :3], rtol=0, atol=2e-8)
    assert np.allclose(out["vel"][:, :3], ref["vel"][:, :3], rtol=1e-6, atol=1e-5)


def test_standalone_passes(gpu_ctx):
    h, prm = solid_case(2500, seed=31)
    out = gpu_run(gpu_ctx, h, prm, "density")
    ref = O.calc_density(prm, h, "f64")
    assert scalar_errors(out["density"], ref["density"])[1] <= TOL
    out = gpu_run(gpu_ctx, h, prm, "balsara")
    ref = O.calc_balsara(prm, h, "f64")
    assert tensor_errors(out["dstress_dt"], ref["dstress_dt"])[1] <= TOL
    out = gpu_run(gpu_ctx, h, prm, "accel")
    ref = O.calc_acceleration(prm, h, "f64")
    check_vec(out["acc"], ref["acc"])


def test_order_and_list_capacity_do_not_change_results(gpu_ctx):
    """The list-overflow fallback walks the same partners in the same slot
    order (bit for bit); the identity order sums in another order (equal to
    FP32 rounding of the compensated sums)."""
    h, prm = solid_case(3000, seed=41)
    base = gpu_run(gpu_ctx, h, prm)
    try:
        gpu_ctx.set_option("reorder_interval", 0)  # identity order
        ident = gpu_run(gpu_ctx, h, prm)
        gpu_ctx.set_option("reorder_interval", 1)
        gpu_ctx.set_option("max_neighbours_grow", 0)  # keep the small capacity
        gpu_ctx.set_option("max_neighbours", 8)  # every list overflows: all-j fallback
        small = gpu_run(gpu_ctx, h, prm)
    finally:
        gpu_ctx.set_option("reorder_interval", 8)  # the defaults
        gpu_ctx.set_option("max_neighbours", 160)
        gpu_ctx.set_option("max_neighbours_grow", 1)
    for f in ("acc", "density", "dstress_dt", "balsara", "maxvsig"):
        assert np.array_equal(base[f], small[f]), f
    mx, nr = vec_errors(ident["acc"], base["acc"])
    assert nr <= 1e-6 and mx <= 1e-5, (mx, nr)
    for f in ("density", "balsara", "maxvsig"):
        mx, nr = scalar_errors(ident[f], base[f])
        assert nr <= 1e-6, (f, mx, nr)
    mx, nr = tensor_errors(ident["dstress_dt"], base["dstress_dt"])
    assert nr <= 1e-5, (mx, nr)


# ------------------------------------------------------------------ integrator
def test_list_capacity_grows_after_overflow(gpu_ctx):
    """An evaluation whose lists overflow (max_neighbours 16 against ~65
    neighbours) takes the exact scan; its largest count reaches the host
    behind an event, and the next evaluation raises max_neighbours to 1.25
    times it (steps of 32).  Both evaluations equal the one with ample
    capacity bit for bit."""
    h, prm = solid_case(3000, seed=43)
    base = gpu_run(gpu_ctx, h, prm)
    ctx = g.Context(0)
    try:
        ctx.set_option("max_neighbours", 16)
        first = gpu_run(ctx, h, prm)
        pairs, mxn, over = ctx.neighbour_stats()
        assert over > 0 and mxn > 16, (pairs, mxn, over)
        assert ctx.get_option("max_neighbours") == 16
        second = gpu_run(ctx, h, prm)  # gpu_run synchronised: the count has landed
        k = ctx.get_option("max_neighbours")
        pairs2, mxn2, over2 = ctx.neighbour_stats()
        print("capacity 16 -> %d (largest count %d)" % (k, mxn), flush=True)
        assert k % 32 == 0 and k >= mxn + mxn // 4 and over2 == 0 and mxn2 == mxn and pairs2 == pairs
        for f in ("acc", "density", "dstress_dt", "balsara", "maxvsig"):
            assert np.array_equal(first[f], base[f]) and np.array_equal(second[f], base[f]), f
    finally:
        ctx.close()


def test_long_lists_four_lanes_per_particle():
    """A list capacity above 512 (grown after a collision's impact) lets four
    lanes share each particle's list.  Results: within FP32 rounding of the
    one-lane walk and of the FP64 oracle, and the same bits whether a
    particle's list overflowed into the exact scan or not (the scan splits
    its partners over the lanes as the list does)."""
    h, prm = solid_case(2500, seed=61)
    prm = prm.copy()
    prm.H *= 2.4  # several hundred to ~1000 neighbours per particle
    runs = {}
    for k in (4096, 640, 512):
        ctx = g.Context(0)
        try:
            ctx.set_option("max_neighbours_grow", 0)
            ctx.set_option("max_neighbours", k)
            runs[k] = gpu_run(ctx, h, prm)
            runs[k]["stats"] = ctx.neighbour_stats()
        finally:
            ctx.close()
    pairs, mxn, over = runs[640]["stats"]
    print("[parity] long lists: %d pairs, largest %d, %d overflowed at 640" % (pairs, mxn, over), flush=True)
    assert mxn > 640 and over > 0 and runs[4096]["stats"][2] == 0
    for f in ("acc", "density", "dstress_dt", "balsara", "maxvsig"):
        assert np.array_equal(runs[640][f], runs[4096][f]), f  # list walk and scan agree bit for bit
    mx, nr = vec_errors(runs[4096]["acc"], runs[512]["acc"])
    assert nr <= 1e-6 and mx <= 1e-5, (mx, nr)  # four lanes against one: FP32 rounding
    ref = O.compute_derivatives(prm, h, "f64")
    check_vec(runs[4096]["acc"], ref["acc"])
    dmx, dnr = scalar_errors(runs[4096]["density"], ref["density"])
    assert dnr <= TOL and dmx <= TOL, (dmx, dnr)


def test_leapfrog_fp64_drift(gpu_ctx):
    h, prm = solid_case(2000, seed=51)
    prm.flags |= g.INTEGRATE_DENSITY
    rng = np.random.default_rng(1)
    n = h["posm"].shape[0]
    h["acc"][:, :3] = rng.normal(size=(n, 3))
    h["dstress_dt"][:] = rng.normal(scale=30.0, size=(6, n))  # drives some particles past yield
    h["density_dt"][:] = rng.normal(scale=5.0, size=n)
    for nfs in (False, True):
        out = gpu_run(gpu_ctx, h, prm, "integrate", not_first_step=nfs)
        ref = O.leapfrog(prm, h, not_first_step=nfs, prec="f64")
        assert np.allclose(out["pos64"].T, ref["posm"][:, :3], rtol=0, atol=1e-9)
        assert np.allclose(out["posm"][:, :3], ref["posm"][:, :3], rtol=0, atol=3e-7)
        assert np.allclose(out["vel"][:, :3], ref["vel"][:, :3], rtol=1e-6, atol=1e-7)
        assert np.allclose(out["dstress"], ref["dstress"], rtol=1e-5, atol=1e-7)
        assert np.allclose(out["density"], ref["density"], rtol=1e-6, atol=1e-7)


def test_leapfrog_fp32_drift_matches_fp32_oracle(gpu_ctx):
    h, prm = solid_case(2000, seed=52)
    rng = np.random.default_rng(2)
    n = h["posm"].shape[0]
    h["acc"][:, :3] = rng.normal(size=(n, 3))
    del h["pos64"]
    out = gpu_run(gpu_ctx, h, prm, "integrate")
    ref = O.leapfrog(prm, h, prec="f32")
    assert np.allclose(out["posm"], ref["posm"], rtol=1e-6, atol=1e-7)
    assert np.allclose(out["vel"], ref["vel"], rtol=1e-6, atol=1e-7)


# ------------------------------------------------------------------ invariants
def test_invariants(gpu_ctx):
    h, prm = solid_case(2000, seed=61)
    d = g.DeviceParticles.from_host(h)
    gpu_ctx.compute_derivatives(d, prm)
    inv = gpu_ctx.invariants(d, prm, half_kick_dt=0.5 * prm.fixed_timestep, with_potential=True)
    hh = d.download()
    d.free()
    ref = O.invariants(prm, hh, half_kick=0.5 * prm.fixed_timestep, phi=O.potential(prm, hh))
    for k in (g.INV_EKIN, g.INV_EINT, g.INV_MASS):
        assert inv[k] == pytest.approx(ref[k], rel=1e-9)
    assert inv[g.INV_EPOT] == pytest.approx(ref[g.INV_EPOT], rel=1e-6)
    scale = np.abs(ref[g.INV_PX:g.INV_LZ + 1]).max()
    assert np.allclose(inv[g.INV_PX:g.INV_LZ + 1], ref[g.INV_PX:g.INV_LZ + 1], atol=1e-9 * scale)


# ------------------------------------------------------------------ errors
def test_errors_are_reported(gpu_ctx):
    h, prm = solid_case(300, seed=71)
    d = g.DeviceParticles.from_host(h)
    bad = prm.copy()
    bad.flags |= g.ARTIFICIAL_STRESS
    with pytest.raises(g.GrasphError) as e:
        gpu_ctx.compute_derivatives(d, bad)
    assert e.value.code == g.E_UNSUPPORTED
    bad = prm.copy()
    bad.flags |= g.CLOHESSY_WILTSHIRE  # meaning left open by an mpUtils operator: rejected
    for call in (gpu_ctx.compute_derivatives, gpu_ctx.calc_acceleration, gpu_ctx.integrate):
        with pytest.raises(g.GrasphError) as e:
            call(d, bad)
        assert e.value.code == g.E_UNSUPPORTED and "CLOHESSY" in str(e.value)
    v = d.view()
    v.dstress = None
    with pytest.raises(g.GrasphError) as e:
        gpu_ctx.compute_derivatives(v, prm)
    assert e.value.code == g.E_INVALID and "dstress" in str(e.value)
    empty = g.Particles()
    empty.n = 0
    gpu_ctx.compute_derivatives(empty, prm)  # zero particles: nothing to do
    d.free()
