"""The variable-step leapfrog (src/integration.h:127-255) through the C ABI
against the FP32 oracle, driven as the reference's main loop drives it
(src/main.cu:256-257): computeDerivatives(pb), then integrate(pb, true),
from the very first step on.

Every acceleration pass sets the next dt to max_timestep
(src/computeDerivatives.h:305-308), so each step's dt is
min(max_timestep, that step's criteria) and grows again when the criteria
loosen.  dt is compared bit for bit: device and oracle evaluate the criteria
sqrt(accel_accuracy H / |a|) and velocity_accuracy H / maxvsig in FP32
rounded as written, on the same accelerations and signal speeds (the oracle
runs its leapfrog on the state downloaded from the device just before each
integrate call, and its own derivative pass on the state before, which
performs the reset).  Positions and velocities are compared to FP32
rounding (the device may fuse a multiply-add the oracle rounds twice).
"""
import numpy as np
import pytest

import grasph2_amd as g
from grasph2_amd import ics
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def variable(prm, initial=0.05, mn=1e-4, mx=0.1, acc=0.01, vel=0.09):
    p = prm.copy()
    p.flags |= g.VARIABLE_TIMESTEP | g.ACCELERATION_CRITERION | g.VELOCITY_CRITERION
    p.initial_timestep, p.min_timestep, p.max_timestep = initial, mn, mx
    p.accel_accuracy, p.velocity_accuracy = acc, vel
    return p


def solid(n=2000, seed=5):
    h, prm = ics.collision(n, seed=seed)
    rng = np.random.default_rng(seed)
    h["vel"][:, :3] += rng.normal(scale=0.1, size=(n, 3))
    h["dstress"][:] = rng.normal(scale=0.02, size=(6, n))
    return h, prm


def criteria_dt(p, state):
    """min(max_timestep, criteria) of a state, in FP64 (integration.h:134,149-154)."""
    a = np.linalg.norm(state["acc"][:, :3].astype(np.float64), axis=1)
    dt = float(p.max_timestep)
    with np.errstate(divide="ignore"):
        dt = min(dt, float(np.min(np.sqrt(float(p.accel_accuracy) * float(p.H) / a))))
        dt = min(dt, float(np.min(float(p.velocity_accuracy) * float(p.H) / state["maxvsig"].astype(np.float64))))
    return max(dt, float(p.min_timestep))


# (accel_accuracy, velocity_accuracy) per step: the criteria tighten twice,
# loosen back, then let max_timestep bind
ACCURACIES = [(0.01, 0.09), (0.001, 0.03), (0.0001, 0.01), (0.01, 0.09), (10.0, 10.0)]


@pytest.mark.parametrize("fp64_drift", [False, True])
def test_main_loop_dt_tightens_and_grows_back(fp64_drift):
    print("variable step main loop, fp64_drift=%s" % fp64_drift, flush=True)
    h, prm = solid()
    if not fp64_drift:
        del h["pos64"]
    ctx = g.Context(0)  # fresh {current, next} dt state, zero as the reference's
    d = g.DeviceParticles.from_host(h)
    st = O.TimestepState("f32")
    dts = []
    for k, (aa, va) in enumerate(ACCURACIES):
        p = variable(prm, acc=aa, vel=va)
        pre = d.download()
        O.compute_derivatives(p, pre, "f32", state=st)  # the oracle's acceleration pass resets next dt
        assert st.next == np.float32(p.max_timestep)
        ctx.compute_derivatives(d, p)
        g.hip.synchronize()
        before = d.download()
        if fp64_drift:
            b64 = dict(before)
            b64["posm"] = before["posm"].astype(np.float64)
            b64["posm"][:, :3] = before["pos64"].T  # the FP64 oracle drifts the FP64 masters
            st64 = O.TimestepState("f64")
            st64.ts[:] = st.ts
            ref64 = O.variable_leapfrog(p, b64, st64, not_first_step=True)
        ref = O.variable_leapfrog(p, before, st, not_first_step=True)
        ctx.integrate(d, p, True)
        dt = ctx.current_timestep(p)
        after = d.download()
        assert dt == st.current, (k, dt, st.current)
        assert dt == pytest.approx(criteria_dt(p, before), rel=1e-6), k
        assert np.allclose(after["vel"], ref["vel"], rtol=1e-6, atol=1e-7), k
        assert np.allclose(after["dstress"], ref["dstress"], rtol=1e-6, atol=1e-7), k
        if fp64_drift:
            # x64 += v dt in FP64 with the FP32 velocity: agrees with the FP64 oracle's drift
            # to the velocity's FP32 rounding times dt: the two half kicks round v + a dt / 2 in FP32,
            # a few ulps of |v| + |a| dt per component
            kick = np.abs(before["acc"][:, :3].astype(np.float64)) * float(dt)
            tol = 2.0 ** -21 * (np.abs(ref64["vel"][:, :3]) + kick) * float(dt) + 1e-15
            assert np.all(np.abs(after["pos64"].T - ref64["posm"][:, :3]) <= tol), k
        else:
            assert np.allclose(after["posm"], ref["posm"], rtol=1e-6, atol=1e-7), k
        dts.append(dt)
    d.free()
    ctx.close()
    print("dt sequence", dts, flush=True)
    assert dts[1] < dts[0] and dts[2] < dts[1]  # tightened
    assert dts[3] > dts[2]  # grew back
    assert dts[4] == pytest.approx(0.1, rel=1e-7)  # max_timestep binds


def test_first_main_loop_step_takes_criteria_not_min_timestep():
    """main.cu:256-257 calls computeDerivatives then integrate(pb, true) from
    the first step on: the acceleration pass has set the next dt to
    max_timestep, so the first dt is min(max_timestep, criteria)."""
    h, prm = solid(500, seed=6)
    p = variable(prm)
    ctx = g.Context(0)
    assert ctx.current_timestep(p) == 0.0  # currentTS before any step
    d = g.DeviceParticles.from_host(h)
    ctx.compute_derivatives(d, p)
    g.hip.synchronize()
    before = d.download()
    ctx.integrate(d, p, True)
    dt = ctx.current_timestep(p)
    assert dt == pytest.approx(criteria_dt(p, before), rel=1e-6)
    assert dt > 10 * p.min_timestep
    d.free()
    ctx.close()


def test_standalone_acceleration_pass_resets_next_dt():
    """calcAcceleration alone (do_for_each_pair_fast<calcAcceleration>) also
    resets the next dt (computeDerivatives.h:305-308): with max_timestep
    binding, two integrate calls with an acceleration pass between them both
    run at max_timestep; without it the second keeps the running minimum."""
    h, prm = solid(400, seed=8)
    p = variable(prm, mx=0.004)  # below every criterion of this state
    loose = variable(prm, mx=0.004, acc=0.01, vel=0.09)
    tight = variable(prm, mx=0.004, acc=1e-7, vel=1e-5)  # criteria far below max_timestep
    ctx = g.Context(0)
    d = g.DeviceParticles.from_host(h)
    ctx.compute_derivatives(d, p)
    ctx.integrate(d, tight, True)  # next := min(0.004, tight criteria)
    t1 = ctx.current_timestep(p)
    assert t1 < 0.004
    ctx.integrate(d, loose, True)  # no acceleration pass: the running minimum stays
    assert ctx.current_timestep(p) == t1
    ctx.calc_acceleration(d, loose)
    ctx.integrate(d, loose, True)
    assert ctx.current_timestep(p) == pytest.approx(0.004, rel=1e-7)
    d.free()
    ctx.close()


def test_velocity_criterion_needs_maxvsig():
    h, prm = solid(300, seed=7)
    del h["maxvsig"]
    p = variable(prm)
    p.flags &= ~g.SPH  # the derivative passes would need maxvsig too
    ctx = g.Context(0)
    d = g.DeviceParticles.from_host(h)
    with pytest.raises(g.GrasphError) as e:
        ctx.integrate(d, p, True)
    assert e.value.code == g.E_INVALID
    d.free()
    ctx.close()
