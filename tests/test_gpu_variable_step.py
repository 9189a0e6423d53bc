"""The variable-step leapfrog (src/integration.h:127-255) through the C ABI
against the FP32 oracle.

dt is compared bit for bit: both evaluate the criteria
sqrt(accel_accuracy H / |a|) and velocity_accuracy H / maxvsig in FP32
rounded as written, on the same accelerations and signal speeds (the
oracle runs on the state downloaded from the device just before each
integrate call).  Positions and velocities are compared to FP32 rounding
(the device may fuse a multiply-add the oracle rounds twice).
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


@pytest.mark.parametrize("reset", [0, 1])
@pytest.mark.parametrize("fp64_drift", [False, True])
def test_variable_step_matches_oracle(reset, fp64_drift):
    h, prm = solid()
    p = variable(prm)
    if not fp64_drift:
        del h["pos64"]
    ctx = g.Context(0)  # fresh {current, next} dt state
    ctx.set_option("timestep_reset", reset)
    d = g.DeviceParticles.from_host(h)
    st = O.TimestepState("f32")
    dts = []
    for k in range(6):
        ctx.compute_derivatives(d, p)
        g.hip.synchronize()
        before = d.download()
        if reset and k > 0:
            st.ts[1] = np.float32(p.max_timestep)
        if fp64_drift:
            b64 = dict(before)
            b64["posm"] = before["posm"].astype(np.float64)
            b64["posm"][:, :3] = before["pos64"].T  # the FP64 oracle drifts the FP64 masters
            st64 = O.TimestepState("f64")
            st64.ts[:] = st.ts
            ref64 = O.variable_leapfrog(p, b64, st64, not_first_step=k > 0)
        ref = O.variable_leapfrog(p, before, st, not_first_step=k > 0)
        ctx.integrate(d, p, k > 0)
        dt = ctx.current_timestep(p)
        after = d.download()
        assert dt == st.current, (k, dt, st.current)
        assert np.allclose(after["vel"], ref["vel"], rtol=1e-6, atol=1e-7), k
        assert np.allclose(after["dstress"], ref["dstress"], rtol=1e-6, atol=1e-7), k
        if fp64_drift:
            # x64 += v dt in FP64 with the FP32 velocity: agrees with the FP64 oracle's drift
            # to the velocity's FP32 rounding times dt
            assert np.allclose(after["pos64"].T, ref64["posm"][:, :3], rtol=0, atol=1e-8), k
        else:
            assert np.allclose(after["posm"], ref["posm"], rtol=1e-6, atol=1e-7), k
        dts.append(dt)
    d.free()
    ctx.close()
    assert dts[0] == pytest.approx(0.05, rel=1e-7)  # initial_timestep
    if not reset:
        assert all(a >= b for a, b in zip(dts, dts[1:]))  # the running minimum never grows
    assert len(set(dts)) > 1


def test_first_call_with_not_first_step_takes_min_timestep():
    """main.cu:257 calls integrate(pb, true) from the first step on: the next
    dt is still zero then, so dt = min_timestep (integration.h:170)."""
    h, prm = solid(500, seed=6)
    p = variable(prm)
    ctx = g.Context(0)
    assert ctx.current_timestep(p) == 0.0  # currentTS before any step
    d = g.DeviceParticles.from_host(h)
    ctx.compute_derivatives(d, p)
    ctx.integrate(d, p, True)
    assert ctx.current_timestep(p) == pytest.approx(1e-4, rel=1e-7)
    ctx.compute_derivatives(d, p)
    ctx.integrate(d, p, True)
    assert ctx.current_timestep(p) == pytest.approx(1e-4, rel=1e-7)
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
