"""The allocation-free derivative entry point, grasph_compute_derivatives_noalloc.

Checked through the C ABI on two calls in a row, the second with fewer
particles, against the allocating grasph_compute_derivatives:
  - bit for bit, every output array, when the allocating call runs the same
    sweep (options sweep_symmetric = 0, reorder_interval = 1 and the j-split
    the allocation-free call took, read back as option "noalloc_slices");
  - within the 1e-5 relative bound of the allocating call with its default
    options, and of the FP64 oracle on sampled particles.
The device's free memory is the same before and after the calls that follow
the first (which loads nothing new either: the context's creation resolved
the scratch array), and a call that would need more than the scratch array
holds is refused before anything is enqueued.
"""
import numpy as np
import pytest

import grasph2_amd as g
from grasph2_amd import ics
from tests._util import sampled_oracle, scalar_errors, vec_errors

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _inputs(kind, n, seed):
    if kind == "solid":
        h, prm = ics.collision(n, seed=seed)
        rng = np.random.default_rng(seed)
        h["vel"][:, :3] += rng.normal(scale=0.05, size=(n, 3))
        h["dstress"][:] = rng.normal(scale=0.02, size=(6, n))
    else:
        h, prm = ics.plummer(n, seed=seed)
    return h, prm


def _run(ctx, h, prm, noalloc):
    """Outputs of one call, and the device's free memory just before and
    after it (the particle arrays are already allocated then)."""
    d = g.DeviceParticles.from_host(h)
    g.hip.synchronize()
    before = g.hip.mem_info()[0]
    if noalloc:
        ctx.compute_derivatives_noalloc(d, prm)
    else:
        ctx.compute_derivatives(d, prm)
    g.hip.synchronize()
    after = g.hip.mem_info()[0]
    out = d.download()
    d.free()
    return out, before, after


@pytest.mark.parametrize("kind,sizes", [("solid", (12000, 5000)), ("gravity", (20000, 3000))])
def test_noalloc_matches_allocating_entry_point(kind, sizes):
    free_ctx = g.Context(0)   # allocation-free calls only
    same_ctx = g.Context(0)   # allocating calls, same sweep configuration
    dflt_ctx = g.Context(0)   # allocating calls, default options
    same_ctx.set_option("sweep_symmetric", 0)
    same_ctx.set_option("reorder_interval", 1)
    try:
        for call, n in enumerate(sizes):
            h, prm = _inputs(kind, n, seed=40 + call)
            print("[noalloc] %s call %d: n = %d" % (kind, call, n), flush=True)
            got, free_before, free_after = _run(free_ctx, h, prm, noalloc=True)
            slices = free_ctx.get_option("noalloc_slices")
            print("[noalloc]   j-slices %d, free memory %d -> %d" % (slices, free_before, free_after), flush=True)
            same_ctx.set_option("sweep_slices", slices)
            ref = _run(same_ctx, h, prm, noalloc=False)[0]
            for f in got:
                assert np.array_equal(got[f].view(np.uint8), ref[f].view(np.uint8)), \
                    "%s call %d: %s differs from the allocating entry point" % (kind, call, f)
            dflt = _run(dflt_ctx, h, prm, noalloc=False)[0]
            mx, nr = vec_errors(got["acc"], dflt["acc"])
            print("[noalloc]   against the default path: acc max %.3g norm %.3g" % (mx, nr), flush=True)
            assert mx <= TOL and nr <= TOL
            sample = np.random.default_rng(call).choice(n, size=min(n, 256), replace=False)
            ora = sampled_oracle(prm, h, sample, "f64")
            mx, nr = vec_errors(got["acc"][sample], ora["acc"][sample])
            print("[noalloc]   against the FP64 oracle: acc max %.3g norm %.3g" % (mx, nr), flush=True)
            assert mx <= TOL and nr <= TOL
            if "density" in ora:
                mx, nr = scalar_errors(got["density"][sample], ora["density"][sample])
                assert mx <= TOL and nr <= TOL
            if call > 0:
                assert free_after == free_before, "the allocation-free call changed the free device memory"
    finally:
        for c in (free_ctx, same_ctx, dflt_ctx):
            c.close()


def test_noalloc_rejects_what_does_not_fit():
    ctx = g.Context(0)
    try:
        h, prm = _inputs("solid", 60000, seed=3)
        d = g.DeviceParticles.from_host(h)
        with pytest.raises(g.GrasphError) as e:
            ctx.compute_derivatives_noalloc(d, prm)
        assert "bytes of scratch" in str(e.value)
        ctx.set_option("max_neighbours", 4096)
        h2, prm2 = _inputs("solid", 4000, seed=4)
        d2 = g.DeviceParticles.from_host(h2)
        with pytest.raises(g.GrasphError):
            ctx.compute_derivatives_noalloc(d2, prm2)
        g.hip.synchronize()
        d.free()
        d2.free()
    finally:
        ctx.close()
