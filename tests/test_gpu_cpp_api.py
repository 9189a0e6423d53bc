"""The C++ template API (include/grasph2/particles.hpp), driven the way the
reference's main.cu drives its own (tests/cpp/api_roundtrip.cpp), gives the
same particles as the C ABI called from Python, and matches the FP64 oracle."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import grasph2_amd as g
from grasph2_amd import ics
from oracle import pyoracle as O
from tests._util import vec_errors

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "api_roundtrip")


def run_cpp(tmp_path, h, prm, steps):
    n = h["posm"].shape[0]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        f.write(bytes(prm))
        f.write(np.int64(n).tobytes())
        f.write(np.int32(steps).tobytes())
        for k in ("posm", "vel", "density", "dstress"):
            f.write(np.ascontiguousarray(h[k], dtype=np.float32).tobytes())
    r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(fout, dtype=np.float32)
    out, o = {}, 0
    for k, size, shape in (("posm", 4 * n, (n, 4)), ("vel", 4 * n, (n, 4)), ("acc", 4 * n, (n, 4)),
                           ("density", n, (n,)), ("balsara", n, (n,)), ("maxvsig", n, (n,)),
                           ("dstress", 6 * n, (6, n)), ("dstress_dt", 6 * n, (6, n))):
        out[k] = raw[o:o + size].reshape(shape)
        o += size
    return out


def case(n=2000, seed=3):
    h, prm = ics.collision(n, seed=seed)
    rng = np.random.default_rng(seed)
    h["vel"][:, :3] += rng.normal(scale=0.05, size=(n, 3))
    h["dstress"][:] = rng.normal(scale=0.02, size=(6, n))
    return h, prm


@pytest.mark.skipif(not os.path.exists(EXE), reason="tests/cpp/api_roundtrip not built (run __graft_entry__.build())")
def test_cpp_api_matches_c_abi_and_oracle(tmp_path, gpu_ctx):
    h, prm = case()
    got = run_cpp(tmp_path, h, prm, 0)
    ref = O.compute_derivatives(prm, h, "f64")
    mx, nr = vec_errors(got["acc"], ref["acc"])
    assert nr <= 1e-5, (mx, nr)
    got = run_cpp(tmp_path, h, prm, 2)
    d = g.DeviceParticles.from_host(h)
    for _ in range(2):
        gpu_ctx.compute_derivatives(d, prm)
        gpu_ctx.integrate(d, prm, True)
    g.hip.synchronize()
    py = d.download()
    d.free()
    for k in ("posm", "vel", "acc", "density", "balsara", "maxvsig", "dstress", "dstress_dt"):
        assert np.array_equal(got[k], py[k]), k
