"""Initial conditions and snapshots on the device path, driven as
src/main.cu:191-266 drives the reference (tests/cpp/ic_device_roundtrip.cpp):
InitGenerator + particle sources -> DeviceParticleBuffer -> steps of
computeDerivatives(pb); integrate(pb, true) on the GPU ->
ResultStorageManager::printToFile(DeviceParticleBuffer) -> ps::TextFile.

The program checks the snapshot read back against the state downloaded when
printToFile was called (the simulation goes on meanwhile).  Here the same
initial conditions, read from what the program generated, are stepped by the
FP64 oracle, and the state after the steps is compared: the acceleration of
the last derivative evaluation at 1e-5 (the accuracy bound), positions and
velocities to the FP32 rounding the steps accumulate.
"""
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
EXE = os.path.join(ROOT, "tests", "cpp", "ic_device_roundtrip")


def params_for(kind, n):
    if kind == "plummer":
        return ics.plummer_params(H=0.01, dt=1e-3)
    d = (4.0 / 3.0 * np.pi / (n // 2)) ** (1.0 / 3.0)  # spacing of n/2 particles in a unit sphere
    return ics.solid_params(2.5 * d, dt=1e-3, gravity=True)


@pytest.mark.skipif(not os.path.exists(EXE), reason="tests/cpp/ic_device_roundtrip not built")
@pytest.mark.parametrize("kind,n", [("plummer", 4096), ("collision", 4096)])
def test_generated_ics_stepped_on_device_and_snapshotted(tmp_path, kind, n):
    print("ic device roundtrip %s n=%d" % (kind, n), flush=True)
    steps = 2
    prm = params_for(kind, n)
    pf = tmp_path / "params.bin"
    pf.write_bytes(bytes(prm))
    r = subprocess.run([EXE, str(tmp_path), str(pf), kind, str(n), str(steps)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "particles match" in r.stdout
    snaps = [f for f in os.listdir(tmp_path) if f.startswith("snap_") and f.endswith(".tsv")]
    assert len(snaps) == 1, snaps
    rows = np.fromfile(tmp_path / "ic.bin", dtype=np.float32).reshape(n, 8)
    fields = ics.SOLID_FIELDS if kind == "collision" else ("posm", "pos64", "vel", "acc")
    h = g.host_particles(n, fields)
    h["posm"][:] = rows[:, :4]
    h["pos64"][:] = rows[:, :3].astype(np.float64).T
    h["vel"][:, :3] = rows[:, 4:7]
    if "density" in h:
        h["density"][:] = rows[:, 7]
    raw = open(tmp_path / "after.bin", "rb").read()
    o = 0
    after = {}
    for k, nb, dt, shape in (("posm", 16, np.float32, (n, 4)), ("vel", 16, np.float32, (n, 4)),
                             ("acc", 16, np.float32, (n, 4)), ("density", 4, np.float32, (n,)),
                             ("pos64", 24, np.float64, (3, n))):
        after[k] = np.frombuffer(raw[o:o + nb * n], dtype=dt).reshape(shape)
        o += nb * n
    # the FP64 oracle over the same steps: it drifts FP64 masters and, as the
    # kernels do, evaluates the derivatives at their FP32 mirror (with FP64
    # arithmetic), so the last evaluation compares like with like
    cur = h
    for _ in range(steps):
        ev = {k: v for k, v in cur.items() if k != "pos64"}
        ev["posm"] = np.array(cur["posm"], dtype=np.float64)
        ev["posm"][:, :3] = np.asarray(cur["pos64"], dtype=np.float64).T.astype(np.float32)
        dv = O.compute_derivatives(prm, ev, "f64")
        for k, v in dv.items():
            if k != "posm":
                cur[k] = v
        lf = O.leapfrog(prm, cur, True, "f64")
        cur["posm"], cur["vel"] = lf["posm"], lf["vel"]
        cur["pos64"] = np.asarray(cur["posm"][:, :3], dtype=np.float64).T.copy()
        if "dstress" in lf:
            cur["dstress"] = lf["dstress"]
    mx, nr = vec_errors(after["acc"], dv["acc"])
    print("  acc of the last evaluation: max %.3g norm %.3g" % (mx, nr), flush=True)
    assert nr <= 1e-5 and mx <= 1e-5, (mx, nr)
    scale = np.abs(cur["posm"][:, :3]).max()
    assert np.abs(after["pos64"].T - cur["posm"][:, :3]).max() <= 1e-6 * scale
    vscale = np.abs(cur["vel"][:, :3]).max()
    assert np.abs(after["vel"][:, :3] - cur["vel"][:, :3]).max() <= 1e-5 * vscale
