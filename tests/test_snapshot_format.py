"""Snapshot text files of the C++ layer (include/grasph2/storage.hpp) against
the reference's format, written out independently here from its rules
(src/ResultStorageManager.cu:101-186): per particle one line, every
attribute value in attribute order printed with std::fixed and 7 decimals
(digits10 + 1 of float) and followed by a tab; vectors x y z, matrices
their 9 entries row-major; file name prefix + start time + "_" + time with
the same decimals + ".tsv".  Reading back through ps::TextFile
(src/initialConditions/particleSources/TextFile.h) reproduces the file.
Host only: no GPU involved.  The reference ships no snapshot files, so the
format is pinned by these rules alone."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "snapshot_roundtrip")

f32 = np.float32


def expected_rows(n=5):
    """The buffer tests/cpp/snapshot_roundtrip.cpp builds, in float32."""
    rows = []
    for i in range(n):
        f = f32(i)
        pos = [f32(0.1) * f - f32(0.25), f32(1.0) / f32(3.0) + f, f32(-2.5e-3) * f]
        mass = f32(1.0) / (f32(1.0) + f)
        vel = [f, -f * f32(0.5), f32(1e-7) * f]
        dens = f32(1000.0) + f
        s = np.zeros(9, dtype=np.float32)
        s[0], s[4], s[8] = 0.5, f32(-0.25) * f, 0.125
        s[1] = s[3] = f32(0.01) * f
        s[2] = s[6] = -0.02
        s[5] = s[7] = 3.0
        vsig = f32(2.0) * f
        rows.append(pos + [mass] + vel + [dens] + list(s) + [vsig])  # POS MASS VEL DENSITY DSTRESS MAXVSIG
    return rows


def render(rows):
    return "".join("".join("%.7f\t" % float(f32(v)) for v in r) + "\n" for r in rows)


@pytest.mark.skipif(not os.path.exists(EXE), reason="tests/cpp/snapshot_roundtrip not built (run __graft_entry__.build())")
def test_snapshot_text_format_and_round_trip(tmp_path):
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    want = render(expected_rows())
    direct = (tmp_path / "direct.tsv").read_text()
    assert direct == want
    assert (tmp_path / "reread.tsv").read_text() == want  # ps::TextFile read back, written again
    name = r.stdout.strip()
    assert re.fullmatch(re.escape(str(tmp_path) + "/snap_") + r"\d{4}-\d{2}-\d{2}_\d{2}-\d{2}_0\.2500000\.tsv", name)
    assert open(name).read() == want
