"""The symmetric sweep's work geometry (grasph2_amd/csrc/sym_geometry.hpp),
checked on the host by tests/cpp/sym_geometry_check: for particle counts
from 1 to 150000 and far-kernel workgroup counts from 1 to 100000, every
pair of super-blocks in different groups is computed exactly once, every
partial-sum index is in bounds and written once, and the combine reads
exactly what was written.  No GPU involved."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "sym_geometry_check")


@pytest.mark.skipif(not os.path.exists(EXE), reason="tests/cpp/sym_geometry_check not built (run __graft_entry__.build())")
def test_far_field_work_covers_every_pair_once():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
