"""The C-ABI library loads and exports every symbol include/*.h declares
(CPU only: no device work)."""
import ctypes
import math
import os
import re
import subprocess

import pytest

import grasph2_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(grasph_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_symbols()
    assert "grasph_compute_derivatives" in names and len(names) >= 15
    lib = g.lib()
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", g._capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (grasph_\w+)", out))
    assert set(names) <= exported


def test_python_binding_covers_the_header():
    assert set(declared_symbols()) <= set(g._capi._SIGNATURES)


def test_version_and_struct_layout():
    assert g.lib().grasph_version() == 10100
    # grasph_params: 2 ints + 15 floats + 3*3 floats + 2 floats + 5 variable-step floats
    assert ctypes.sizeof(g.Params) == 4 * (2 + 15 + 9 + 2 + 5)
    assert ctypes.sizeof(g.Particles) == 8 * 12


def test_defaults_follow_settings_h():
    p = g.default_params()
    pr = 0.1 * 16384 ** (-1.0 / 3.0)  # settings.h:97-98
    assert p.H == pytest.approx(2.5 * pr, rel=1e-6)
    rho0 = 1.0 / 16384 / (4.0 / 3.0 * pr ** 3 * math.pi)
    assert p.rho0 == pytest.approx(rho0, rel=1e-6)
    assert p.SOUNDSPEED == pytest.approx(math.sqrt(1024 / rho0), rel=1e-6)
    assert p.BULK == 1024 and p.shear == 128 and p.tanfr == pytest.approx(1.0)
    assert p.flags == (g.SELF_GRAVITY | g.SPH | g.INTEGRATE_DENSITY | g.SOLIDS | g.PLASTICITY_MC |
                       g.ARTIFICIAL_VISCOSITY | g.BALSARA_SWITCH | g.VARIABLE_TIMESTEP |
                       g.ACCELERATION_CRITERION | g.VELOCITY_CRITERION)
    assert p.fixed_timestep == pytest.approx(0.002)
    # settings.h:56-62
    assert (p.initial_timestep, p.min_timestep, p.max_timestep) == pytest.approx((0.002, 0.00002, 0.02))
    assert (p.accel_accuracy, p.velocity_accuracy) == pytest.approx((0.01, 0.09))
    assert list(p.constant_acceleration) == [0.0, -1.0, 0.0]


def test_null_context_is_rejected():
    lib = g.lib()
    assert lib.grasph_compute_derivatives(None, None, None, None) == g.E_INVALID
    assert lib.grasph_destroy(None) == g.E_INVALID
    assert lib.grasph_last_error(None)  # never NULL


def test_host_layout():
    h = g.host_particles(5)
    assert h["posm"].shape == (5, 4) and h["pos64"].shape == (3, 5) and h["dstress"].shape == (6, 5)
