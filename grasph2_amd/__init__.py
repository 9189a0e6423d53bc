"""grasph2_amd -- MI355X-native GraSPH2 derivative + leapfrog path.

The product is the HIP library libgrasph2_amd.so behind the C ABI of
include/grasph2.h (C++ callers use include/grasph2/*.hpp).  This package is
the Python plumbing around it: a ctypes binding, device particle arrays and
a context object, used by the tests and the benchmark.
"""
import ctypes

import numpy as np

from . import hip
from ._capi import (  # noqa: F401
    ARTIFICIAL_STRESS, ARTIFICIAL_VISCOSITY, BALSARA_SWITCH, CLOHESSY_WILTSHIRE, CONSTANT_ACCELERATION,
    DEAL_WITH_NO_PARTNERS, INTEGRATE_DENSITY, PLASTICITY_MC, PLASTICITY_MIESE, SELF_GRAVITY, SIMPLE_BOX_BOUNDARY,
    SOLIDS, SPH, XSPH, E_HIP, E_INVALID, E_NOMEM, E_UNSUPPORTED, NINV, INV_EKIN, INV_EPOT, INV_EINT, INV_PX, INV_PY,
    INV_PZ, INV_LX, INV_LY, INV_LZ, INV_MASS, KERNEL_CLASSES, NKERNEL_CLASSES, GrasphError, Params, Particles,
    default_params, lib,
)

# (shape as a function of n, dtype) of every particle array of the C ABI
FIELDS = {
    "posm": (lambda n: (n, 4), np.float32),
    "pos64": (lambda n: (3, n), np.float64),
    "vel": (lambda n: (n, 4), np.float32),
    "acc": (lambda n: (n, 4), np.float32),
    "xvel": (lambda n: (n, 4), np.float32),
    "density": (lambda n: (n,), np.float32),
    "density_dt": (lambda n: (n,), np.float32),
    "balsara": (lambda n: (n,), np.float32),
    "maxvsig": (lambda n: (n,), np.float32),
    "dstress": (lambda n: (6, n), np.float32),
    "dstress_dt": (lambda n: (6, n), np.float32),
}


def host_particles(n, fields=tuple(FIELDS)):
    """Zero-filled host arrays (numpy) in the C-ABI layout."""
    return {f: np.zeros(FIELDS[f][0](n), dtype=FIELDS[f][1]) for f in fields}


class DeviceParticles:
    """Device copies of the particle arrays; owns the memory."""

    def __init__(self, n, fields=tuple(FIELDS)):
        self.n = int(n)
        self.arrays = {}
        for f in fields:
            shape, dt = FIELDS[f][0](self.n), FIELDS[f][1]
            self.arrays[f] = hip.DeviceArray(int(np.prod(shape)) * np.dtype(dt).itemsize)
        self._view = None

    @classmethod
    def from_host(cls, h, fields=None):
        n = h["posm"].shape[0]
        fields = tuple(fields or h.keys())
        d = cls(n, fields)
        d.upload(h)
        return d

    def upload(self, h):
        for f, a in h.items():
            if f in self.arrays:
                shape, dt = FIELDS[f][0](self.n), FIELDS[f][1]
                a = np.ascontiguousarray(a, dtype=dt).reshape(shape)
                self.arrays[f].upload(a)

    def download(self, fields=None):
        out = {}
        for f in fields or self.arrays:
            shape, dt = FIELDS[f][0](self.n), FIELDS[f][1]
            out[f] = self.arrays[f].download(dt, int(np.prod(shape))).reshape(shape)
        return out

    def view(self):
        v = Particles()
        v.n = self.n
        for f, _ in Particles._fields_[1:]:
            setattr(v, f, self.arrays[f].value if f in self.arrays else None)
        return v

    def free(self):
        for a in self.arrays.values():
            a.free()
        self.arrays = {}


class Context:
    """One grasph_ctx on a device; methods raise GrasphError on failure."""

    def __init__(self, device=0):
        self._lib = lib()
        h = ctypes.c_void_p()
        r = self._lib.grasph_create(device, ctypes.byref(h))
        if r != 0:
            raise GrasphError(r, self._lib.grasph_last_error(None).decode())
        self.handle = h

    def _check(self, r):
        if r != 0:
            raise GrasphError(r, self._lib.grasph_last_error(self.handle).decode())
        return r

    def set_option(self, key, value):
        return self._check(self._lib.grasph_set_option(self.handle, key.encode(), int(value)))

    @staticmethod
    def _s(stream):
        return stream.handle if stream is not None else None

    def compute_derivatives(self, parts, prm, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_compute_derivatives(self.handle, ctypes.byref(v), ctypes.byref(prm),
                                                                self._s(stream)))

    def calc_density(self, parts, prm, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_calc_density(self.handle, ctypes.byref(v), ctypes.byref(prm),
                                                         self._s(stream)))

    def calc_balsara_density_dt_dstress_dt(self, parts, prm, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_calc_balsara_density_dt_dstress_dt(
            self.handle, ctypes.byref(v), ctypes.byref(prm), self._s(stream)))

    def calc_acceleration(self, parts, prm, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_calc_acceleration(self.handle, ctypes.byref(v), ctypes.byref(prm),
                                                              self._s(stream)))

    def integrate(self, parts, prm, not_first_step=True, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_integrate(self.handle, ctypes.byref(v), ctypes.byref(prm),
                                                      1 if not_first_step else 0, self._s(stream)))

    def current_timestep(self, prm):
        return self._lib.grasph_get_current_timestep(self.handle, ctypes.byref(prm))

    def invariants(self, parts, prm, half_kick_dt=0.0, with_potential=True, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        out = (ctypes.c_double * NINV)()
        self._check(self._lib.grasph_reduce_invariants(self.handle, ctypes.byref(v), ctypes.byref(prm),
                                                       float(half_kick_dt), 1 if with_potential else 0, out,
                                                       self._s(stream)))
        return np.array(out[:], dtype=np.float64)

    def kernel_times(self, reset=True):
        ms = (ctypes.c_double * NKERNEL_CLASSES)()
        n = (ctypes.c_int64 * NKERNEL_CLASSES)()
        self._check(self._lib.grasph_get_kernel_times(self.handle, ms, n, 1 if reset else 0))
        return {k: (ms[i], n[i]) for i, k in enumerate(KERNEL_CLASSES)}

    def neighbour_stats(self):
        """(ordered neighbour pairs, largest count, particles over capacity) of the last list."""
        pairs, mx, over = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64()
        self._check(self._lib.grasph_get_neighbour_stats(self.handle, ctypes.byref(pairs), ctypes.byref(mx),
                                                         ctypes.byref(over)))
        return pairs.value, mx.value, over.value

    def close(self):
        if self.handle is not None and self.handle.value:
            self._lib.grasph_destroy(self.handle)
        self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
