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
    SOLIDS, SPH, XSPH, VARIABLE_TIMESTEP, ACCELERATION_CRITERION, VELOCITY_CRITERION, E_HIP, E_INVALID, E_NOMEM, E_UNSUPPORTED, NINV, INV_EKIN, INV_EPOT, INV_EINT, INV_PX, INV_PY,
    INV_PZ, INV_LX, INV_LY, INV_LZ, INV_MASS, KERNEL_CLASSES, NKERNEL_CLASSES, Batch, GrasphError, Params, Particles,
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


def pack_batch(instances, fields=None):
    """Concatenate per-instance host arrays into the batch layout of
    grasph_batch: every instance keeps its own single-instance layout, so
    planar arrays hold instance k's planes back to back at planes * offset_k.
    Returns (host dict shaped like one n_total instance, offsets)."""
    fields = fields or [f for f in FIELDS if f in instances[0]]
    sizes = [h["posm"].shape[0] for h in instances]
    offsets = np.zeros(len(sizes) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)
    n = int(offsets[-1])
    out = {}
    for f in fields:
        flat = np.concatenate([np.ascontiguousarray(h[f], dtype=FIELDS[f][1]).ravel() for h in instances])
        out[f] = flat.reshape(FIELDS[f][0](n))
    return out, offsets


def unpack_batch(h, offsets):
    """Inverse of pack_batch: a list of per-instance host dicts."""
    out = []
    for k in range(len(offsets) - 1):
        a, b = int(offsets[k]), int(offsets[k + 1])
        d = {}
        for f, arr in h.items():
            flat = np.ascontiguousarray(arr).ravel()
            per = int(np.prod(FIELDS[f][0](1)))  # values per particle (planes for planar arrays)
            d[f] = flat[per * a:per * b].reshape(FIELDS[f][0](b - a))
        out.append(d)
    return out


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

    def get_option(self, key):
        v = ctypes.c_int64()
        self._check(self._lib.grasph_get_option(self.handle, key.encode(), ctypes.byref(v)))
        return v.value

    @staticmethod
    def _s(stream):
        return stream.handle if stream is not None else None

    def compute_derivatives(self, parts, prm, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_compute_derivatives(self.handle, ctypes.byref(v), ctypes.byref(prm),
                                                                self._s(stream)))

    def compute_derivatives_noalloc(self, parts, prm, stream=None):
        """grasph_compute_derivatives_noalloc: the same passes with every work
        array cut from the library's static per-GPU scratch (no allocation)."""
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_compute_derivatives_noalloc(self.handle, ctypes.byref(v),
                                                                        ctypes.byref(prm), self._s(stream)))

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
        """getCurrentTimestep(): fixed_timestep, or the context's current variable dt."""
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

    # ---- batched ensembles
    def batch(self, offsets, params):
        """A grasph_batch; keeps its host arrays alive on the returned object."""
        offs = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        arr = (Params * len(params))(*params)
        b = Batch()
        b.n_instances = len(params)
        b.offsets = offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        b.params = arr
        b._keep = (offs, arr)
        return b

    def batch_compute_derivatives(self, parts, b, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_batch_compute_derivatives(self.handle, ctypes.byref(v), ctypes.byref(b),
                                                                      self._s(stream)))

    def batch_integrate(self, parts, b, not_first_step=True, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_batch_integrate(self.handle, ctypes.byref(v), ctypes.byref(b),
                                                            1 if not_first_step else 0, self._s(stream)))

    def batch_step(self, parts, b, n_steps=1, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_batch_step(self.handle, ctypes.byref(v), ctypes.byref(b), int(n_steps),
                                                       self._s(stream)))

    def batch_step_loop(self, parts, b, n_steps=1, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_batch_step_loop(self.handle, ctypes.byref(v), ctypes.byref(b),
                                                            int(n_steps), self._s(stream)))

    def autotune(self, parts, b, cache_path=None, stream=None):
        v = parts.view() if isinstance(parts, DeviceParticles) else parts
        return self._check(self._lib.grasph_autotune(self.handle, ctypes.byref(v), ctypes.byref(b),
                                                     cache_path.encode() if cache_path else None, self._s(stream)))

    def batch_current_timesteps(self, b):
        out = (ctypes.c_double * b.n_instances)()
        self._check(self._lib.grasph_batch_get_current_timesteps(self.handle, ctypes.byref(b), out))
        return np.array(out[:], dtype=np.float64)

    def batch_tuning(self):
        ni = (ctypes.c_int32 * 4)()
        w, bp = ctypes.c_int32(), ctypes.c_int32()
        self._check(self._lib.grasph_get_batch_tuning(self.handle, ni, ctypes.byref(w), ctypes.byref(bp)))
        sym = self.get_option("batch_symmetric")
        return {"ni": list(ni), "waves_per_block": w.value, "blocks_per_cu": bp.value,
                "symmetric_min": self.get_option("batch_symmetric_min") if sym else None}

    def close(self):
        if self.handle is not None and self.handle.value:
            self._lib.grasph_destroy(self.handle)
        self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
