"""ctypes view of the C ABI in include/grasph2.h.

Loads the in-tree grasph2_amd/libgrasph2_amd.so; raises if it is missing, so
a box without the built HIP library fails loudly instead of running anything
else in its place.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgrasph2_amd.so")

# physics switches (settings/settings.h), same bits as include/grasph2.h
SELF_GRAVITY = 1 << 0
SPH = 1 << 1
INTEGRATE_DENSITY = 1 << 2
DEAL_WITH_NO_PARTNERS = 1 << 3
SOLIDS = 1 << 4
PLASTICITY_MIESE = 1 << 5
PLASTICITY_MC = 1 << 6
CLOHESSY_WILTSHIRE = 1 << 7
CONSTANT_ACCELERATION = 1 << 8
SIMPLE_BOX_BOUNDARY = 1 << 9
ARTIFICIAL_VISCOSITY = 1 << 10
BALSARA_SWITCH = 1 << 11
ARTIFICIAL_STRESS = 1 << 12
XSPH = 1 << 13
VARIABLE_TIMESTEP = 1 << 14
ACCELERATION_CRITERION = 1 << 15
VELOCITY_CRITERION = 1 << 16

E_INVALID, E_HIP, E_UNSUPPORTED, E_NOMEM = -1, -2, -3, -4

INV_EKIN, INV_EPOT, INV_EINT, INV_PX, INV_PY, INV_PZ, INV_LX, INV_LY, INV_LZ, INV_MASS = range(10)
NINV = 10
NKERNEL_CLASSES = 8
KERNEL_CLASSES = ("sweep", "density", "strain", "accel", "integrate", "reorder", "other", "far")


class Params(ctypes.Structure):
    _fields_ = [
        ("flags", ctypes.c_uint32),
        ("dimension", ctypes.c_int32),
        ("H", ctypes.c_float),
        ("rho0", ctypes.c_float),
        ("BULK", ctypes.c_float),
        ("dBULKdP", ctypes.c_float),
        ("SOUNDSPEED", ctypes.c_float),
        ("shear", ctypes.c_float),
        ("Y", ctypes.c_float),
        ("tanfr", ctypes.c_float),
        ("cohesion", ctypes.c_float),
        ("alpha", ctypes.c_float),
        ("mateps", ctypes.c_float),
        ("matexp", ctypes.c_float),
        ("normalsep", ctypes.c_float),
        ("xsph_factor", ctypes.c_float),
        ("cw_n", ctypes.c_float),
        ("constant_acceleration", ctypes.c_float * 3),
        ("box_min", ctypes.c_float * 3),
        ("box_max", ctypes.c_float * 3),
        ("box_reflectiveness", ctypes.c_float),
        ("fixed_timestep", ctypes.c_float),
        ("initial_timestep", ctypes.c_float),
        ("min_timestep", ctypes.c_float),
        ("max_timestep", ctypes.c_float),
        ("accel_accuracy", ctypes.c_float),
        ("velocity_accuracy", ctypes.c_float),
    ]

    def copy(self):
        c = Params()
        ctypes.pointer(c)[0] = self
        return c

    def as_dict(self):
        out = {}
        for f, _ in self._fields_:
            v = getattr(self, f)
            out[f] = list(v) if hasattr(v, "__len__") else v
        return out


class Particles(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("posm", ctypes.c_void_p),
        ("pos64", ctypes.c_void_p),
        ("vel", ctypes.c_void_p),
        ("acc", ctypes.c_void_p),
        ("xvel", ctypes.c_void_p),
        ("density", ctypes.c_void_p),
        ("density_dt", ctypes.c_void_p),
        ("balsara", ctypes.c_void_p),
        ("maxvsig", ctypes.c_void_p),
        ("dstress", ctypes.c_void_p),
        ("dstress_dt", ctypes.c_void_p),
    ]


class Batch(ctypes.Structure):
    _fields_ = [
        ("n_instances", ctypes.c_int32),
        ("offsets", ctypes.POINTER(ctypes.c_int64)),
        ("params", ctypes.POINTER(Params)),
    ]


# every symbol include/grasph2.h declares, with (restype, argtypes)
_P = ctypes.POINTER
_SIGNATURES = {
    "grasph_version": (ctypes.c_int, []),
    "grasph_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "grasph_params_default": (None, [_P(Params)]),
    "grasph_create": (ctypes.c_int, [ctypes.c_int, _P(ctypes.c_void_p)]),
    "grasph_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "grasph_set_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64]),
    "grasph_compute_derivatives": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Params), ctypes.c_void_p]),
    "grasph_compute_derivatives_noalloc": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Params), ctypes.c_void_p]),
    "grasph_calc_density": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Params), ctypes.c_void_p]),
    "grasph_calc_balsara_density_dt_dstress_dt": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Params), ctypes.c_void_p]),
    "grasph_calc_acceleration": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Params), ctypes.c_void_p]),
    "grasph_integrate": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Params), ctypes.c_int, ctypes.c_void_p]),
    "grasph_get_current_timestep": (ctypes.c_double, [ctypes.c_void_p, _P(Params)]),
    "grasph_reduce_invariants": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Params), ctypes.c_double,
                                                ctypes.c_int, _P(ctypes.c_double), ctypes.c_void_p]),
    "grasph_get_kernel_times": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_double), _P(ctypes.c_int64), ctypes.c_int]),
    "grasph_get_neighbour_stats": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int64), _P(ctypes.c_int32),
                                                  _P(ctypes.c_int64)]),
    "grasph_batch_compute_derivatives": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Batch), ctypes.c_void_p]),
    "grasph_batch_integrate": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Batch), ctypes.c_int,
                                              ctypes.c_void_p]),
    "grasph_batch_step": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Batch), ctypes.c_int, ctypes.c_void_p]),
    "grasph_batch_step_loop": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Batch), ctypes.c_int,
                                              ctypes.c_void_p]),
    "grasph_autotune": (ctypes.c_int, [ctypes.c_void_p, _P(Particles), _P(Batch), ctypes.c_char_p, ctypes.c_void_p]),
    "grasph_get_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _P(ctypes.c_int64)]),
    "grasph_get_batch_tuning": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int32), _P(ctypes.c_int32),
                                               _P(ctypes.c_int32)]),
    "grasph_batch_get_current_timesteps": (ctypes.c_int, [ctypes.c_void_p, _P(Batch), _P(ctypes.c_double)]),
}

_LIB = None


def lib():
    """The loaded HIP library; raises RuntimeError when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("HIP library %s is missing: run __graft_entry__.build() (or python "
                               "grasph2_amd/build.py) first" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class GrasphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("grasph error %d: %s" % (code, msg))
        self.code = code


def default_params():
    p = Params()
    lib().grasph_params_default(ctypes.byref(p))
    return p
