"""ctypes binding of the CPU oracle oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Arrays use the C-ABI layout of include/grasph2.h (dict of numpy arrays as
made by grasph2_amd.host_particles); the oracle runs in float32 ("f32", the
reference's default precision) or float64 ("f64", the FP64 oracle).  Inputs
are converted to the requested precision, copied, and the updated copies are
returned; the caller's arrays are not modified.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_FIELDS = ("posm", "vel", "acc", "xvel", "density", "density_dt", "balsara", "maxvsig", "dstress", "dstress_dt")


class OracleParticles(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64)] + [(f, ctypes.c_void_p) for f in _FIELDS]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library %s is missing: run make -C oracle" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(OracleParticles)
        for prec in ("f32", "f64"):
            for name in ("calc_density", "calc_balsara"):
                fn = getattr(L, "oracle_%s_%s" % (name, prec))
                fn.restype = ctypes.c_int
                fn.argtypes = [ctypes.c_void_p, P, ctypes.c_void_p, ctypes.c_int64]
            fn = getattr(L, "oracle_calc_acceleration_" + prec)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_void_p, P, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
            fn = getattr(L, "oracle_compute_derivatives_" + prec)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_void_p, P, ctypes.c_void_p]
            fn = getattr(L, "oracle_fixed_ts_leapfrog_" + prec)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_void_p, P, ctypes.c_int]
            fn = getattr(L, "oracle_variable_ts_leapfrog_" + prec)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_void_p, P, ctypes.c_int, ctypes.c_void_p]
            fn = getattr(L, "oracle_potential_" + prec)
            fn.restype = None
            fn.argtypes = [ctypes.c_void_p, P, ctypes.c_void_p]
            fn = getattr(L, "oracle_invariants_" + prec)
            fn.restype = None
            fn.argtypes = [ctypes.c_void_p, P, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        for name in ("wspline", "dwspline"):
            fn = getattr(L, "oracle_%s_f64" % name)
            fn.restype = ctypes.c_double
            fn.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
            fn = getattr(L, "oracle_%s_f32" % name)
            fn.restype = ctypes.c_float
            fn.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int]
        L.oracle_eos_liquid_f64.restype = ctypes.c_double
        L.oracle_eos_liquid_f64.argtypes = [ctypes.c_double] * 3
        L.oracle_eos_murnaghan_f64.restype = ctypes.c_double
        L.oracle_eos_murnaghan_f64.argtypes = [ctypes.c_double] * 4
        D = ctypes.POINTER(ctypes.c_double)
        L.oracle_artificial_viscosity_f64.restype = ctypes.c_double
        L.oracle_artificial_viscosity_f64.argtypes = [D, ctypes.c_double, ctypes.c_double, ctypes.c_double, D, D,
                                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_double]
        L.oracle_balsara_switch_f64.restype = ctypes.c_double
        L.oracle_balsara_switch_f64.argtypes = [ctypes.c_double, D, ctypes.c_double, ctypes.c_double]
        L.oracle_plasticity_f64.restype = None
        L.oracle_plasticity_f64.argtypes = [D, ctypes.c_double]
        L.oracle_dstress_dt_f64.restype = None
        L.oracle_dstress_dt_f64.argtypes = [D, D, D, ctypes.c_double, D]
        L.oracle_strain_rotation_f64.restype = None
        L.oracle_strain_rotation_f64.argtypes = [D, D, ctypes.c_double, ctypes.c_double, D, D]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_get_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def set_threads(n):
    lib().oracle_set_threads(int(n))


def threads():
    return lib().oracle_get_threads()


class _Bound:
    """Precision-converted copies of the arrays plus the ctypes view of them."""

    def __init__(self, h, prec):
        dt = np.float32 if prec == "f32" else np.float64
        self.arrays = {}
        n = h["posm"].shape[0]
        for f in _FIELDS:
            if f in h and h[f] is not None:
                self.arrays[f] = np.ascontiguousarray(np.array(h[f], dtype=dt, copy=True))
        if prec == "f64" and "pos64" in h and h["pos64"] is not None:
            # the FP64 masters are the positions of the FP64 oracle
            self.arrays["posm"][:, :3] = np.asarray(h["pos64"], dtype=np.float64).T
        self.view = OracleParticles()
        self.view.n = n
        for f in _FIELDS:
            setattr(self.view, f, self.arrays[f].ctypes.data if f in self.arrays else None)


def _check(r):
    if r != 0:
        raise RuntimeError("oracle returned %d" % r)


def _sample(sample):
    if sample is None:
        return None, 0, None
    s = np.ascontiguousarray(np.asarray(sample, dtype=np.int64))
    return s.ctypes.data, len(s), s


def calc_density(prm, h, prec="f64", sample=None):
    b = _Bound(h, prec)
    sp, ns, keep = _sample(sample)
    _check(getattr(lib(), "oracle_calc_density_" + prec)(ctypes.byref(prm), ctypes.byref(b.view), sp, ns))
    return b.arrays


def calc_balsara(prm, h, prec="f64", sample=None):
    b = _Bound(h, prec)
    sp, ns, keep = _sample(sample)
    _check(getattr(lib(), "oracle_calc_balsara_" + prec)(ctypes.byref(prm), ctypes.byref(b.view), sp, ns))
    return b.arrays


def _ts(state, prec):
    if state is None:
        return None
    if state.prec != prec:
        raise ValueError("time-step state precision %s, oracle precision %s" % (state.prec, prec))
    return state.ts.ctypes.data


def calc_acceleration(prm, h, prec="f64", sample=None, state=None):
    """calcAcceleration; with the variable step, `state` (a TimestepState of
    the same precision) gets next dt = max_timestep as do_after does."""
    b = _Bound(h, prec)
    sp, ns, keep = _sample(sample)
    _check(getattr(lib(), "oracle_calc_acceleration_" + prec)(ctypes.byref(prm), ctypes.byref(b.view), sp, ns,
                                                              _ts(state, prec)))
    return b.arrays


def compute_derivatives(prm, h, prec="f64", state=None):
    b = _Bound(h, prec)
    _check(getattr(lib(), "oracle_compute_derivatives_" + prec)(ctypes.byref(prm), ctypes.byref(b.view),
                                                                _ts(state, prec)))
    return b.arrays


def leapfrog(prm, h, not_first_step=True, prec="f64"):
    b = _Bound(h, prec)
    _check(getattr(lib(), "oracle_fixed_ts_leapfrog_" + prec)(ctypes.byref(prm), ctypes.byref(b.view),
                                                              1 if not_first_step else 0))
    return b.arrays


class TimestepState:
    """The reference's __device__ currentTS / nextTS (src/integration.h:127-128),
    zero until the first integrate call, in the oracle's precision."""

    def __init__(self, prec="f64"):
        self.prec = prec
        self.ts = np.zeros(2, dtype=np.float32 if prec == "f32" else np.float64)

    @property
    def current(self):
        return float(self.ts[0])

    @property
    def next(self):
        return float(self.ts[1])


def variable_leapfrog(prm, h, state, not_first_step=True):
    """integrate(pb, notFirstStep) with VARIABLE_TIMESTEP_LEAPFROG; updates
    `state` (a TimestepState) in place and returns the new arrays."""
    b = _Bound(h, state.prec)
    _check(getattr(lib(), "oracle_variable_ts_leapfrog_" + state.prec)(
        ctypes.byref(prm), ctypes.byref(b.view), 1 if not_first_step else 0, state.ts.ctypes.data))
    return b.arrays


def potential(prm, h, prec="f64"):
    b = _Bound(h, prec)
    phi = np.zeros(h["posm"].shape[0], dtype=np.float64)
    getattr(lib(), "oracle_potential_" + prec)(ctypes.byref(prm), ctypes.byref(b.view), phi.ctypes.data)
    return phi


def invariants(prm, h, half_kick=0.0, phi=None, prec="f64"):
    b = _Bound(h, prec)
    out = np.zeros(10, dtype=np.float64)
    pos64 = None
    if "pos64" in h and h["pos64"] is not None:
        p64 = np.ascontiguousarray(h["pos64"], dtype=np.float64)
        pos64 = p64.ctypes.data
    ph = np.ascontiguousarray(phi, dtype=np.float64) if phi is not None else None
    getattr(lib(), "oracle_invariants_" + prec)(ctypes.byref(prm), ctypes.byref(b.view), pos64, float(half_kick),
                                                ph.ctypes.data if ph is not None else None, out.ctypes.data)
    return out


# scalar physics (FP64) for the closed-form pins
def W(r, h, dim=3):
    return lib().oracle_wspline_f64(r, h, dim)


def dW(r, h, dim=3):
    return lib().oracle_dwspline_f64(r, h, dim)


def W32(r, h, dim=3):
    return lib().oracle_wspline_f32(r, h, dim)


def eos_liquid(rho, rho0, c2):
    return lib().oracle_eos_liquid_f64(rho, rho0, c2)


def eos_murnaghan(rho, rho0, K0, dK0):
    return lib().oracle_eos_murnaghan_f64(rho, rho0, K0, dK0)


def _dp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def artificial_viscosity(alpha, di, dj, vij, rij, r, ci, cj, bi=1.0, bj=1.0, maxvsig=None):
    v, vp = _dp(vij)
    x, xp = _dp(rij)
    if maxvsig is None:
        return lib().oracle_artificial_viscosity_f64(None, alpha, di, dj, vp, xp, r, ci, cj, bi, bj), None
    mv = ctypes.c_double(maxvsig)
    II = lib().oracle_artificial_viscosity_f64(ctypes.pointer(mv), alpha, di, dj, vp, xp, r, ci, cj, bi, bj)
    return II, mv.value


def balsara_switch(divv, curl, c, h):
    cc, cp = _dp(curl)
    return lib().oracle_balsara_switch_f64(divv, cp, c, h)


def plasticity(S, Y):
    s, sp = _dp(np.array(S, dtype=np.float64).reshape(9))
    lib().oracle_plasticity_f64(sp, Y)
    return s.reshape(3, 3)


def dstress_dt(edot, rdot, S, shear):
    e, ep = _dp(np.array(edot).reshape(9))
    r, rp = _dp(np.array(rdot).reshape(9))
    s, sp = _dp(np.array(S).reshape(9))
    o, op = _dp(np.zeros(9))
    lib().oracle_dstress_dt_f64(ep, rp, sp, shear, op)
    return o.reshape(3, 3)


def strain_rotation(edot, rdot, mj, dj, vij, gradw):
    e, ep = _dp(np.array(edot, dtype=np.float64).reshape(9))
    r, rp = _dp(np.array(rdot, dtype=np.float64).reshape(9))
    v, vp = _dp(vij)
    g, gp = _dp(gradw)
    lib().oracle_strain_rotation_f64(ep, rp, mj, dj, vp, gp)
    return e.reshape(3, 3), r.reshape(3, 3)
