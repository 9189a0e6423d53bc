"""Shared helpers of the test suite: error metrics, the sampled oracle and
sub-views of batched particle arrays."""
import ctypes

import numpy as np


def vec_errors(got, ref, floor=1e-3):
    """Errors of a per-particle vector field (n, >=3) against a reference.

    Returns (max_rel, norm_rel): the largest per-particle |d a| / |a_ref| over
    particles whose |a_ref| is at least `floor` times the median |a_ref|
    (per-particle relative error is meaningless where forces cancel to ~0),
    and the norm-wise ||d a|| / ||a_ref|| over all particles.
    """
    g = np.asarray(got, dtype=np.float64)[:, :3]
    r = np.asarray(ref, dtype=np.float64)[:, :3]
    d = np.linalg.norm(g - r, axis=1)
    m = np.linalg.norm(r, axis=1)
    norm_rel = np.linalg.norm(d) / max(np.linalg.norm(m), 1e-300)
    keep = m >= floor * np.median(m) if len(m) else m > 0
    max_rel = float(np.max(d[keep] / m[keep])) if np.any(keep) else 0.0
    return max_rel, float(norm_rel)


def vec_errors_floored(got, ref, floor=0.1):
    """Largest per-particle error of a vector field relative to
    max(|a_ref|, floor * median |a_ref|), over every particle.

    Where a particle's pair terms cancel to a small fraction of their sum of
    magnitudes, its relative error measures the conditioning of the sum, not
    the arithmetic (random stresses leave particles with |a| at 0.5% of the
    median whose FP32 terms are each as exact as FP32 allows).  This metric
    holds those particles to an absolute error of `floor` x median x bound
    and every other particle to the relative bound.
    """
    g = np.asarray(got, dtype=np.float64)[:, :3]
    r = np.asarray(ref, dtype=np.float64)[:, :3]
    d = np.linalg.norm(g - r, axis=1)
    m = np.linalg.norm(r, axis=1)
    if not len(m):
        return 0.0
    return float(np.max(d / np.maximum(m, floor * np.median(m))))


def scalar_errors(got, ref, floor=1e-3):
    g = np.asarray(got, dtype=np.float64).ravel()
    r = np.asarray(ref, dtype=np.float64).ravel()
    d = np.abs(g - r)
    m = np.abs(r)
    norm_rel = np.linalg.norm(d) / max(np.linalg.norm(m), 1e-300)
    keep = m >= floor * np.median(m) if len(m) else m > 0
    max_rel = float(np.max(d[keep] / m[keep])) if np.any(keep) else 0.0
    return max_rel, float(norm_rel)


def tensor_errors(got, ref):
    """Symmetric-tensor planes (6, n): error per particle relative to the
    largest reference tensor entry, and norm-wise."""
    g = np.asarray(got, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    scale = max(np.abs(r).max(), 1e-300)
    max_rel = float(np.abs(g - r).max() / scale)
    norm_rel = float(np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-300))
    return max_rel, norm_rel


def sampled_oracle(prm, h, sample, prec):
    """Oracle acceleration (and density) of `sample`, computed through the
    whole derivative chain on the particles it depends on."""
    import grasph2_amd as g
    from scipy.spatial import cKDTree

    from oracle import pyoracle as O

    f = prm.flags
    if not (f & g.SPH):
        return O.calc_acceleration(prm, h, prec, sample=sample)
    pos = np.asarray(h["posm"][:, :3], dtype=np.float64)
    tree = cKDTree(pos)
    r = float(prm.H) * 1.0001

    def ring(idx):
        nb = tree.query_ball_point(pos[idx], r)
        return np.unique(np.concatenate([np.asarray(idx)] + [np.asarray(x, dtype=np.int64) for x in nb]))

    u1 = ring(sample)
    u2 = ring(u1)
    cur = dict(h)
    if not (f & g.INTEGRATE_DENSITY):
        cur["density"] = O.calc_density(prm, cur, prec, sample=u2)["density"]
    if (f & g.SOLIDS) or (f & g.INTEGRATE_DENSITY) or (f & g.BALSARA_SWITCH):
        o = O.calc_balsara(prm, cur, prec, sample=u1)
        cur["balsara"] = o["balsara"] if "balsara" in o else cur.get("balsara")
    out = O.calc_acceleration(prm, cur, prec, sample=sample)
    out["density"] = cur["density"]
    return out


# bytes per particle of each array of the C ABI (planar arrays: planes * 4 or 8)
_BYTES = {"posm": 16, "pos64": 24, "vel": 16, "acc": 16, "xvel": 16, "density": 4, "density_dt": 4,
          "balsara": 4, "maxvsig": 4, "dstress": 24, "dstress_dt": 24}


def sub_view(view, off, n):
    """The grasph_particles view of instance [off, off + n) of a batch laid
    out as include/grasph2.h describes (every array holds each instance's
    own single-instance layout back to back)."""
    import grasph2_amd as g

    v = g.Particles()
    v.n = n
    for f, _ in g.Particles._fields_[1:]:
        p = getattr(view, f)
        addr = p if isinstance(p, int) else (ctypes.cast(p, ctypes.c_void_p).value if p else None)
        setattr(v, f, addr + _BYTES[f] * off if addr else None)
    return v
