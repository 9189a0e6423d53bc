"""Shared helpers of the test suite: error metrics and GPU round trips."""
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
