"""Shared fixtures data and comparisons for the tests."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
VARIANTS = ("numeric_direct", "numeric_gmres", "analytic_direct", "analytic_gmres")
FIELDS = ("BVec", "XVec", "TRANSVel", "ROTVel", "POS", "ORIENT")


def golden(variant, step, field):
    """tests/reference/100_<variant>/<step>_<field>.ref of the reference (F16.8 values)."""
    return np.loadtxt(os.path.join(GOLDEN, "reference", f"100_{variant}", f"{step}_{field}.ref"))


def ref100():
    from rigidfibers_amd.config import read_fibers

    return read_fibers(os.path.join(GOLDEN, "XcT_ref100.in"))


def ref_config(validate=True):
    from rigidfibers_amd.config import read_configuration

    return read_configuration(os.path.join(GOLDEN, "config.ini"), validate=validate)


def random_scene(M, side, seed, min_dist=0.4):
    """Centres uniform in a cube (rejection for a minimum centre distance),
    orientations uniform on the sphere."""
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < M:
        c = rng.uniform(-side / 2, side / 2, size=3)
        if all(np.linalg.norm(c - q) >= min_dist for q in pts):
            pts.append(c)
    ori = rng.normal(size=(M, 3))
    ori /= np.linalg.norm(ori, axis=1, keepdims=True)
    return np.array(pts).ravel(), ori.ravel()


def rel_err(a, ref):
    """max |a - ref| / max |ref| (norm-wise relative error of an array)."""
    a, ref = np.asarray(a, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    scale = np.abs(ref).max()
    return float(np.abs(a - ref).max() / (scale if scale > 0 else 1.0))
