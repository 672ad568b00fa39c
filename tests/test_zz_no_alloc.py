"""rf_get_system_noalloc (the entry point that allocates no device memory)
against rf_get_system, through the C ABI: the same dense matrix, b and x,
bit for bit, on two calls in a row, the second with a smaller system."""
import ctypes as C

import numpy as np
import pytest

from helpers import random_scene, ref_config

pytestmark = pytest.mark.gpu


def _both(sim):
    """(A, b, x) from rf_get_system and from rf_get_system_noalloc."""
    from rigidfibers_amd import _lib

    out = []
    for name in ("rf_get_system", "rf_get_system_noalloc"):
        A = np.full((sim.n, sim.n), np.nan, dtype=np.float32, order="F")
        b, x = np.full(sim.n, np.nan), np.full(sim.n, np.nan)
        fn = getattr(sim._lib, name)
        _lib.check(fn(sim._h, A.ctypes.data_as(C.POINTER(C.c_float)), b.ctypes.data_as(C.POINTER(C.c_double)),
                      x.ctypes.data_as(C.POINTER(C.c_double))))
        out.append((A, b, x))
    return out


@pytest.mark.parametrize("matrix_format", ["theta", "dense"])
def test_noalloc_system_matches_rf_get_system(matrix_format):
    """M = 150 (the theta form is expanded in two ranges of source fibers
    through the 16 MiB staging buffer), then M = 40 (one range), each after a
    full step, so the matrix belongs to the orientations it was assembled
    from, not to the advanced state."""
    from rigidfibers_amd.simulation import Simulation

    for M, seed in ((150, 41), (40, 42)):
        pos, ori = random_scene(M, side=5.0 * (M / 60) ** (1 / 3), seed=seed, min_dist=0.4)
        params = ref_config(validate=False)
        params.NUMBER_OF_FIBERS = M
        with Simulation(params, pos, ori, solver="gmres", integration="analytical",
                        matrix_format=matrix_format) as sim:
            sim.step(0)
            (A0, b0, x0), (A1, b1, x1) = _both(sim)
            print(f"{matrix_format} M={M}: n={sim.n}, max|A|={np.abs(A0).max():.3e}")
            assert np.isfinite(A0).all() and np.isfinite(b0).all() and np.isfinite(x0).all()
            assert np.array_equal(A0, A1) and np.array_equal(b0, b1) and np.array_equal(x0, x1)
            # the identity diagonal blocks, and A x = b to the solver's tolerance
            assert np.array_equal(A0[:15, :15], np.eye(15, dtype=np.float32))
            r = b0 - A0.astype(np.float64) @ x0
            assert np.linalg.norm(r) <= 1e-4 * np.linalg.norm(b0)
