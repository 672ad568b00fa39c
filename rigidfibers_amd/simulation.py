"""Host-side mirror of the reference's Simulation class (src/simulation.h:38-98).

``Simulation(parameters, positions, orientations, solver=..., integration=...)``
owns one rf_simulation of librigidfibers.so; ``step(t)`` is
Simulation::step (src/simulation.cu:286-329).  The phase methods
(assemble_system, solve_system, update_velocities, update_fibers, advance)
and the getters expose what the reference dumps under VALIDATE.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .config import Parameters


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_config(params: Parameters, solver="gmres", integration="numerical", velocity_integration="follow",
                external_force=(0.0, 0.0, -1.0), refinement_tolerance=1e-10, refinement_max_iterations=50,
                initial_guess="previous", matrix_format="auto", velocity_operator="stored", inner_tolerance=0.0,
                separate_residual=False, confirm_residual=False):
    if solver not in _lib.SOLVERS:
        raise ValueError(f"unknown solver {solver!r}")
    if integration not in _lib.INTEGRATIONS:
        raise ValueError(f"unknown integration {integration!r}")
    cfg = _lib.RfConfig()
    _lib.check(_lib.lib().rf_config_init(C.byref(cfg)))
    cfg.number_of_fibers = params.NUMBER_OF_FIBERS
    cfg.number_of_terms_in_force_expansion = params.NUMBER_OF_TERMS_IN_FORCE_EXPANSION
    cfg.number_of_quadrature_intervals = params.NUMBER_OF_QUADRATURE_INTERVALS
    cfg.number_of_quadrature_points_per_interval = params.NUMBER_OF_QUADRATURE_POINTS_PER_INTERVAL
    cfg.timestep = params.TIMESTEP
    cfg.slenderness = params.SLENDERNESS
    for a in range(3):
        cfg.external_force[a] = float(external_force[a])
    cfg.solver = _lib.SOLVERS[solver]
    cfg.integration = _lib.INTEGRATIONS[integration]
    cfg.velocity_integration = _lib.VELOCITY_INTEGRATIONS[velocity_integration]
    cfg.gmres_restart = params.GMRES_RESTART
    cfg.gmres_max_iterations = params.GMRES_MAX_ITERATIONS
    cfg.gmres_tolerance = params.GMRES_TOLERANCE
    cfg.bicgstab_max_iterations = params.BICGSTAB_MAX_ITERATIONS
    cfg.bicgstab_tolerance = params.BICGSTAB_TOLERANCE
    cfg.refinement_tolerance = refinement_tolerance
    cfg.refinement_max_iterations = refinement_max_iterations
    cfg.initial_guess = _lib.INITIAL_GUESSES[initial_guess]
    cfg.matrix_format = _lib.MATRIX_FORMATS[matrix_format]
    cfg.velocity_operator = _lib.VELOCITY_OPERATORS[velocity_operator]
    cfg.inner_tolerance = inner_tolerance
    cfg.separate_residual = 1 if separate_residual else 0
    cfg.confirm_residual = 1 if confirm_residual else 0
    return cfg


class Simulation:
    def __init__(self, params: Parameters, positions, orientations, solver="gmres", integration="numerical",
                 velocity_integration="follow", external_force=(0.0, 0.0, -1.0), **kw):
        self.params = params
        self.M = int(params.NUMBER_OF_FIBERS)
        pos = np.ascontiguousarray(positions, dtype=np.float64).ravel()
        ori = np.ascontiguousarray(orientations, dtype=np.float64).ravel()
        if pos.size != 3 * self.M or ori.size != 3 * self.M:
            raise ValueError(f"expected {3 * self.M} position and orientation values")
        self._lib = _lib.lib()
        self.config = make_config(params, solver, integration, velocity_integration, external_force, **kw)
        handle = C.c_void_p()
        _lib.check(self._lib.rf_create(C.byref(self.config), _dptr(pos), _dptr(ori), C.byref(handle)))
        self._h = handle
        self.n = int(self._lib.rf_system_size(self._h))

    # -- lifetime --------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.rf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the timestep ------------------------------------------------------------
    def step(self, current_timestep: int) -> int:
        """Returns RF_OK or RF_WARN_NOT_CONVERGED; raises on errors."""
        return _lib.check(self._lib.rf_step(self._h, int(current_timestep)), allow_warning=True)

    def assemble_system(self):
        _lib.check(self._lib.rf_assemble_system(self._h))

    def solve_system(self):
        _lib.check(self._lib.rf_solve_system(self._h), allow_warning=True)

    def update_velocities(self):
        _lib.check(self._lib.rf_update_velocities(self._h))

    def update_fibers(self, first_timestep: bool):
        _lib.check(self._lib.rf_update_fibers(self._h, 1 if first_timestep else 0))

    def advance(self):
        _lib.check(self._lib.rf_advance(self._h))

    # -- state ---------------------------------------------------------------------
    def state(self):
        p, o = np.empty(3 * self.M), np.empty(3 * self.M)
        _lib.check(self._lib.rf_get_state(self._h, _dptr(p), _dptr(o)))
        return p, o

    def next_state(self):
        p, o = np.empty(3 * self.M), np.empty(3 * self.M)
        _lib.check(self._lib.rf_get_next_state(self._h, _dptr(p), _dptr(o)))
        return p, o

    def velocities(self):
        t, r = np.empty(3 * self.M), np.empty(3 * self.M)
        _lib.check(self._lib.rf_get_velocities(self._h, _dptr(t), _dptr(r)))
        return t, r

    def set_state(self, positions, orientations, previous=None):
        """previous = (positions, orientations, translational, rotational) of the step before."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).ravel()
        ori = np.ascontiguousarray(orientations, dtype=np.float64).ravel()
        if previous is None:
            _lib.check(self._lib.rf_set_state(self._h, _dptr(pos), _dptr(ori), 0, None, None, None, None))
        else:
            prev = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in previous]
            _lib.check(self._lib.rf_set_state(self._h, _dptr(pos), _dptr(ori), 1, *(_dptr(a) for a in prev)))

    def system(self, with_matrix=False):
        """(A or None, b, x): A column-major n x n FP32 as a Fortran-ordered array."""
        b, x = np.empty(self.n), np.empty(self.n)
        A = np.empty((self.n, self.n), dtype=np.float32, order="F") if with_matrix else None
        aptr = A.ctypes.data_as(C.POINTER(C.c_float)) if with_matrix else None
        _lib.check(self._lib.rf_get_system(self._h, aptr, _dptr(b), _dptr(x)))
        return A, b, x

    def solver_stats(self):
        st = _lib.RfSolverStats()
        _lib.check(self._lib.rf_get_solver_stats(self._h, C.byref(st)))
        return {"iterations": st.iterations, "converged": bool(st.converged),
                "relative_residual": st.relative_residual}

    def timings(self):
        t = _lib.RfTimings()
        _lib.check(self._lib.rf_get_timings(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _lib.RfTimings._fields_ if not k.startswith("reserved")}

    def device_bytes(self):
        return int(self._lib.rf_device_bytes(self._h))
