"""rigidfibers_amd -- the Tornberg-Gustavsson rigid-fiber Stokes-flow timestep
on MI355X (gfx950).

The compute path is librigidfibers.so (hand-written HIP kernels behind the C
ABI of include/rigidfibers.h); this package is the host side: the reference's
.ini and fibers-file formats, a Simulation class mirroring
src/simulation.h:38-98, and the fibers_runner.py command line.
"""
from .config import Parameters, read_configuration, read_fibers, write_fibers  # noqa: F401

__version__ = "0.1.0"


def library():
    """Loads librigidfibers.so; raises ImportError if it has not been built."""
    from . import _lib

    return _lib.lib()


def Simulation(*args, **kwargs):  # noqa: N802 - mirrors the reference class name
    from .simulation import Simulation as _Simulation

    return _Simulation(*args, **kwargs)
