"""The reference's input formats: the .ini configuration and the fibers file.

Configuration keys are those runner/cuda_runner.py:18-47 reads
(tests/config.ini:18-37 of the reference); the fibers file ("layout v1") is
parsed as src/parameters.cc:75-116 and fortran/fibers.f95:156-168 read it:
the fiber count on the first line, then per fiber the position and the
orientation, 3 values each, in any line breaking.
"""
from __future__ import annotations

import configparser
import dataclasses

import numpy as np


@dataclasses.dataclass
class Parameters:
    NUMBER_OF_TIMESTEPS: int = 10
    STATE_SAVE_INTERVAL: int = 1
    VELOCITY_SAVE_INTERVAL: int = 1
    TIMESTEP: float = 0.1
    SLENDERNESS: float = 0.01
    NUMBER_OF_TERMS_IN_FORCE_EXPANSION: int = 5
    NUMBER_OF_QUADRATURE_INTERVALS: int = 8
    NUMBER_OF_QUADRATURE_POINTS_PER_INTERVAL: int = 3
    GMRES_RESTART: int = 10
    GMRES_MAX_ITERATIONS: int = 1000
    GMRES_TOLERANCE: float = 1e-5
    BICGSTAB_MAX_ITERATIONS: int = 1000
    BICGSTAB_TOLERANCE: float = 1e-5
    NUMBER_OF_FIBERS: int = 0


_KEYS = {
    "NUMBER_OF_TIMESTEPS": ("RUN", "NUMBER_OF_TIMESTEPS", int),
    "STATE_SAVE_INTERVAL": ("RUN", "STATE_SAVE_INTERVAL", int),
    "VELOCITY_SAVE_INTERVAL": ("RUN", "VELOCITY_SAVE_INTERVAL", int),
    "TIMESTEP": ("SIMULATION", "TIMESTEP", float),
    "SLENDERNESS": ("SIMULATION", "SLENDERNESS", float),
    "NUMBER_OF_TERMS_IN_FORCE_EXPANSION": ("SIMULATION", "NUMBER_OF_TERMS_IN_FORCE_EXPANSION", int),
    "NUMBER_OF_QUADRATURE_INTERVALS": ("SIMULATION", "NUMBER_OF_QUADRATURE_INTERVALS", int),
    "NUMBER_OF_QUADRATURE_POINTS_PER_INTERVAL": ("SIMULATION", "NUMBER_OF_QUADRATURE_POINTS_PER_INTERVAL", int),
    "GMRES_RESTART": ("GMRES", "RESTART", int),
    "GMRES_MAX_ITERATIONS": ("GMRES", "MAX_ITERATIONS", int),
    "GMRES_TOLERANCE": ("GMRES", "TOLERANCE", float),
    "BICGSTAB_MAX_ITERATIONS": ("BICGSTAB", "MAX_ITERATIONS", int),
    "BICGSTAB_TOLERANCE": ("BICGSTAB", "TOLERANCE", float),
}


class ConfigError(ValueError):
    pass


def read_configuration(path: str, validate: bool = False) -> Parameters:
    """Parses the .ini file; ``validate`` forces 2 timesteps as the reference's
    validate command does (runner/cuda_runner.py:26-28)."""
    cp = configparser.ConfigParser(inline_comment_prefixes=(";", "#"))
    with open(path, "r") as fh:
        cp.read_file(fh)
    p = Parameters()
    for attr, (section, key, typ) in _KEYS.items():
        try:
            raw = cp.get(section, key)
        except (configparser.NoSectionError, configparser.NoOptionError) as exc:
            raise ConfigError(f"{path}: missing [{section}] {key}") from exc
        try:
            setattr(p, attr, typ(raw))
        except ValueError as exc:
            raise ConfigError(f"{path}: [{section}] {key}={raw!r} is not a valid {typ.__name__}") from exc
    if validate:
        p.NUMBER_OF_TIMESTEPS = 2
    return p


def read_fibers(path: str):
    """Returns (M, positions[3M], orientations[3M]) in FP64."""
    with open(path, "r") as fh:
        first = fh.readline()
        try:
            M = int(first.split()[0])
        except (IndexError, ValueError) as exc:
            raise ConfigError(f"{path}: first line must hold the number of fibers") from exc
        values = np.array(fh.read().split(), dtype=np.float64)
    if M < 1:
        raise ConfigError(f"{path}: number of fibers must be >= 1")
    if values.size < 6 * M:
        raise ConfigError(f"{path}: expected {6 * M} values for {M} fibers, found {values.size}")
    v = values[: 6 * M].reshape(M, 2, 3)
    return M, np.ascontiguousarray(v[:, 0, :].ravel()), np.ascontiguousarray(v[:, 1, :].ravel())


def write_fibers(path: str, positions, orientations):
    """Writes a fibers file in the format src/simulation.cu:530-566 saves
    (M, then "x y z" and "tx ty tz" lines, fixed 8 decimals)."""
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    ori = np.asarray(orientations, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as fh:
        fh.write(f"{pos.shape[0]}\n")
        for p, o in zip(pos, ori):
            fh.write(f"{p[0]:.8f} {p[1]:.8f} {p[2]:.8f}\n{o[0]:.8f} {o[1]:.8f} {o[2]:.8f}\n")
