"""CPU-side checks of the C ABI and the host logic (no device calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "rigidfibers.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"RF_API\s+[\w\s\*]+?\b(rf_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from rigidfibers_amd import _lib

    lib = _lib.load()
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTS)


def test_config_struct_layout_and_defaults():
    """rf_config_init runs host code only; its defaults are the values of the
    reference's tests/config.ini."""
    from rigidfibers_amd import _lib

    lib = _lib.load()
    cfg = _lib.RfConfig()
    assert lib.rf_config_init(C.byref(cfg)) == 0
    assert (cfg.number_of_terms_in_force_expansion, cfg.number_of_quadrature_intervals,
            cfg.number_of_quadrature_points_per_interval) == (5, 8, 3)
    assert (cfg.timestep, cfg.slenderness) == (0.1, 0.01)
    assert list(cfg.external_force) == [0.0, 0.0, -1.0]
    assert (cfg.gmres_restart, cfg.gmres_max_iterations, cfg.gmres_tolerance) == (10, 1000, 1e-5)
    assert (cfg.bicgstab_max_iterations, cfg.bicgstab_tolerance) == (1000, 1e-5)
    assert cfg.initial_guess == 1
    assert C.sizeof(_lib.RfConfig) == 152  # sizeof(rf_config) in C (gcc, x86-64)
    assert lib.rf_version().decode().startswith("rigidfibers_amd")
    assert lib.rf_config_init(None) == _lib.RF_ERR_INVALID_ARGUMENT
    assert b"NULL" in lib.rf_last_error()


def test_read_configuration_and_validate_override():
    from rigidfibers_amd.config import ConfigError, read_configuration

    p = read_configuration(os.path.join(GOLDEN, "config.ini"))
    assert p.NUMBER_OF_TIMESTEPS == 10 and p.GMRES_RESTART == 10 and p.SLENDERNESS == 0.01
    assert read_configuration(os.path.join(GOLDEN, "config.ini"), validate=True).NUMBER_OF_TIMESTEPS == 2


def test_read_configuration_errors(tmp_path):
    from rigidfibers_amd.config import ConfigError, read_configuration

    bad = tmp_path / "bad.ini"
    bad.write_text("[RUN]\nNUMBER_OF_TIMESTEPS=3\n")
    with pytest.raises(ConfigError, match="STATE_SAVE_INTERVAL"):
        read_configuration(str(bad))
    text = open(os.path.join(GOLDEN, "config.ini")).read().replace("TIMESTEP=0.1", "TIMESTEP=abc")
    bad.write_text(text)
    with pytest.raises(ConfigError, match="TIMESTEP"):
        read_configuration(str(bad))


def test_fibers_file_roundtrip(tmp_path):
    from rigidfibers_amd.config import read_fibers, write_fibers

    M, pos, ori = read_fibers(os.path.join(GOLDEN, "XcT_ref100.in"))
    assert M == 100 and pos.shape == (300,)
    assert np.isclose(pos[0], 0.4359164870203274) and np.isclose(ori[2], -0.0699232812799387)
    out = tmp_path / "XcT_res_00001.out"
    write_fibers(str(out), pos, ori)
    M2, pos2, ori2 = read_fibers(str(out))
    assert M2 == M and np.abs(pos2 - pos).max() < 5e-9 and np.abs(ori2 - ori).max() < 5e-9
    # any line breaking is accepted (fortran/fibers.f95:162-168 list-directed read)
    one_line = tmp_path / "one.in"
    one_line.write_text("2\n" + " ".join(str(v) for v in np.arange(12.0)) + "\n")
    M3, p3, o3 = read_fibers(str(one_line))
    assert M3 == 2 and list(p3) == [0, 1, 2, 6, 7, 8] and list(o3) == [3, 4, 5, 9, 10, 11]


def test_fibers_file_errors(tmp_path):
    from rigidfibers_amd.config import ConfigError, read_fibers

    f = tmp_path / "short.in"
    f.write_text("3\n0 0 0\n0 0 1\n")
    with pytest.raises(ConfigError, match="expected 18 values"):
        read_fibers(str(f))
    with pytest.raises(FileNotFoundError):
        read_fibers(str(tmp_path / "missing.in"))
