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
    assert (cfg.separate_residual, cfg.confirm_residual) == (0, 0)
    assert C.sizeof(_lib.RfConfig) == 152  # sizeof(rf_config) in C (gcc, x86-64)
    assert lib.rf_version().decode().startswith("rigidfibers_amd")
    assert lib.rf_config_init(None) == _lib.RF_ERR_INVALID_ARGUMENT
    assert b"NULL" in lib.rf_last_error()


def test_struct_layouts_match_the_header(tmp_path):
    """Every struct of include/rigidfibers.h has the size and field offsets the
    ctypes binding assumes (a C program compiled against the header prints them)."""
    import subprocess

    from rigidfibers_amd import _lib

    structs = {"rf_config": _lib.RfConfig, "rf_solver_stats": _lib.RfSolverStats, "rf_timings": _lib.RfTimings,
               "rf_server_config": _lib.RfServerConfig, "rf_request_info": _lib.RfRequestInfo,
               "rf_server_stats": _lib.RfServerStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rigidfibers.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src, exe = tmp_path / "layout.c", tmp_path / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)


def test_request_stream_is_seeded():
    from rigidfibers_amd.serve import request_stream, summarize

    a = request_stream(count=6, m_range=(16, 40), rate=50.0, steps=3, seed=4)
    b = request_stream(count=6, m_range=(16, 40), rate=50.0, steps=3, seed=4)
    assert [r[1] for r in a] == [r[1] for r in b] and all(np.array_equal(x[2], y[2]) for x, y in zip(a, b))
    assert all(16 <= r[1] <= 40 for r in a) and all(t1 > t0 for (t0, *_), (t1, *_) in zip(a, a[1:]))
    for arrival, M, pos, ori, steps in a:
        side = (M / 0.1) ** (1 / 3)
        assert np.abs(pos).max() <= side / 2 and np.allclose(np.linalg.norm(ori.reshape(-1, 3), axis=1), 1.0)
    res = [{"fibers": 10, "steps": 2, "step_ms": [1.0, 3.0], "arrival": 0.0, "admitted": 0.5, "finished": 1.0}]
    s = summarize(res, 2.0)
    assert s["fiber_steps"] == 20 and s["fiber_steps_per_s"] == 10.0 and s["step_ms_p50"] == 2.0


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


def test_confirm_residual_switch_reaches_the_config():
    """--confirm-residual of both runner surfaces sets rf_config.confirm_residual,
    and rf_create rejects a value other than 0 or 1 before it touches a device."""
    from rigidfibers_amd import _lib
    from rigidfibers_amd.runner import build_parser
    from rigidfibers_amd.simulation import make_config

    from helpers import ref_config

    parser = build_parser()
    for platform, solver in (("cuda", "--gmres"), ("fortran", "--gmres")):
        args = parser.parse_args([platform, "run", "config.ini", "fibers.in", solver, "--analytical",
                                  "--confirm-residual"] + (["--D1"] if platform == "cuda" else []))
        assert args.confirm_residual is True
        args = parser.parse_args([platform, "run", "config.ini", "fibers.in", solver, "--analytical"]
                                 + (["--D1"] if platform == "cuda" else []))
        assert args.confirm_residual is False
    params = ref_config()
    params.NUMBER_OF_FIBERS = 4  # the fibers file sets it in a run
    assert make_config(params).confirm_residual == 0
    cfg = make_config(params, confirm_residual=True)
    assert cfg.confirm_residual == 1
    cfg.confirm_residual = 2
    M = params.NUMBER_OF_FIBERS
    pos = np.zeros(3 * M)
    ori = np.tile([0.0, 0.0, 1.0], M)
    sim = C.c_void_p()
    lib = _lib.load()
    assert lib.rf_create(C.byref(cfg), pos.ctypes.data_as(C.POINTER(C.c_double)),
                         ori.ctypes.data_as(C.POINTER(C.c_double)), C.byref(sim)) == _lib.RF_ERR_INVALID_ARGUMENT
    assert b"confirm_residual" in lib.rf_last_error()
    assert not sim.value
